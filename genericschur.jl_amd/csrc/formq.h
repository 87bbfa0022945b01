// Formation of the orthogonal / unitary factor Q from the Householder
// reflectors of a Hessenberg (or Hermitian tridiagonal, lower triangle)
// reduction, in place over the reflector storage: reflector p in column p
// below the subdiagonal, tau[p] its scalar, Q = diag(1, H_0 H_1 ... H_{n-2})
// (hessenberg.jl:150-166: the unblocked backward accumulation of xUNG2R on
// the reflectors shifted one column right).  One workgroup of NT threads.
#pragma once
#include "dev_common.h"

namespace gs {

// Complex.  Per reflector: the dot products (finished by the lanes holding
// them), then the rank-1 update of columns c+1.. together with column c
// itself (disjoint columns).
template <int NT, int RM = 0> GS_DEV void z_form_q_inplace(zd* H, int64_t ld, int n, const zd* tau, zd* wk) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  for (int k = n - 2; k >= 0; --k) {
    const int c = k + 1;  // column of Q produced by reflector k
    const zd t = tau[k];
    // apply H_k to Q(c:n-1, c+1:n-1); Q(c, j>c) is zero at this point:
    // s_j = t (Q[c, j] + sum_{r>c} conj(v_r) Q[r, j]); column j -= s_j [1; v]
    zdots4e<NT>(c + 1, n, c + 1, n,
                [&](int r, int j) { return zconj(H[r + (int64_t)k * ld]) * H[r + (int64_t)j * ld]; },
                [&](int j, zd acc) { wk[j] = t * (H[c + (int64_t)j * ld] + acc); });
    __syncthreads();
    if constexpr (RM > 0) {
      zd vr[RM];
#pragma unroll
      for (int q = 0; q < RM; ++q) {
        const int r = c + lane + 64 * q;
        vr[q] = (r < n) ? ((r == c) ? zr(1.0) : H[r + (int64_t)k * ld]) : zr(0.0);
      }
#pragma unroll 4
      for (int j = c + 1 + wave; j < n; j += NT / 64) {
        const zd sj = wk[j];
#pragma unroll
        for (int q = 0; q < RM; ++q) {
          const int r = c + lane + 64 * q;
          if (r < n) H[r + (int64_t)j * ld] = H[r + (int64_t)j * ld] - sj * vr[q];
        }
      }
    } else {
#pragma unroll 4
      for (int j = c + 1 + wave; j < n; j += NT / 64) {
        const zd sj = wk[j];
        for (int r = c + lane; r < n; r += 64) {
          const zd h = H[r + (int64_t)j * ld];
          H[r + (int64_t)j * ld] = (r == c) ? h - sj : h - sj * H[r + (int64_t)k * ld];
        }
      }
    }
    // column c: -tau * v below, 1 - tau on the diagonal, zero above (reads
    // column k, which the rank-1 update only reads)
    for (int r = tid; r < n; r += NT) {
      zd q;
      if (r < c) q = zr(0.0);
      else if (r == c) q = zr(1.0) - t;
      else q = -(t * H[r + (int64_t)k * ld]);
      H[r + (int64_t)c * ld] = q;
    }
    __syncthreads();
  }
  for (int r = tid; r < n; r += NT) H[r] = (r == 0) ? zr(1.0) : zr(0.0);
  __syncthreads();
}

// Real: the same with real reflectors.
template <int NT, int RM = 0> GS_DEV void d_form_q_inplace(double* H, int64_t ld, int n, const double* tau, double* wk) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  for (int k = n - 2; k >= 0; --k) {
    const int c = k + 1;
    const double t = tau[k];
    // s_j = t (Q[c, j] + sum_{r>c} v_r Q[r, j]); column j -= s_j [1; v]
    ddots4<NT>(c + 1, n, c + 1, n, [&](int r, int j) { return H[r + (int64_t)k * ld] * H[r + (int64_t)j * ld]; }, wk);
    __syncthreads();
    for (int j = c + 1 + tid; j < n; j += NT) wk[j] = t * (H[c + (int64_t)j * ld] + wk[j]);
    __syncthreads();
    if constexpr (RM > 0) {
      double vr[RM];
#pragma unroll
      for (int q = 0; q < RM; ++q) {
        const int r = c + lane + 64 * q;
        vr[q] = (r < n) ? ((r == c) ? 1.0 : H[r + (int64_t)k * ld]) : 0.0;
      }
      #pragma unroll 4
      for (int j = c + 1 + wave; j < n; j += NT / 64) {
        const double sj = wk[j];
#pragma unroll
        for (int q = 0; q < RM; ++q) {
          const int r = c + lane + 64 * q;
          if (r < n) H[r + (int64_t)j * ld] -= sj * vr[q];
        }
      }
    } else {
      #pragma unroll 4
      for (int j = c + 1 + wave; j < n; j += NT / 64) {
        const double sj = wk[j];
        for (int r = c + lane; r < n; r += 64) {
          const double h = H[r + (int64_t)j * ld];
          H[r + (int64_t)j * ld] = (r == c) ? h - sj : h - sj * H[r + (int64_t)k * ld];
        }
      }
    }
    __syncthreads();
    for (int r = tid; r < n; r += NT) {
      double q;
      if (r < c) q = 0.0;
      else if (r == c) q = 1.0 - t;
      else q = -(t * H[r + (int64_t)k * ld]);
      H[r + (int64_t)c * ld] = q;
    }
    __syncthreads();
  }
  for (int r = tid; r < n; r += NT) H[r] = (r == 0) ? 1.0 : 0.0;
  __syncthreads();
}

}  // namespace gs
