// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64), used by the
// blocked Hessenberg reduction and Q formation of large matrices:
//
//   C(m x n) = beta * C + alpha * op(A) * op(B),   op(X) = X or X^T,
//
// all column-major.  A workgroup of 256 threads (4 wavefronts) owns a
// 128 x 128 tile of C; each wavefront owns a 64 x 64 quarter as 4 x 4 MFMA
// blocks of 16 x 16 (16 accumulators of 4 doubles per lane).  K advances in
// slabs of 16 staged through LDS (k-major, padded so the 16-lane row groups
// of one MFMA operand read fall into different banks).
//
// Operand layout of v_mfma_f64_16x16x4_f64 (lane l, 0 <= l < 64):
//   A (16 x 4):  lane l supplies A[l % 16][l / 16]
//   B (4 x 16):  lane l supplies B[l / 16][l % 16]
//   D (16 x 16): lane l holds D[4 * (l / 16) + r][l % 16] in register r.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.h"

namespace gs {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int GBM = 128, GBN = 128, GBK = 16, GPAD = 16;
constexpr int GLD = GBM + GPAD;  // LDS pitch (doubles) of one k row

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void dgemm_kernel(int m, int n, int k, double alpha, const double* __restrict__ A,
                                                   int64_t lda, const double* __restrict__ B, int64_t ldb,
                                                   double beta, double* __restrict__ C, int64_t ldc) {
  __shared__ double As[GBK * GLD];
  __shared__ double Bs[GBK * GLD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * GBM, n0 = blockIdx.y * GBN;
  const int wr = (wave & 1) * 64, wc = (wave >> 1) * 64;
  v4d acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  // the next slab's global loads are issued before the current slab's MFMAs
  double ra[(GBM * GBK) / 256], rb[(GBN * GBK) / 256];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int r = 0; r < (GBM * GBK) / 256; ++r) {
      int i, kk;
      if (!TA) {
        i = tid & (GBM - 1);
        kk = (tid >> 7) + 2 * r;
      } else {
        kk = tid & (GBK - 1);
        i = (tid >> 4) + 16 * r;
      }
      const int gi = m0 + i, gk = k0 + kk;
      double v = 0.0;
      if (gi < m && gk < k) v = TA ? A[(int64_t)gk + (int64_t)gi * lda] : A[(int64_t)gi + (int64_t)gk * lda];
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < (GBN * GBK) / 256; ++r) {
      int j, kk;
      if (TB) {
        j = tid & (GBN - 1);
        kk = (tid >> 7) + 2 * r;
      } else {
        kk = tid & (GBK - 1);
        j = (tid >> 4) + 16 * r;
      }
      const int gj = n0 + j, gk = k0 + kk;
      double v = 0.0;
      if (gj < n && gk < k) v = TB ? B[(int64_t)gj + (int64_t)gk * ldb] : B[(int64_t)gk + (int64_t)gj * ldb];
      rb[r] = v;
    }
  };
  // op(A)[m0:m0+128, k0:k0+16] goes to As[kk][i], op(B)[k0:k0+16, n0:n0+128] to Bs[kk][j]
  auto stage = [&]() {
#pragma unroll
    for (int r = 0; r < (GBM * GBK) / 256; ++r) {
      const int i = TA ? (tid >> 4) + 16 * r : tid & (GBM - 1);
      const int kk = TA ? tid & (GBK - 1) : (tid >> 7) + 2 * r;
      As[kk * GLD + i] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < (GBN * GBK) / 256; ++r) {
      const int j = TB ? tid & (GBN - 1) : (tid >> 4) + 16 * r;
      const int kk = TB ? (tid >> 7) + 2 * r : tid & (GBK - 1);
      Bs[kk * GLD + j] = rb[r];
    }
  };
  if (k > 0) fetch(0);
  for (int k0 = 0; k0 < k; k0 += GBK) {
    stage();
    __syncthreads();
    if (k0 + GBK < k) fetch(k0 + GBK);
#pragma unroll
    for (int k4 = 0; k4 < GBK / 4; ++k4) {
      const int kk = 4 * k4 + (lane >> 4);
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk * GLD + wr + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk * GLD + wc + 16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // write back: lane holds rows 4 * (lane / 16) + r, column lane % 16 of each block
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int gj = n0 + wc + 16 * j + (lane & 15);
      if (gj >= n) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gi = m0 + wr + 16 * i + 4 * (lane >> 4) + r;
        if (gi >= m) continue;
        double* c = C + (int64_t)gi + (int64_t)gj * ldc;
        const double v = alpha * acc[i][j][r];
        *c = (beta == 0.0) ? v : beta * (*c) + v;
      }
    }
}

hipError_t launch_dgemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  dim3 grid((m + GBM - 1) / GBM, (n + GBN - 1) / GBN);
  if (!ta && !tb)
    hipLaunchKernelGGL((dgemm_kernel<false, false>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C,
                       ldc);
  else if (!ta && tb)
    hipLaunchKernelGGL((dgemm_kernel<false, true>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C,
                       ldc);
  else if (ta && !tb)
    hipLaunchKernelGGL((dgemm_kernel<true, false>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C,
                       ldc);
  else
    hipLaunchKernelGGL((dgemm_kernel<true, true>), grid, dim3(256), 0, s, m, n, k, alpha, A, lda, B, ldb, beta, C,
                       ldc);
  return hipGetLastError();
}

}  // namespace gs
