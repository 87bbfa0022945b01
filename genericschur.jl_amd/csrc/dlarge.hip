// Real Schur decomposition of one large matrix (order beyond the
// LDS-resident range), spread over the whole GPU.
//
//   1. Scaling (util.jl:14-29) with a grid-wide max |a_ij|.
//   2. Blocked Householder Hessenberg reduction: panels of LNB columns are
//      factored with the "Y = A V T" scheme (Quintana-Orti & van de Geijn,
//      ACM TOMS 32(2), 2006) and the trailing matrix is updated with
//      compact-WY block reflectors as FP64 MFMA GEMMs (dgemm.hip).  The
//      reflectors are those of hessenberg.jl:3-17 (xLARFG convention,
//      householder.jl:12-54); only the summation order differs.
//   3. Q = H_1 ... H_{n-2} formed backwards from the same block reflectors
//      (hessenberg.jl:150-195 hands this to LAPACK's orgqr).
//   4. Francis double-shift QR (GenericSchur.jl:513-699, 837-952).  One
//      workgroup runs the control logic (deflation, shifts, 2x2 standard
//      form, start row) between sweeps.  A sweep's bulge chase runs in
//      chunks of up to LKC steps: the chunk's window of H (rows and columns
//      k0-1 .. kend+3) sits in LDS of one wavefront that chases the bulge
//      and logs the reflectors, while the other workgroups of the same
//      launch apply the previous chunk's reflectors to the parts outside
//      its window (rows above it, columns right of it, and Z) as
//      independent chains, one lane per row or column.  Every element of H
//      and Z receives the same reflector updates in the same order and with
//      the same expressions as in the unblocked sweep, so the chunking does
//      not change the arithmetic.
#include <algorithm>
#include <cstring>

#include "dev_common.h"
#include "chains.h"
#include "dchase.h"
#include "gs2x2.h"
#include "internal.h"

namespace gs {

constexpr int LNB = 64;  // Hessenberg panel width
constexpr int LW = 64;   // bulge-chase window (rows and columns)
constexpr int LKC = LW - 4;  // bulge steps per chunk
constexpr int LWL = LW + 1;  // odd LDS pitch of a window column (bank spread)

// ---------------------------------------------------------------- scaling
template <int NT> __global__ void __launch_bounds__(NT) labsmax_kernel(const double* A, int64_t lda, int n,
                                                                       double* part) {
  __shared__ double red[NT / 64 + 1];
  double m = 0.0;
  for (int c = blockIdx.x; c < n; c += gridDim.x) {
    const double* col = A + (int64_t)c * lda;
    for (int r = threadIdx.x; r < n; r += NT) m = fmax(m, fabs(col[r]));
  }
  m = block_max<NT>(m, red);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// _scale! decision (util.jl:14-29): scl = (cscale, anrm), cscale = 0 when
// no scaling is needed
template <int NT> __global__ void __launch_bounds__(NT) lscale_decide_kernel(const double* part, int np, int scale,
                                                                             double* scl) {
  __shared__ double red[NT / 64 + 1];
  double m = 0.0;
  for (int i = threadIdx.x; i < np; i += NT) m = fmax(m, part[i]);
  const double an = block_max<NT>(m, red);
  if (threadIdx.x == 0) {
    const double smlnum = sqrt(GS_FMIN) / GS_EPS;
    const double bignum = 1.0 / smlnum;
    double cs = 0.0;
    if (scale) {
      if (an > 0.0 && an < smlnum) cs = smlnum;
      else if (an > bignum) cs = bignum;
    }
    scl[0] = cs;
    scl[1] = cs != 0.0 ? an : 0.0;
  }
}

// H = A (scaled by cscale / anrm with safescale!'s multipliers when asked)
template <int NT> __global__ void __launch_bounds__(NT) lcopy_scale_kernel(const double* A, int64_t lda, double* H,
                                                                           int64_t ldh, int n, const double* scl) {
  const double cs = scl[0], an = scl[1];
  double mul[4];
  const int nm = cs != 0.0 ? safescale_muls(an, cs, mul) : 0;
  for (int c = blockIdx.x; c < n; c += gridDim.x) {
    for (int r = threadIdx.x; r < n; r += NT) {
      double v = A[r + (int64_t)c * lda];
      for (int q = 0; q < nm; ++q) v *= mul[q];
      H[r + (int64_t)c * ldh] = v;
    }
  }
}

// ------------------------------------------------------- blocked Hessenberg
struct PanelArgs {
  double* H;
  int64_t ld;
  int n;
  int i;       // first column of the panel (0-based)
  int j;       // column of the panel being reduced
  double* Y;   // n x LNB, ld n
  double* V;   // n x LNB, ld n: explicit unit lower trapezoidal reflectors
  double* Tm;  // LNB x LNB, ld LNB: upper triangular block factor
  double* tau; // n
  double* tv;  // LNB: V(:, 0:j)^T v_j for the Y update
  double* part;  // partial products of the Y gemv, nch x n
};

// out[q] = sum_{r0 <= r < r1} V(r, q) x[r] for q < jq (block reduction)
template <int NT> GS_DEV void vt_times(const double* V, int n, const double* x, int r0, int r1, int jq, double* out,
                                       double* red) {
  // sixteen columns of V per pass (the accumulators stay in registers)
  constexpr int QB = 16;
  const int w = threadIdx.x >> 6;
  for (int qb = 0; qb < jq; qb += QB) {
    double acc[QB];
#pragma unroll
    for (int q = 0; q < QB; ++q) acc[q] = 0.0;
    for (int r = r0 + (int)threadIdx.x; r < r1; r += NT) {
      const double xv = x[r];
#pragma unroll
      for (int q = 0; q < QB; ++q)
        if (qb + q < jq) acc[q] += V[r + (int64_t)(qb + q) * n] * xv;
    }
#pragma unroll
    for (int q = 0; q < QB; ++q) {
      if (qb + q < jq) {
        const double s = wave_sum(acc[q]);
        if ((threadIdx.x & 63) == 0) red[w * LNB + qb + q] = s;
      }
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < jq) {
    double s = 0.0;
    for (int ww = 0; ww < NT / 64; ++ww) s += red[ww * LNB + threadIdx.x];
    out[threadIdx.x] = s;
  }
  __syncthreads();
}

// Panel column j (global column c = i + j): bring column c up to date with
// the panel's earlier reflectors (right: minus Y V(c,:)^T; left: minus
// V T^T V^T b), generate its reflector (hessenberg.jl:8-9 with
// householder.jl:12-54), and extend V, T and the inputs of Y's column j.
template <int NT> __global__ void __launch_bounds__(NT) lpanel_col_kernel(PanelArgs p) {
  __shared__ double red[(NT / 64) * LNB];
  __shared__ double redm[NT / 64 + 1];
  __shared__ double vc[LNB], w[LNB], w2[LNB], tvs[LNB];
  const int tid = threadIdx.x;
  const int n = p.n, i = p.i, j = p.j, c = i + j;
  const int64_t ld = p.ld;
  double* col = p.H + (int64_t)c * ld;
  double* Tm = p.Tm;
  if (j > 0) {
    if (tid < j) vc[tid] = p.V[c + (int64_t)tid * n];
    __syncthreads();
    for (int r = i + 1 + tid; r < n; r += NT) {
      double s = col[r];
#pragma unroll 8
      for (int q = 0; q < j; ++q) s -= p.Y[r + (int64_t)q * n] * vc[q];
      col[r] = s;
    }
    __syncthreads();
    vt_times<NT>(p.V, n, col, i + 1, n, j, w, red);
    if (tid < j) {
      double s = 0.0;
      for (int q = 0; q <= tid; ++q) s += Tm[q + tid * LNB] * w[q];
      w2[tid] = s;
    }
    __syncthreads();
    for (int r = i + 1 + tid; r < n; r += NT) {
      double s = col[r];
#pragma unroll 8
      for (int q = 0; q < j; ++q) s -= p.V[r + (int64_t)q * n] * w2[q];
      col[r] = s;
    }
    __syncthreads();
  }
  // reflector annihilating col[c+2 .. n-1] against col[c+1]
  const double sfmin = 2.0 * GS_FMIN / GS_EPS;
  const int r0 = c + 2;
  double t_ = 0.0;
  if (n - 1 - c > 1) {
    double amax = 0.0;
    for (int r = r0 + tid; r < n; r += NT) amax = fmax(amax, fabs(col[r]));
    amax = block_max<NT>(amax, redm);
    double xnorm = 0.0;
    if (amax > 0.0) {
      const double ra = 1.0 / amax;
      double s = 0.0;
      for (int r = r0 + tid; r < n; r += NT) {
        const double a = col[r] * ra;
        s += a * a;
      }
      s = block_sum<NT>(s, redm);
      xnorm = amax * sqrt(s);
    }
    if (xnorm != 0.0) {
      double alpha = col[c + 1];
      double beta = -copysign(hypot(alpha, xnorm), alpha);
      int kount = 0;
      if (fabs(beta) < sfmin) {
        const double rsfmin = 1.0 / sfmin;
        bool smallb = true;
        while (smallb) {
          kount += 1;
          __syncthreads();
          for (int r = r0 + tid; r < n; r += NT) col[r] *= rsfmin;
          beta *= rsfmin;
          alpha *= rsfmin;
          smallb = (fabs(beta) < sfmin) && (kount < 20);
        }
        __syncthreads();
        double am2 = 0.0;
        for (int r = r0 + tid; r < n; r += NT) am2 = fmax(am2, fabs(col[r]));
        am2 = block_max<NT>(am2, redm);
        xnorm = 0.0;
        if (am2 > 0.0) {
          const double ra = 1.0 / am2;
          double s = 0.0;
          for (int r = r0 + tid; r < n; r += NT) {
            const double a = col[r] * ra;
            s += a * a;
          }
          s = block_sum<NT>(s, redm);
          xnorm = am2 * sqrt(s);
        }
        beta = -copysign(hypot(alpha, xnorm), alpha);
      }
      t_ = (beta - alpha) / beta;
      const double sc = 1.0 / (alpha - beta);
      __syncthreads();
      for (int r = r0 + tid; r < n; r += NT) col[r] *= sc;
      for (int q = 0; q < kount; ++q) beta *= sfmin;
      __syncthreads();
      if (tid == 0) col[c + 1] = beta;
    }
  }
  __syncthreads();
  double* vj = p.V + (int64_t)j * n;
  for (int r = tid; r < n; r += NT) vj[r] = (r <= c) ? 0.0 : (r == c + 1 ? 1.0 : col[r]);
  if (tid == 0) p.tau[c] = t_;
  __syncthreads();
  // t = V(:, 0:j)^T v_j and the new column of T (T(0:j, j) = -tau T t)
  vt_times<NT>(p.V, n, vj, c + 1, n, j, tvs, red);
  if (tid < LNB) {
    double val = 0.0;
    if (tid < j) {
      double s = 0.0;
      for (int q = tid; q < j; ++q) s += Tm[tid + q * LNB] * tvs[q];
      val = -t_ * s;
    } else if (tid == j) {
      val = t_;
    }
    Tm[tid + j * LNB] = val;
    if (tid < j) p.tv[tid] = tvs[tid];
  }
}

// part[ch][r] = sum over the 64 columns of chunk ch of H(r, col) v_j(col),
// columns c+1 .. n-1 (still those of the panel's starting matrix), rows
// i+1 .. n-1
template <int NT> __global__ void __launch_bounds__(NT) lpanel_gemv_kernel(PanelArgs p) {
  __shared__ double vs[64];
  const int n = p.n, i = p.i, j = p.j, c = i + j;
  const int col0 = c + 1 + blockIdx.y * 64;
  const int ncol = min(64, n - col0);
  if ((int)threadIdx.x < 64) vs[threadIdx.x] = ((int)threadIdx.x < ncol) ? p.V[col0 + threadIdx.x + (int64_t)j * n] : 0.0;
  __syncthreads();
  const int r = i + 1 + blockIdx.x * NT + threadIdx.x;
  if (r >= n) return;
  const double* Hc = p.H + r + (int64_t)col0 * p.ld;
  double s = 0.0;
  for (int q = 0; q < ncol; ++q) s += Hc[(int64_t)q * p.ld] * vs[q];
  p.part[(int64_t)blockIdx.y * n + r] = s;
}

// Y(r, j) = tau_c (A v_j - Y(:, 0:j) t)(r), rows i+1 .. n-1
template <int NT> __global__ void __launch_bounds__(NT) lpanel_yfin_kernel(PanelArgs p, int nch) {
  const int n = p.n, i = p.i, j = p.j, c = i + j;
  const int r = i + 1 + blockIdx.x * NT + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int ch = 0; ch < nch; ++ch) s += p.part[(int64_t)ch * n + r];
  for (int q = 0; q < j; ++q) s -= p.Y[r + (int64_t)q * n] * p.tv[q];
  p.Y[r + (int64_t)j * n] = p.tau[c] * s;
}

// Y(r, q) = sum_{s <= q} X(r, s) Tm(s, q) for rows r < m, q < ib
template <int NT> __global__ void __launch_bounds__(NT) ltrmm_right_upper_kernel(const double* X, int64_t ldx,
                                                                                 double* Y, int64_t ldy,
                                                                                 const double* Tm, int m, int ib) {
  __shared__ double ts[LNB * LNB];
  for (int e = threadIdx.x; e < LNB * LNB; e += NT) ts[e] = (e % LNB < ib && e / LNB < ib) ? Tm[e] : 0.0;
  __syncthreads();
  const int r = blockIdx.x * NT + threadIdx.x;
  if (r >= m) return;
  double x[LNB];
#pragma unroll
  for (int s = 0; s < LNB; ++s) x[s] = (s < ib) ? X[r + (int64_t)s * ldx] : 0.0;
#pragma unroll
  for (int q = 0; q < LNB; ++q) {
    double acc = 0.0;
#pragma unroll
    for (int s = 0; s <= q; ++s) acc += x[s] * ts[s + q * LNB];
    if (q < ib) Y[r + (int64_t)q * ldy] = acc;
  }
}

// Wo(:, col) = op(Tm) Wi(:, col) for the upper triangular ib x ib Tm,
// op = transpose (TRANS) or none; columns col < ncols, ld of Wi/Wo = LNB
template <int NT, bool TRANS> __global__ void __launch_bounds__(NT) ltrmm_left_upper_kernel(const double* Wi,
                                                                                            double* Wo,
                                                                                            const double* Tm,
                                                                                            int ib, int ncols) {
  __shared__ double ts[LNB * LNB];
  for (int e = threadIdx.x; e < LNB * LNB; e += NT) ts[e] = (e % LNB < ib && e / LNB < ib) ? Tm[e] : 0.0;
  __syncthreads();
  const int cidx = blockIdx.x * NT + threadIdx.x;
  if (cidx >= ncols) return;
  double x[LNB];
#pragma unroll
  for (int s = 0; s < LNB; ++s) x[s] = (s < ib) ? Wi[s + (int64_t)cidx * LNB] : 0.0;
#pragma unroll
  for (int q = 0; q < LNB; ++q) {
    double acc = 0.0;
    if (TRANS) {
#pragma unroll
      for (int s = 0; s <= q; ++s) acc += ts[s + q * LNB] * x[s];
    } else {
#pragma unroll
      for (int s = q; s < LNB; ++s) acc += ts[q + s * LNB] * x[s];
    }
    if (q < ib) Wo[q + (int64_t)cidx * LNB] = acc;
  }
}

// explicit V of the panel starting at column i (ib reflectors) from the
// factors stored below H's subdiagonal
template <int NT> __global__ void __launch_bounds__(NT) lbuild_v_kernel(const double* H, int64_t ld, int n, int i,
                                                                        int ib, double* V) {
  const int jj = blockIdx.y;
  if (jj >= ib) return;
  const int c = i + jj;
  for (int r = blockIdx.x * NT + threadIdx.x; r < n; r += gridDim.x * NT)
    V[r + (int64_t)jj * n] = (r <= c) ? 0.0 : (r == c + 1 ? 1.0 : H[r + (int64_t)c * ld]);
}

template <int NT> __global__ void __launch_bounds__(NT) lidentity_kernel(double* Z, int64_t ldz, int n) {
  for (int c = blockIdx.x; c < n; c += gridDim.x)
    for (int r = threadIdx.x; r < n; r += NT) Z[r + (int64_t)c * ldz] = (r == c) ? 1.0 : 0.0;
}

// triu!(HH, -1) (GenericSchur.jl:540)
template <int NT> __global__ void __launch_bounds__(NT) lzero_lower_kernel(double* H, int64_t ld, int n) {
  for (int c = blockIdx.x; c < n; c += gridDim.x)
    for (int r = c + 2 + threadIdx.x; r < n; r += NT) H[r + (int64_t)c * ld] = 0.0;
}

// ------------------------------------------------------------- QR phase
struct LState {
  int64_t iter, nsweep, nstep;
  int iend, istart, iterqr, iwcur, phase, status, mx, pad;
  double v0, v1, v2;
};

struct LQArgs {
  double* H;
  int64_t ld;
  int n;
  double* Z;
  int64_t ldz;
  int wantz;
  zd* W;
  LState* st;
  double* lg;  // 3 per bulge step of the current sweep
  int64_t maxiter;
};

// Control logic of gschur!(H::Hessenberg{<:AbstractFloat}) (GenericSchur.jl:
// 545-699) between sweeps: deflation search, 1x1 and 2x2 deflations (with
// _gs2x2! and its Givens rotation on H and Z), shifts, and the start row and
// first column of the next double-shift sweep (GenericSchur.jl:851-876).
// Resumes from and saves to the state record; returns when a sweep is due
// (status 0), when all eigenvalues are found (1) or at maxiter (2).
template <int NT> __global__ void __launch_bounds__(NT) lctrl_kernel(LQArgs a) {
  __shared__ int redi[NT / 64 + 1];
  const int tid = threadIdx.x;
  const int n = a.n;
  const int64_t ld = a.ld;
  double* Hg = a.H;
  auto H = [&](int r, int c) -> double& { return Hg[(r - 1) + (int64_t)(c - 1) * ld]; };
  const double eps = GS_EPS;
  const double smallnum = GS_FMIN * ((double)n / eps);
  LState s = *a.st;
  // launched again after the iteration ended (the host runs a few sweeps
  // ahead of the state it has seen): nothing to do
  if (s.status != 0) return;
  while (true) {
    if (s.phase == 0) {
      if (s.iend < 1) {
        s.status = 1;
        break;
      }
      s.istart = 1;
      s.iterqr = 0;
      s.phase = 1;
    }
    s.iter += 1;
    if (s.iter > a.maxiter) {
      s.status = 2;
      break;
    }
    const int iend = s.iend;
    int istart = s.istart;
    int found = -1;
    for (int k = iend - tid; k >= istart + 1; k -= NT) {
      bool split = fabs(H(k, k - 1)) < smallnum;
      if (!split) {
        const double Hkk = H(k, k);
        const double Hkm1km1 = H(k - 1, k - 1);
        double t = fabs(Hkm1km1) + fabs(Hkk);
        if (t == 0.0) {
          if (k > 2) t += fabs(H(k - 1, k - 2));
          if (k + 1 <= n) t += fabs(H(k + 1, k));
        }
        const double aHkkm1 = fabs(H(k, k - 1));
        if (aHkkm1 <= t * eps) {
          const double aHkm1k = fabs(H(k - 1, k));
          const double ab = fmax(aHkkm1, aHkm1k);
          const double ba = fmin(aHkkm1, aHkm1k);
          const double aa = fmax(fabs(Hkk), fabs(Hkm1km1 - Hkk));
          const double bb = fmin(fabs(Hkk), fabs(Hkm1km1 - Hkk));
          const double ss = aa + bb;
          split = ba * (ab / ss) <= fmax(smallnum, eps * (bb * (aa / ss)));
        }
      }
      if (split) {
        found = k;
        break;
      }
    }
    found = block_max_int<NT>(found, redi);
    istart = (found > 0) ? found : 1;
    s.istart = istart;
    if (istart > 1) {
      __syncthreads();
      if (tid == 0) H(istart, istart - 1) = 0.0;
      __syncthreads();
    }
    if (istart >= iend - 1) {
      // deflation (GenericSchur.jl:668-687)
      if (istart >= iend) {
        if (tid == 0) a.W[s.iwcur - 1] = zr(H(iend, iend));
        s.iwcur -= 1;
      } else {
        double aa = H(iend - 1, iend - 1), bb = H(iend - 1, iend), cc = H(iend, iend - 1), dd = H(iend, iend);
        double cs, sn;
        zd w1, w2;
        gs2x2_dev(aa, bb, cc, dd, cs, sn, w1, w2);
        if (tid == 0) {
          a.W[s.iwcur - 1] = w2;
          a.W[s.iwcur - 2] = w1;
        }
        s.iwcur -= 2;
        __syncthreads();
        for (int j = istart + tid; j <= n; j += NT) {
          const double a1 = H(iend - 1, j), a2 = H(iend, j);
          H(iend - 1, j) = cs * a1 + sn * a2;
          H(iend, j) = -sn * a1 + cs * a2;
        }
        __syncthreads();
        for (int i = 1 + tid; i <= iend; i += NT) {
          const double a1 = H(i, iend - 1), a2 = H(i, iend);
          H(i, iend - 1) = a1 * cs + a2 * sn;
          H(i, iend) = -a1 * sn + a2 * cs;
        }
        __syncthreads();
        if (tid == 0) {
          H(iend - 1, iend - 1) = aa;
          H(iend - 1, iend) = bb;
          H(iend, iend - 1) = cc;
          H(iend, iend) = dd;
          if (iend > 2) H(iend - 1, iend - 2) = 0.0;
        }
        if (a.wantz) {
          for (int i = 1 + tid; i <= n; i += NT) {
            double* z1 = a.Z + (i - 1) + (int64_t)(iend - 2) * a.ldz;
            double* z2 = a.Z + (i - 1) + (int64_t)(iend - 1) * a.ldz;
            const double a1 = *z1, a2 = *z2;
            *z1 = a1 * cs + a2 * sn;
            *z2 = -a1 * sn + a2 * cs;
          }
        }
        __syncthreads();
      }
      s.iend = istart - 1;
      s.phase = 0;
      continue;
    }
    // shifts (GenericSchur.jl:613-665)
    s.iterqr += 1;
    const double threeq = 0.75, m7_16 = -7.0 / 16.0;
    double H11, H12, H21, H22;
    if (s.iterqr == 10) {
      const double sx = fabs(H(istart + 1, istart)) + fabs(H(istart + 2, istart + 1));
      H11 = threeq * sx + H(istart, istart);
      H12 = m7_16 * sx;
      H21 = sx;
      H22 = H11;
    } else if (s.iterqr == 20) {
      const double sx = fabs(H(iend, iend - 1)) + fabs(H(iend - 1, iend - 2));
      H11 = threeq * sx + H(iend, iend);
      H12 = m7_16 * sx;
      H21 = sx;
      H22 = H11;
    } else {
      H11 = H(iend - 1, iend - 1);
      H21 = H(iend, iend - 1);
      H12 = H(iend - 1, iend);
      H22 = H(iend, iend);
    }
    const double sh = fabs(H11) + fabs(H12) + fabs(H21) + fabs(H22);
    double r1r, r2r, r1i, r2i;
    if (sh == 0.0) {
      r1r = r2r = r1i = r2i = 0.0;
    } else {
      H11 /= sh;
      H12 /= sh;
      H21 /= sh;
      H22 /= sh;
      const double tr = (H11 + H22) / 2.0;
      const double d = (H11 - tr) * (H22 - tr) - H12 * H21;
      const double rtd = sqrt(fabs(d));
      if (d >= 0.0) {
        r1r = tr * sh;
        r2r = r1r;
        r1i = rtd * sh;
        r2i = -r1i;
      } else {
        r1r = tr + rtd;
        r2r = tr - rtd;
        if (fabs(r1r - H22) <= fabs(r2r - H22)) {
          r1r = r1r * sh;
          r2r = r1r;
        } else {
          r2r = r2r * sh;
          r1r = r2r;
        }
        r1i = r2i = 0.0;
      }
    }
    s.nsweep += 1;
    // start row and first column of (H - s1)(H - s2) (GenericSchur.jl:851-876)
    auto firstcol = [&](int m, double& v0, double& v1, double& v2) {
      double H21s = H(m + 1, m);
      double sx = fabs(H(m, m) - r2r) + fabs(r2i) + fabs(H21s);
      H21s = H21s / sx;
      v0 = H21s * H(m, m + 1) + (H(m, m) - r1r) * ((H(m, m) - r2r) / sx) - r1i * (r2i / sx);
      v1 = H21s * (H(m, m) + H(m + 1, m + 1) - r1r - r2r);
      v2 = H21s * H(m + 2, m + 1);
      sx = fabs(v0) + fabs(v1) + fabs(v2);
      v0 = v0 / sx;
      v1 = v1 / sx;
      v2 = v2 / sx;
    };
    int fm = -1;
    for (int m = iend - 2 - tid; m >= istart + 1; m -= NT) {
      double a0, a1, a2;
      firstcol(m, a0, a1, a2);
      if (fabs(H(m, m - 1)) * (fabs(a1) + fabs(a2)) <=
          eps * fabs(a0) * (fabs(H(m - 1, m - 1)) + fabs(H(m, m)) + fabs(H(m + 1, m + 1)))) {
        fm = m;
        break;
      }
    }
    fm = block_max_int<NT>(fm, redi);
    const int mx = (fm > 0) ? fm : istart;
    double v0, v1, v2;
    firstcol(mx, v0, v1, v2);
    s.mx = mx;
    s.v0 = v0;
    s.v1 = v1;
    s.v2 = v2;
    s.nstep += iend - mx;
    s.status = 0;
    break;
  }
  __syncthreads();
  if (tid == 0) *a.st = s;
}


struct LChunk {
  int k0, kend;    // this chunk's bulge steps (k0 > kend: no chase, flush only)
  int pk0, pkend;  // the previous chunk's steps (pk0 > pkend: none)
  int istart, iend, mx;
};

// Chunk c (0-based) of the sweep the state record describes: bulge steps
// mx + c LKC .. (through iend - 1), the last one (c = number of chunks) the
// flush of the previous chunk's updates.  Returns false when the iteration
// has ended or c lies beyond the flush.
GS_DEV bool chunk_of(const LState* st, int c, LChunk& ck) {
  if (st->status != 0) return false;
  ck.istart = st->istart;
  ck.iend = st->iend;
  ck.mx = st->mx;
  const int nch = (ck.iend - ck.mx + LKC - 1) / LKC;  // steps mx .. iend - 1
  if (c > nch) return false;
  if (c < nch) {
    ck.k0 = ck.mx + c * LKC;
    ck.kend = min(ck.k0 + LKC - 1, ck.iend - 1);
  } else {
    ck.k0 = 1;
    ck.kend = 0;
  }
  if (c > 0) {
    ck.pk0 = ck.mx + (c - 1) * LKC;
    ck.pkend = min(ck.pk0 + LKC - 1, ck.iend - 1);
  } else {
    ck.pk0 = 1;
    ck.pkend = 0;
  }
  return true;
}

// One chunk of a sweep.  Workgroup 0 (one wavefront): apply the previous
// chunk's left update to the columns this window adds, load the window
// H(wb:we, wb:we) into LDS, chase the bulge through steps k0 .. kend
// (GenericSchur.jl:877-949, restricted to the window) and log the
// reflectors.  Workgroups 1..: the previous chunk's updates outside its
// window: columns right of the current window, rows above the previous
// window, and Z.  The regions are disjoint.
__global__ void __launch_bounds__(64) lchunk_kernel(LQArgs a, int c) {
  LChunk ck;
  if (!chunk_of(a.st, c, ck)) return;
  __shared__ double win[LWL * LW];
  // the chunk's reflectors, logged in LDS during the chase and written to
  // the global log after it: a global store per step would make every
  // step's LDS ordering point wait for the store to complete
  __shared__ double lgs[3 * LKC];
  const int lane = threadIdx.x;
  const int n = a.n, iend = ck.iend, mx = ck.mx;
  const int64_t ld = a.ld;
  const bool chase = ck.k0 <= ck.kend, prev = ck.pk0 <= ck.pkend;
  const int we = chase ? min(ck.kend + 3, iend) : 0;
  const int pwe = prev ? min(ck.pkend + 3, iend) : 0;
  const int plast = prev ? min(ck.pkend + 2, iend) : 0;
  // a full chunk of three-element reflectors: fixed-shape chains
  const bool pfull = prev && ck.pkend - ck.pk0 + 1 == LKC && ck.pkend <= iend - 2;
  if (blockIdx.x == 0) {
    if (!chase) return;
    if (prev) {
      const int col = pwe + 1 + lane;
      if (col <= we) {
        if (pfull)
          chain_line_full<LKC, 16>(a.H + (int64_t)(col - 1) * ld, 1, a.lg, mx, ck.pk0);
        else
          chain_line<LKC, 16>(a.H + (int64_t)(col - 1) * ld, 1, a.lg, mx, ck.pk0, ck.pkend, iend);
      }
      __syncthreads();
    }
    const int wb = max(ck.k0 - 1, 1);
    const int W = we - wb + 1;
    // the sweep's first reflector input, read before the barrier so that it
    // has landed when the chase starts
    const double sv0 = a.st->v0, sv1 = a.st->v1, sv2 = a.st->v2;
    win_load<double, LW, LWL>(a.H + (wb - 1) + (int64_t)(wb - 1) * ld, ld, W, win, lane);
    __syncthreads();
    dchase_chunk<LWL, true>(win, wb, we, ck.k0, ck.kend, mx, ck.istart, iend, sv0, sv1, sv2, lgs);
    __syncthreads();
    win_store<double, LW, LWL>(a.H + (wb - 1) + (int64_t)(wb - 1) * ld, ld, W, win, lane);
    for (int e = lane; e < 3 * (ck.kend - ck.k0 + 1); e += 64) a.lg[3 * (ck.k0 - mx) + e] = lgs[e];
    return;
  }
  if (!prev) return;
  // the previous chunk's updates outside both windows
  const int lcol0 = (chase ? we : pwe) + 1;
  const int pwb = max(ck.pk0 - 1, 1);
  const int nL = (n >= lcol0) ? (n - lcol0 + 64) / 64 : 0;
  const int nR = (pwb - 1 + 63) / 64;
  const int nZ = a.wantz ? (n + 63) / 64 : 0;
  const int ntask = nL + nR + nZ;
  for (int t = blockIdx.x - 1; t < ntask; t += gridDim.x - 1) {
    if (t < nL) {
      const int col = lcol0 + 64 * t + lane;
      if (col <= n) {
        if (pfull)
          chain_line_full<LKC, 16>(a.H + (int64_t)(col - 1) * ld, 1, a.lg, mx, ck.pk0);
        else
          chain_line<LKC, 16>(a.H + (int64_t)(col - 1) * ld, 1, a.lg, mx, ck.pk0, ck.pkend, iend);
      }
    } else if (t < nL + nR) {
      const int row = 1 + 64 * (t - nL) + lane;
      if (row <= pwb - 1) {
        if (pfull)
          chain_line_full<LKC, 16>(a.H + (row - 1), ld, a.lg, mx, ck.pk0);
        else
          chain_line<LKC, 16>(a.H + (row - 1), ld, a.lg, mx, ck.pk0, ck.pkend, iend);
      }
    } else {
      const int row = 1 + 64 * (t - nL - nR) + lane;
      if (row <= n) {
        if (pfull)
          chain_line_full<LKC, 16>(a.Z + (row - 1), a.ldz, a.lg, mx, ck.pk0);
        else
          chain_line<LKC, 16>(a.Z + (row - 1), a.ldz, a.lg, mx, ck.pk0, ck.pkend, iend);
      }
    }
  }
  (void)plast;
}

// T = triu(HH, -1) and safescale! of T and the values (GenericSchur.jl:
// 697, 830-833)
template <int NT> __global__ void __launch_bounds__(NT) lfinish_kernel(double* H, int64_t ld, int n, zd* W,
                                                                       const double* scl) {
  const double cs = scl[0], an = scl[1];
  double umul[4];
  const int nu = cs != 0.0 ? safescale_muls(cs, an, umul) : 0;
  for (int c = blockIdx.x; c < n; c += gridDim.x) {
    for (int r = threadIdx.x; r < n; r += NT) {
      double v = (r > c + 1) ? 0.0 : H[r + (int64_t)c * ld];
      for (int q = 0; q < nu; ++q) v *= umul[q];
      H[r + (int64_t)c * ld] = v;
    }
  }
  if (nu > 0 && blockIdx.x == 0) {
    for (int r = threadIdx.x; r < n; r += NT) {
      zd w = W[r];
      for (int q = 0; q < nu; ++q) w = w * umul[q];
      W[r] = w;
    }
  }
}

// ------------------------------------------------------------ host driver
static int lgrid(int n) { return std::min(std::max(n, 1), 2048); }

hipError_t large_schur_d(const LargeArgs& a, LargeWork& w, hipStream_t s, LargeTimer* tm) {
  const int n = (int)a.n;
  const int64_t ldt = a.ldt;
  double* H = a.T;
  hipError_t e;
#define LCHK(x)                     \
  do {                              \
    if ((e = (x)) != hipSuccess) {  \
      w.where = #x;                 \
      return e;                     \
    }                               \
  } while (0)
#define LLAUNCH(k, ...)                           \
  do {                                            \
    hipLaunchKernelGGL(k, __VA_ARGS__);           \
    if ((e = hipGetLastError()) != hipSuccess) {  \
      w.where = #k;                               \
      return e;                                   \
    }                                             \
  } while (0)
  const bool wantz = (a.flags & 0x01) != 0;
  const bool scale = (a.flags & 0x02) != 0;
  double* scl = w.scl;
  // 1. copy and scale
  if (tm) LCHK(tm->mark(s, 0));
  const int ng = lgrid(n);
  LLAUNCH(labsmax_kernel<256>, dim3(ng), dim3(256), 0, s, a.A, a.lda, n, w.part);
  LLAUNCH(lscale_decide_kernel<256>, dim3(1), dim3(256), 0, s, w.part, ng, scale ? 1 : 0, scl);
  LLAUNCH(lcopy_scale_kernel<256>, dim3(ng), dim3(256), 0, s, a.A, a.lda, H, ldt, n, scl);
  if (tm) LCHK(tm->mark(s, 1));
  // 2. blocked Hessenberg
  PanelArgs p;
  p.H = H;
  p.ld = ldt;
  p.n = n;
  p.Y = w.Y;
  p.V = w.V;
  p.tau = w.tau;
  p.tv = w.tv;
  p.part = w.part;
  const int npanel = (n - 1 + LNB - 1) / LNB;  // columns 0 .. n-2
  for (int pi = 0; pi < npanel; ++pi) {
    const int i = pi * LNB;
    const int ib = std::min(LNB, n - 1 - i);
    p.i = i;
    p.Tm = w.Tall + (int64_t)pi * LNB * LNB;
    for (int j = 0; j < ib; ++j) {
      p.j = j;
      const int c = i + j;
      LLAUNCH(lpanel_col_kernel<1024>, dim3(1), dim3(1024), 0, s, p);
      const int nrow = n - i - 1;
      const int nch = (n - c - 1 + 63) / 64;
      if (nrow > 0 && nch > 0) {
        LLAUNCH(lpanel_gemv_kernel<256>, dim3((nrow + 255) / 256, nch), dim3(256), 0, s, p);
      }
      if (nrow > 0) LLAUNCH(lpanel_yfin_kernel<256>, dim3((nrow + 255) / 256), dim3(256), 0, s, p, nch);
    }
    // rows 0..i of Y: Y(0:i+1, :) = H(0:i+1, i+1:n) V(i+1:n, :) T
    LCHK(tm_gemm(tm, s, false, false, i + 1, ib, n - i - 1, 1.0, H + (int64_t)(i + 1) * ldt, ldt, w.V + (i + 1), n,
                 0.0, w.Ytmp, n, w.gk));
    LLAUNCH(ltrmm_right_upper_kernel<256>, dim3((i + 1 + 255) / 256), dim3(256), 0, s, w.Ytmp, (int64_t)n, w.Y,
            (int64_t)n, p.Tm, i + 1, ib);
    // rows 0..i of the panel's columns i+1 .. i+ib-1
    if (ib > 1)
      LCHK(tm_gemm(tm, s, false, true, i + 1, ib - 1, ib, -1.0, w.Y, n, w.V + (i + 1), n, 1.0,
                   H + (int64_t)(i + 1) * ldt, ldt));
    const int ntr = n - i - ib;
    if (ntr > 0) {
      // right update of the trailing columns: H(:, i+ib:n) -= Y V(i+ib:n, :)^T
      LCHK(tm_gemm(tm, s, false, true, n, ntr, ib, -1.0, w.Y, n, w.V + (i + ib), n, 1.0,
                   H + (int64_t)(i + ib) * ldt, ldt));
      // left update: H(i+1:n, i+ib:n) -= V T^T (V^T H(i+1:n, i+ib:n))
      LCHK(tm_gemm(tm, s, true, false, ib, ntr, n - i - 1, 1.0, w.V + (i + 1), n,
                   H + (i + 1) + (int64_t)(i + ib) * ldt, ldt, 0.0, w.Wb, LNB, w.gk));
      LLAUNCH((ltrmm_left_upper_kernel<256, true>), dim3((ntr + 255) / 256), dim3(256), 0, s, w.Wb, w.W2, p.Tm, ib,
              ntr);
      LCHK(tm_gemm(tm, s, false, false, n - i - 1, ntr, ib, -1.0, w.V + (i + 1), n, w.W2, LNB, 1.0,
                   H + (i + 1) + (int64_t)(i + ib) * ldt, ldt));
    }
  }
  if (tm) LCHK(tm->mark(s, 2));
  // 3. Q = B_0 B_1 ... B_last, accumulated backwards into Z
  if (wantz) {
    LLAUNCH(lidentity_kernel<256>, dim3(ng), dim3(256), 0, s, a.Z, a.ldz, n);
    for (int pi = npanel - 1; pi >= 0; --pi) {
      const int i = pi * LNB;
      const int ib = std::min(LNB, n - 1 - i);
      const double* Tm = w.Tall + (int64_t)pi * LNB * LNB;
      const int m = n - i - 1;
      LLAUNCH(lbuild_v_kernel<256>, dim3((n + 255) / 256, ib), dim3(256), 0, s, H, ldt, n, i, ib, w.V);
      double* Zs = a.Z + (i + 1) + (int64_t)(i + 1) * a.ldz;
      LCHK(tm_gemm(tm, s, true, false, ib, m, m, 1.0, w.V + (i + 1), n, Zs, a.ldz, 0.0, w.Wb, LNB, w.gk));
      LLAUNCH((ltrmm_left_upper_kernel<256, false>), dim3((m + 255) / 256), dim3(256), 0, s, w.Wb, w.W2, Tm, ib, m);
      LCHK(tm_gemm(tm, s, false, false, m, m, ib, -1.0, w.V + (i + 1), n, w.W2, LNB, 1.0, Zs, a.ldz));
    }
  }
  if (tm) LCHK(tm->mark(s, 3));
  LLAUNCH(lzero_lower_kernel<256>, dim3(ng), dim3(256), 0, s, H, ldt, n);
  // 4. QR iterations
  LQArgs q;
  q.H = H;
  q.ld = ldt;
  q.n = n;
  q.Z = a.Z;
  q.ldz = a.ldz;
  q.wantz = wantz ? 1 : 0;
  q.W = reinterpret_cast<zd*>(a.W);
  q.st = static_cast<LState*>(w.st);
  q.lg = w.lg;
  q.maxiter = a.maxiter > 0 ? a.maxiter : 100 * (int64_t)n;
  LState st0;
  std::memset(&st0, 0, sizeof st0);
  st0.iend = n;
  st0.iwcur = n;
  st0.istart = 1;
  LCHK(hipMemcpyAsync(q.st, &st0, sizeof st0, hipMemcpyHostToDevice, s));
  const int nwork = (n + 63) / 64;
  const int grid = 1 + std::min(3 * nwork, 1024);
  // Sweep control stays on the device: each sweep is lctrl_kernel and then
  // chunk launches that read their steps from the state record it wrote.
  // The host does not wait for a sweep's state: it issues up to kAhead
  // sweeps ahead of the newest state it has seen (copied back with an event
  // per sweep), launching for each the chunks an upper bound allows: iend
  // never grows, so ceil((iend - 1) / LKC) chunks of a state seen earlier
  // cover any later sweep (surplus launches return at once).  Launches after
  // the iteration has ended return at once as well.
  constexpr int kAhead = 4;
  LState* hst = static_cast<LState*>(w.host_state);  // kAhead slots
  hipEvent_t ev[kAhead];
  for (int i = 0; i < kAhead; ++i) LCHK(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
  auto fail = [&](hipError_t err) {
    for (int i = 0; i < kAhead; ++i) (void)hipEventDestroy(ev[i]);
    return err;
  };
  int status = 0, seen_iend = n;
  int64_t issued = 0, done = 0;  // sweeps issued / whose state was read
  LState last;
  std::memset(&last, 0, sizeof last);
  while (true) {
    if (issued - done == kAhead) {
      // read the oldest outstanding state
      const int slot = (int)(done % kAhead);
      if ((e = hipEventSynchronize(ev[slot])) != hipSuccess) return fail(e);
      last = hst[slot];
      done += 1;
      status = last.status;
      seen_iend = std::min(seen_iend, last.iend);
      if (status != 0) break;
    }
    const int slot = (int)(issued % kAhead);
    hipLaunchKernelGGL(lctrl_kernel<1024>, dim3(1), dim3(1024), 0, s, q);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    const int nch = (seen_iend - 1 + LKC - 1) / LKC;
    for (int c = 0; c <= nch; ++c) {
      hipLaunchKernelGGL(lchunk_kernel, dim3(grid), dim3(64), 0, s, q, c);
      if ((e = hipGetLastError()) != hipSuccess) return fail(e);
    }
    if ((e = hipMemcpyAsync(hst + slot, q.st, sizeof(LState), hipMemcpyDeviceToHost, s)) != hipSuccess) return fail(e);
    if ((e = hipEventRecord(ev[slot], s)) != hipSuccess) return fail(e);
    issued += 1;
  }
  // the sweeps issued after the last one read return at once; wait for them
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return fail(e);
  for (int i = 0; i < kAhead; ++i) (void)hipEventDestroy(ev[i]);
  hst[0] = last;
  if (tm) LCHK(tm->mark(s, 4));
  // 5. results
  const int32_t info = status == 2 ? (int32_t)std::min<int64_t>(hst->iter, 0x7fffffff) : 0;
  const int64_t stats[2] = {hst->nsweep, hst->nstep};
  LCHK(hipMemcpyAsync(a.info, &info, sizeof info, hipMemcpyHostToDevice, s));
  if (a.stats) LCHK(hipMemcpyAsync(a.stats, stats, sizeof stats, hipMemcpyHostToDevice, s));
  if (status == 1) LLAUNCH(lfinish_kernel<256>, dim3(ng), dim3(256), 0, s, H, ldt, n, q.W, scl);
  LCHK(hipStreamSynchronize(s));
  if (tm) LCHK(tm->mark(s, 5));
#undef LLAUNCH
#undef LCHK
  return hipSuccess;
}

size_t large_work_doubles(int64_t n, LargeWorkSizes* sz) {
  const int64_t npanel = (n - 1 + LNB - 1) / LNB;
  sz->Y = n * LNB;
  sz->V = n * LNB;
  sz->Ytmp = n * LNB;
  sz->Tall = std::max<int64_t>(npanel, 1) * LNB * LNB;
  sz->Wb = (int64_t)LNB * n;
  sz->W2 = (int64_t)LNB * n;
  sz->tau = n;
  sz->tv = LNB;
  sz->part = std::max<int64_t>((n + 63) / 64 * n, 2048);
  sz->scl = 2;
  sz->lg = 3 * n + 3;
  sz->st = (sizeof(LState) + 7) / 8;
  sz->gk = (int64_t)dgemm_splitk_doubles(LNB, (int)n, (int)n);
  // every segment is carved at a multiple of 32 doubles (256 bytes)
  int64_t* segs[] = {&sz->Y, &sz->V, &sz->Ytmp, &sz->Tall, &sz->Wb, &sz->W2, &sz->tau,
                     &sz->tv, &sz->part, &sz->scl, &sz->lg, &sz->st, &sz->gk};
  int64_t total = 0;
  for (int64_t* q : segs) {
    *q = (*q + 31) / 32 * 32;
    total += *q;
  }
  return (size_t)total;
}

size_t large_state_bytes() { return sizeof(LState); }

}  // namespace gs
