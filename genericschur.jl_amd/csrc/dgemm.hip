// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64), used by the
// blocked Hessenberg reduction and Q formation of large matrices:
//
//   C(m x n) = beta * C + alpha * op(A) * op(B),   op(X) = X or X^T,
//
// all column-major.  A workgroup of 256 threads (4 wavefronts) owns a BM x BN
// tile of C, the wavefronts arranged WR x (4 / WR); each wavefront owns
// (BM / WR) x (BN * WR / 4) as 16 x 16 MFMA blocks.  K advances in slabs of
// 16 staged through LDS (k-major, padded); the next slab's global loads are
// issued before the current slab's MFMAs.
//
//   64 x 64 tiles (WR = 2) for the trailing updates (large m and n, k = 32);
//   128 x 128 tiles (WR = 2) for large m, n and k;
//   32 x 128 tiles (WR = 1), split over K, for the W = V^T C products
//   (m = 32, long k) of the left update and of the Q formation;
//   128 x 32 tiles (WR = 4), split over K, for Y's top rows A V (n = 32,
//   long k).
//
// Operand layout of v_mfma_f64_16x16x4_f64 (lane l, 0 <= l < 64):
//   A (16 x 4):  lane l supplies A[l % 16][l / 16]
//   B (4 x 16):  lane l supplies B[l / 16][l % 16]
//   D (16 x 16): lane l holds D[l / 16 + 4 r][l % 16] in register r
//                (established on gfx950 by tests/test_gpu_large.py::test_dgemm)
// Each block is computed transposed, (A B)^T = B^T A^T, so that the 16 lanes
// that share an accumulator register hold 16 consecutive rows of C.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "internal.h"

namespace gs {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int GBK = 16, GPAD = 16;

template <bool TA, bool TB, int BM, int BN, int WR>
__global__ __launch_bounds__(256) void dgemm_kernel(int m, int n, int k, int kchunk, double alpha,
                                                   const double* __restrict__ A, int64_t lda,
                                                   const double* __restrict__ B, int64_t ldb, double beta,
                                                   double* __restrict__ C, int64_t ldc, int64_t split_stride) {
  constexpr int WC = 4 / WR;
  constexpr int WM = BM / WR, WN = BN / WC;  // wavefront tile
  constexpr int IM = WM / 16, JN = WN / 16;  // MFMA blocks per wavefront
  constexpr int LDA_ = BM + GPAD, LDB_ = BN + GPAD;
  constexpr int NA = BM * GBK / 256, NB = BN * GBK / 256;  // staged elements per thread
  static_assert(IM >= 1 && JN >= 1 && NA >= 1 && NB >= 1, "tile shape");
  __shared__ double As[GBK * LDA_];
  __shared__ double Bs[GBK * LDB_];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int wr = (wave % WR) * WM, wc = (wave / WR) * WN;
  // split over K: blockIdx.z takes the k range [kb, ke) and writes its
  // partial product to C + z * split_stride
  const int kb = blockIdx.z * kchunk, ke = min(k, kb + kchunk);
  C += (int64_t)blockIdx.z * split_stride;
  v4d acc[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

  double ra[NA], rb[NB];
  // op(A)[m0 + i, k0 + kk] and op(B)[k0 + kk, n0 + j]; the thread-to-element
  // maps walk the contiguous direction of the stored operand
  auto a_idx = [&](int r, int& i, int& kk) {
    if (!TA) {
      i = tid % BM;
      kk = tid / BM + (256 / BM) * r;
    } else {
      kk = tid & (GBK - 1);
      i = (tid >> 4) + 16 * r;
    }
  };
  auto b_idx = [&](int r, int& j, int& kk) {
    if (TB) {
      j = tid % BN;
      kk = tid / BN + (256 / BN) * r;
    } else {
      kk = tid & (GBK - 1);
      j = (tid >> 4) + 16 * r;
    }
  };
  auto fetch = [&](int k0) {
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      int i, kk;
      a_idx(r, i, kk);
      const int gi = m0 + i, gk = k0 + kk;
      double v = 0.0;
      if (gi < m && gk < ke) v = TA ? A[(int64_t)gk + (int64_t)gi * lda] : A[(int64_t)gi + (int64_t)gk * lda];
      ra[r] = v;
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      int j, kk;
      b_idx(r, j, kk);
      const int gj = n0 + j, gk = k0 + kk;
      double v = 0.0;
      if (gj < n && gk < ke) v = TB ? B[(int64_t)gj + (int64_t)gk * ldb] : B[(int64_t)gk + (int64_t)gj * ldb];
      rb[r] = v;
    }
  };
  auto stage = [&]() {
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      int i, kk;
      a_idx(r, i, kk);
      As[kk * LDA_ + i] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      int j, kk;
      b_idx(r, j, kk);
      Bs[kk * LDB_ + j] = rb[r];
    }
  };
  if (kb < ke) fetch(kb);
  for (int k0 = kb; k0 < ke; k0 += GBK) {
    stage();
    __syncthreads();
    if (k0 + GBK < ke) fetch(k0 + GBK);
#pragma unroll
    for (int k4 = 0; k4 < GBK / 4; ++k4) {
      const int kk = 4 * k4 + (lane >> 4);
      double a[IM], b[JN];
#pragma unroll
      for (int i = 0; i < IM; ++i) a[i] = As[kk * LDA_ + wr + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < JN; ++j) b[j] = Bs[kk * LDB_ + wc + 16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // register r of lane l holds C row 16 i + l % 16, column 16 j + l / 16 + 4 r
  // of the wavefront's block (i, j)
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      const int gi = m0 + wr + 16 * i + (lane & 15);
      if (gi >= m) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gj = n0 + wc + 16 * j + (lane >> 4) + 4 * r;
        if (gj >= n) continue;
        double* c = C + (int64_t)gi + (int64_t)gj * ldc;
        const double v = alpha * acc[i][j][r];
        *c = (beta == 0.0) ? v : beta * (*c) + v;
      }
    }
}

// C = beta C + the sum of the nsplit partial products (each m x n, ld = m)
__global__ __launch_bounds__(256) void dgemm_reduce_kernel(int m, int n, int nsplit, const double* __restrict__ P,
                                                           double beta, double* __restrict__ C, int64_t ldc) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)m * n) return;
  const int j = (int)(e / m), i = (int)(e - (int64_t)j * m);
  double s = 0.0;
  for (int q = 0; q < nsplit; ++q) s += P[(int64_t)q * m * n + e];
  double* c = C + (int64_t)i + (int64_t)j * ldc;
  *c = (beta == 0.0) ? s : beta * (*c) + s;
}

template <bool TA, bool TB, int BM, int BN, int WR>
static hipError_t launch_tile(int m, int n, int k, int nsplit, int kchunk, double alpha, const double* A, int64_t lda,
                              const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int64_t split_stride,
                              hipStream_t s) {
  dim3 grid((m + BM - 1) / BM, (n + BN - 1) / BN, nsplit);
  hipLaunchKernelGGL((dgemm_kernel<TA, TB, BM, BN, WR>), grid, dim3(256), 0, s, m, n, k, kchunk, alpha, A, lda, B,
                     ldb, beta, C, ldc, split_stride);
  return hipGetLastError();
}

template <int BM, int BN, int WR>
static hipError_t launch_shape(bool ta, bool tb, int m, int n, int k, int nsplit, int kchunk, double alpha,
                               const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                               int64_t ldc, int64_t split_stride, hipStream_t s) {
  if (!ta && !tb)
    return launch_tile<false, false, BM, BN, WR>(m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                                 split_stride, s);
  if (!ta && tb)
    return launch_tile<false, true, BM, BN, WR>(m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                                split_stride, s);
  if (ta && !tb)
    return launch_tile<true, false, BM, BN, WR>(m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                                split_stride, s);
  return launch_tile<true, true, BM, BN, WR>(m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                             split_stride, s);
}

// the tile shape for an m x n product: 32 x 128 for m <= 32, 128 x 32 for
// n <= 32, else 128 x 128
static hipError_t launch_any(bool ta, bool tb, int m, int n, int k, int nsplit, int kchunk, double alpha,
                             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                             int64_t ldc, int64_t split_stride, hipStream_t s) {
  if (m <= 32)
    return launch_shape<32, 128, 1>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                    split_stride, s);
  if (n <= 32)
    return launch_shape<128, 32, 4>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                    split_stride, s);
  // short K (the rank-32 trailing updates): the C read-modify-write is the
  // cost, so smaller tiles and more resident workgroups per CU
  if (k <= 64)
    return launch_shape<64, 64, 2>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                   split_stride, s);
  return launch_shape<128, 128, 2>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                   split_stride, s);
}

hipError_t launch_dgemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  return launch_any(ta, tb, m, n, k, 1, k, alpha, A, lda, B, ldb, beta, C, ldc, 0, s);
}

static int splitk_parts(int k) { return std::min(16, std::max(1, k / 256)); }

size_t dgemm_splitk_doubles(int m, int n, int k) { return (size_t)splitk_parts(k) * m * n; }

hipError_t launch_dgemm_splitk(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int64_t lda,
                               const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* ws,
                               hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  const int nsplit = splitk_parts(k);
  if (nsplit == 1 || (m > 32 && n > 32)) return launch_dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
  // K chunks are whole slabs; every launched split gets at least one
  const int kchunk = ((k + nsplit - 1) / nsplit + GBK - 1) / GBK * GBK;
  const int used = (k + kchunk - 1) / kchunk;
  hipError_t e = launch_any(ta, tb, m, n, k, used, kchunk, alpha, A, lda, B, ldb, 0.0, ws, m, (int64_t)m * n, s);
  if (e != hipSuccess) return e;
  const int64_t tot = (int64_t)m * n;
  hipLaunchKernelGGL(dgemm_reduce_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, m, n, used, ws, beta,
                     C, ldc);
  return hipGetLastError();
}

}  // namespace gs
