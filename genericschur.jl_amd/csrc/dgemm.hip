// FP64 GEMM on the CDNA4 matrix cores (v_mfma_f64_16x16x4_f64), used by the
// blocked Hessenberg reduction and Q formation of large matrices:
//
//   C(m x n) = beta * C + alpha * op(A) * op(B),   op(X) = X or X^T,
//
// all column-major.  A workgroup of 256 threads (4 wavefronts) owns a BM x BN
// tile of C, the wavefronts arranged WR x (4 / WR); each wavefront owns
// (BM / WR) x (BN * WR / 4) as 16 x 16 MFMA blocks.  K advances in slabs of
// 32 (16 for tiles with a side of 128) staged through two LDS buffers
// (k-major, padded); a ring of registers keeps the global loads of the next
// slabs in flight during the current slab's MFMAs, and the small tiles
// read their part of C before the K loop.
//
//   64 x 64 tiles (WR = 2) for the trailing updates (large m and n, k = 32);
//   128 x 128 tiles (WR = 2) for large m, n and k;
//   32 x 128 / 64 x 128 (WR = 1) and 128 x 32 / 128 x 64 (WR = 4) tiles
//   for products with a short side;
//   split over K, 32 x 64 or 64 x 64 tiles (WR = 1) for the W = V^T C
//   products (m = 32 or 64, long k) of the left update and of the Q
//   formation, 64 x 32 or 64 x 64 tiles (WR = 4) for Y's top rows A V
//   (n = 32 or 64, long k).
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

constexpr int GPAD = 16;

// Slab depth: 32 for the tiles up to 64 x 64, 16 for those with a side of 128
// (whose double-buffered slabs would not leave room for a second workgroup).
template <int BM, int BN> constexpr int slab_depth() { return (BM > 64 || BN > 64) ? 16 : 32; }

template <bool TA, bool TB, int BM, int BN, int WR, int PD>
__global__ __launch_bounds__(256) void dgemm_kernel(int m, int n, int k, int kchunk, double alpha,
                                                   const double* __restrict__ A, int64_t lda,
                                                   const double* __restrict__ B, int64_t ldb, double beta,
                                                   double* __restrict__ C, int64_t ldc, int64_t split_stride) {
  constexpr int KB = slab_depth<BM, BN>();
  constexpr int WC = 4 / WR;
  constexpr int WM = BM / WR, WN = BN / WC;  // wavefront tile
  constexpr int IM = WM / 16, JN = WN / 16;  // MFMA blocks per wavefront
  constexpr int LDA_ = BM + GPAD, LDB_ = BN + GPAD;
  constexpr int NA = BM * KB / 256, NB = BN * KB / 256;  // staged elements per thread
  // C read ahead of the K loop for the small tiles (the rank-k updates)
  constexpr bool CPRE = IM * JN <= 4;
  static_assert(IM >= 1 && JN >= 1 && NA >= 1 && NB >= 1 && PD % 2 == 0, "tile shape");
  // two LDS buffers: slab s + 1 is staged into one while slab s's MFMAs
  // read the other, one barrier per slab
  __shared__ double As[2][KB * LDA_];
  __shared__ double Bs[2][KB * LDB_];
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

  // op(A)[m0 + i, k0 + kk] and op(B)[k0 + kk, n0 + j]; the thread-to-element
  // maps walk the contiguous direction of the stored operand
  auto a_idx = [&](int r, int& i, int& kk) {
    if (!TA) {
      i = tid % BM;
      kk = tid / BM + (256 / BM) * r;
    } else {
      kk = tid % KB;
      i = tid / KB + (256 / KB) * r;
    }
  };
  auto b_idx = [&](int r, int& j, int& kk) {
    if (TB) {
      j = tid % BN;
      kk = tid / BN + (256 / BN) * r;
    } else {
      kk = tid % KB;
      j = tid / KB + (256 / KB) * r;
    }
  };
  // PD slabs in flight: slab s lives in ring slot s % PD from its fetch
  // until it is staged into LDS
  double ra[PD][NA], rb[PD][NB];
  // Loads are unconditional, from clamped in-range addresses (a load under a
  // branch would make the compiler wait for every load in flight), and the
  // entries outside the operands are zeroed when staged.
  auto fetch = [&](int k0, double* xa, double* xb) {
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      int i, kk;
      a_idx(r, i, kk);
      const int gi = min(m0 + i, m - 1), gk = min(k0 + kk, ke - 1);
      xa[r] = TA ? A[(int64_t)gk + (int64_t)gi * lda] : A[(int64_t)gi + (int64_t)gk * lda];
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      int j, kk;
      b_idx(r, j, kk);
      const int gj = min(n0 + j, n - 1), gk = min(k0 + kk, ke - 1);
      xb[r] = TB ? B[(int64_t)gj + (int64_t)gk * ldb] : B[(int64_t)gk + (int64_t)gj * ldb];
    }
  };
  auto stage = [&](int k0, const double* xa, const double* xb, int buf) {
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      int i, kk;
      a_idx(r, i, kk);
      As[buf][kk * LDA_ + i] = (m0 + i < m && k0 + kk < ke) ? xa[r] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NB; ++r) {
      int j, kk;
      b_idx(r, j, kk);
      Bs[buf][kk * LDB_ + j] = (n0 + j < n && k0 + kk < ke) ? xb[r] : 0.0;
    }
  };
  auto mma = [&](int buf) {
#pragma unroll
    for (int k4 = 0; k4 < KB / 4; ++k4) {
      const int kk = 4 * k4 + (lane >> 4);
      double a[IM], b[JN];
#pragma unroll
      for (int i = 0; i < IM; ++i) a[i] = As[buf][kk * LDA_ + wr + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < JN; ++j) b[j] = Bs[buf][kk * LDB_ + wc + 16 * j + (lane & 15)];
#pragma unroll
      for (int i = 0; i < IM; ++i)
#pragma unroll
        for (int j = 0; j < JN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j], a[i], acc[i][j], 0, 0, 0);
    }
  };
  // register r of lane l holds C row 16 i + l % 16, column 16 j + l / 16 + 4 r
  // of the wavefront's block (i, j) (clamped in-range reads; beta = 0 never
  // reads C)
  double cpre[CPRE ? IM : 1][CPRE ? JN : 1][4];
  if (CPRE && beta != 0.0) {
#pragma unroll
    for (int i = 0; i < IM; ++i)
#pragma unroll
      for (int j = 0; j < JN; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gi = min(m0 + wr + 16 * i + (lane & 15), m - 1);
          const int gj = min(n0 + wc + 16 * j + (lane >> 4) + 4 * r, n - 1);
          cpre[CPRE ? i : 0][CPRE ? j : 0][r] = C[(int64_t)gi + (int64_t)gj * ldc];
        }
  }
  // The slabs are processed in rounds of PD (the K range padded to a
  // multiple of PD slabs, the padding staged as zeros): a steady-state loop
  // that fetches PD slabs ahead, then a last round without fetches.  Both
  // are straight-line code with the same slabs in flight on entry, so each
  // stage waits for its own slab's loads only.  Slab s is staged into
  // buffer s % 2 during slab s - 1's MFMAs; the barrier at the top of slab
  // s orders that stage before s's MFMAs and every MFMA of slab s - 1
  // before the stage of slab s + 1 into the same buffer.
  const int nround = (ke - kb + PD * KB - 1) / (PD * KB);
#pragma unroll
  for (int u = 0; u < PD; ++u) fetch(kb + u * KB, ra[u], rb[u]);
  stage(kb, ra[0], rb[0], 0);
  for (int rd = 0; rd + 1 < nround; ++rd) {
#pragma unroll
    for (int u = 0; u < PD; ++u) {
      const int sl = rd * PD + u;
      __syncthreads();
      fetch(kb + (sl + PD) * KB, ra[u], rb[u]);
      stage(kb + (sl + 1) * KB, ra[(u + 1) % PD], rb[(u + 1) % PD], (u + 1) & 1);
      mma(u & 1);
    }
  }
#pragma unroll
  for (int u = 0; u < PD; ++u) {
    const int sl = (nround - 1) * PD + u;
    __syncthreads();
    if (u + 1 < PD) stage(kb + (sl + 1) * KB, ra[u + 1], rb[u + 1], (u + 1) & 1);
    mma(u & 1);
  }
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
        const double cv = CPRE ? cpre[CPRE ? i : 0][CPRE ? j : 0][r] : (beta == 0.0 ? 0.0 : *c);
        *c = (beta == 0.0) ? v : beta * cv + v;
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

// Prefetch depth: four slabs in flight (each stage then has three slabs'
// MFMAs to land), two for K ranges of at most two slabs and for the
// 128 x 128 tiles.
template <bool TA, bool TB, int BM, int BN, int WR>
static hipError_t launch_tile(int m, int n, int k, int nsplit, int kchunk, double alpha, const double* A, int64_t lda,
                              const double* B, int64_t ldb, double beta, double* C, int64_t ldc, int64_t split_stride,
                              hipStream_t s) {
  dim3 grid((m + BM - 1) / BM, (n + BN - 1) / BN, nsplit);
  if (BM * BN >= 128 * 128 || std::min(k, kchunk) <= 2 * slab_depth<BM, BN>())
    hipLaunchKernelGGL((dgemm_kernel<TA, TB, BM, BN, WR, 2>), grid, dim3(256), 0, s, m, n, k, kchunk, alpha, A, lda,
                       B, ldb, beta, C, ldc, split_stride);
  else
    hipLaunchKernelGGL((dgemm_kernel<TA, TB, BM, BN, WR, 4>), grid, dim3(256), 0, s, m, n, k, kchunk, alpha, A, lda,
                       B, ldb, beta, C, ldc, split_stride);
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

// Tiles of a product split over K: small along the long side, so that a
// launch has many workgroups without many K slices.
static void split_tile(int m, int n, int& bm, int& bn) {
  bm = m <= 32 ? 32 : 64;
  bn = (m > 64 && n <= 32) ? 32 : 64;
}

// The tile shape for an m x n product: 32 x 128 for m <= 32, 128 x 32 for
// n <= 32, 64 x 128 for m <= 64, 128 x 64 for n <= 64, else 128 x 128 (or
// 64 x 64 for short K); split over K, those of split_tile.
static hipError_t launch_any(bool ta, bool tb, int m, int n, int k, int nsplit, int kchunk, double alpha,
                             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                             int64_t ldc, int64_t split_stride, hipStream_t s) {
  if (nsplit > 1) {
    // 32 x 64 and 64 x 64 with the wavefronts along n for a short m;
    // 64 x 32 and 64 x 64 along m for a short n
    if (m <= 32)
      return launch_shape<32, 64, 1>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                     split_stride, s);
    if (m <= 64)
      return launch_shape<64, 64, 1>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                     split_stride, s);
    if (n <= 32)
      return launch_shape<64, 32, 4>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                     split_stride, s);
    return launch_shape<64, 64, 4>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                   split_stride, s);
  }
  if (m <= 32)
    return launch_shape<32, 128, 1>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                    split_stride, s);
  if (n <= 32)
    return launch_shape<128, 32, 4>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                    split_stride, s);
  if (m <= 64)
    return launch_shape<64, 128, 1>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
                                    split_stride, s);
  if (n <= 64)
    return launch_shape<128, 64, 4>(ta, tb, m, n, k, nsplit, kchunk, alpha, A, lda, B, ldb, beta, C, ldc,
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

// Split over K for products with a short side (m or n <= 64) and a long K:
// enough K slices that the launch has about 512 workgroups, but no more
// than keeps the partial products' traffic (written once, read once) at
// half the long operand's, k / (4 * short side), and each slice at least 128
// deep.
constexpr int kMaxSplit = 64;
static int splitk_parts(int m, int n, int k) {
  const int sh = std::min(m, n);
  if (sh > 64) return 1;
  int bm, bn;
  split_tile(m, n, bm, bn);
  const int tiles = ((m + bm - 1) / bm) * ((n + bn - 1) / bn);
  const int want = (512 + tiles - 1) / tiles;
  return std::max(1, std::min({want, kMaxSplit, k / (4 * sh), k / 128}));
}

size_t dgemm_splitk_doubles(int m, int n, int k) {
  (void)k;
  return (size_t)kMaxSplit * m * n;
}

hipError_t launch_dgemm_splitk(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int64_t lda,
                               const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* ws,
                               hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  const int nsplit = splitk_parts(m, n, k);
  if (nsplit == 1) return launch_dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
  // K chunks are whole rounds of slabs (128); every launched split gets one
  const int kchunk = ((k + nsplit - 1) / nsplit + 127) / 128 * 128;
  const int used = (k + kchunk - 1) / kchunk;
  hipError_t e = launch_any(ta, tb, m, n, k, used, kchunk, alpha, A, lda, B, ldb, 0.0, ws, m, (int64_t)m * n, s);
  if (e != hipSuccess) return e;
  const int64_t tot = (int64_t)m * n;
  hipLaunchKernelGGL(dgemm_reduce_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, m, n, used, ws, beta,
                     C, ldc);
  return hipGetLastError();
}

}  // namespace gs
