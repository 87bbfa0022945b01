// Shared pieces of the eigenvector path for orders above the
// one-workgroup kernels (vectors.jl:12-20 takes any order): T's transpose,
// which turns the row walks of the left solves into coalesced column
// reads, the strictly-upper column norms (one wavefront per column), and
// the accessor that reads T through the transpose.
#pragma once
#include "dev_common.h"

namespace gs {

// Tt(j, i) = T(i, j) for 0 <= i, j < n (Tt with ld n); 32 x 32 tiles
// through LDS, grid ((n + 31) / 32, (n + 31) / 32), 256 threads.
template <class E> __global__ void __launch_bounds__(256) trevc_transpose_kernel(int n, const E* T, int64_t ldt, E* Tt) {
  __shared__ E tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  for (int k = ty; k < 32; k += 8) {
    const int r = bx + tx, c = by + k;
    if (r < n && c < n) tile[k][tx] = T[r + (int64_t)c * ldt];
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    const int r = by + tx, c = bx + k;  // Tt(r, c) = T(c, r) = tile[tx][k]
    if (r < n && c < n) Tt[r + (int64_t)c * n] = tile[tx][k];
  }
}

GS_DEV double tmag(double x, bool) { return fabs(x); }
GS_DEV double tmag(zd z, bool abs1) { return abs1 ? zabs1(z) : zabs(z); }

// cn[j] = sum_{i < j} mag(T(i, j)), mag = |.| or abs1; one wavefront per
// column, grid (n + 3) / 4 of 256 threads.
template <class E> __global__ void __launch_bounds__(256) trevc_colnorm_kernel(int n, const E* T, int64_t ldt, bool abs1,
                                                                              double* cn) {
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= n) return;
  double s = 0.0;
  for (int i = lane; i < j; i += 64) s += tmag(T[i + (int64_t)j * ldt], abs1);
  s = wave_sum(s);
  if (lane == 0) cn[j] = s;
}

// T(i, j) read from the transpose Tt (ld n): the left solves' row walks
// T(j, i), i along the lanes, become column reads of Tt.
template <class E> struct TransG {
  const E* p;
  int64_t ld;
  GS_DEV E operator()(int i, int j) const { return p[j + (int64_t)i * ld]; }
};

// Workspace segments of the large-order path, 256-byte aligned.
inline size_t trevc_seg(size_t bytes) { return (bytes + 255) & ~size_t(255); }

}  // namespace gs
