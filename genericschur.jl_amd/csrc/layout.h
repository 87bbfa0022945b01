// Storage layouts of the working Hessenberg matrix H used by the QR kernels.
// Both map 1-based (i, j) to an element offset.
//
//   FullLayout  column-major with leading dimension ld.
//   BandLayout  only the entries with j >= i - B: column j keeps rows
//               1 .. min(j + B, n), columns back to back.  B = 2 holds the
//               complex single-shift bulge (which lives on the second
//               subdiagonal), B = 3 the real double-shift bulge.  It halves
//               the LDS footprint of a resident matrix, which doubles the
//               number of resident matrices per CU.
#pragma once
#include <stdint.h>

struct FullLayout {
  int64_t ld;
  __device__ __forceinline__ int64_t idx(int i, int j) const { return (int64_t)(i - 1) + (int64_t)(j - 1) * ld; }
};

template <int B> struct BandLayout {
  int n;
  int nb;       // n - B: the last column index whose length is j + B
  int64_t cut;  // offset of column nb + 1
  __device__ __forceinline__ void init(int n_) {
    n = n_;
    nb = n_ - B;
    if (nb < 0) nb = 0;
    cut = (int64_t)nb * (nb + 1 + 2 * B) / 2;
  }
  // 32-bit arithmetic: band matrices live in LDS, far below 2^31 elements
  __device__ __forceinline__ int col(int j) const {
    const int jm = j - 1;
    if (jm <= nb) return (jm * (jm + 1 + 2 * B)) >> 1;
    return (int)cut + (jm - nb) * n;
  }
  __device__ __forceinline__ int idx(int i, int j) const { return col(j) + (i - 1); }
  // number of stored elements for order n
  static __host__ __device__ int64_t size(int n_) {
    int nb_ = n_ - B;
    if (nb_ < 0) nb_ = 0;
    return (int64_t)nb_ * (nb_ + 1 + 2 * B) / 2 + (int64_t)(n_ - nb_) * n_;
  }
};
