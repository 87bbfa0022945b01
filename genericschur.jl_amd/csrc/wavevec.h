// A vector of up to 64*R entries distributed over one wavefront: lane l
// holds entries l, l+64, ... (0-based) in registers.  Entry access by a
// uniform index broadcasts through a cross-lane shuffle, so back- and
// forward-substitutions keep the solution vector out of LDS.
#pragma once
#include "dev_common.h"

template <int R> struct WaveVecZ {
  zd v[R];
  GS_DEV zd& slot(int q) { return v[q]; }
  // broadcast entry i (0-based, uniform)
  GS_DEV zd get(int i) const {
    const int q = i >> 6, l = i & 63;
    zd s = v[0];
#pragma unroll
    for (int k = 1; k < R; ++k)
      if (q == k) s = v[k];
    zd r;
    r.x = __shfl(s.x, l, 64);
    r.y = __shfl(s.y, l, 64);
    return r;
  }
  // set entry i (0-based, uniform) to a uniform value
  GS_DEV void set(int i, zd val) {
    const int q = i >> 6, l = i & 63;
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int k = 0; k < R; ++k)
      if (q == k && lane == l) v[k] = val;
  }
  GS_DEV void scale(double s) {
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = v[k] * s;
  }
};

// The same vector kept in LDS (64 R entries at p, entry i at p[i]), for
// orders whose vectors would not fit the registers.  v[q] is lane l's
// entry l + 64 q as in WaveVecZ; get() reads the entry every lane sees
// (the wavefront's LDS accesses complete in order, so a lane reads what
// another lane of the wavefront stored before).
struct LdsLaneRef {
  zd* p;
  GS_DEV zd& operator[](int q) const { return p[(threadIdx.x & 63) + 64 * q]; }
};
template <int R> struct WaveVecL {
  LdsLaneRef v;
  GS_DEV zd get(int i) const { return v.p[i]; }
  GS_DEV void set(int i, zd val) {
    if ((threadIdx.x & 63) == 0) v.p[i] = val;
  }
  GS_DEV void scale(double s) {
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = v[k] * s;
  }
};

// wave-wide maximum over lanes (all lanes receive it)
GS_DEV double wmax(double x) { return wave_max(x); }
