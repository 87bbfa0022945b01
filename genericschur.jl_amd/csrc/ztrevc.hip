// Eigenvectors of a complex Schur factorization (T upper triangular, Z) on
// gfx950, following GenericSchur.jl:
//   right   vectors.jl:45-131   with _usolve!  util.jl:128-293 (zlatrs)
//   left    vectors.jl:372-460  with _cusolve! util.jl:301-461
//   unit 2-norm, largest entry real: _enormalize! util.jl:572-592
//
// One workgroup per matrix, NT/64 wavefronts.  T is held in LDS in compact
// upper-triangular storage.  Each wavefront solves whole eigenvectors: the
// right-hand side / solution lives in registers across its 64 lanes
// (WaveVecZ), the substitution walks T's columns (right) or rows (left) in
// LDS, and the scalar control of the scaled solver runs uniformly in every
// lane.  The solutions are parked in the output V, then each column is
// back-transformed by Z (V[:, k] = Z x) and normalised.
#include <type_traits>

#include "dev_common.h"
#include "internal.h"
#include "trevc_big.h"
#include "wavevec.h"

namespace gs {


// compact upper triangle: column j (0-based) holds rows 0..j at j(j+1)/2
GS_DEV int tri(int i, int j) { return (j * (j + 1)) / 2 + i; }

// Entry (i, j), i <= j, of T: from the compact copy in LDS (orders whose
// triangle fits a CU's LDS) or from T itself in global memory (larger
// orders; the reads stay in L2 across the substitutions).
struct TriLdsZ {
  const zd* p;
  GS_DEV zd operator()(int i, int j) const { return p[tri(i, j)]; }
};
struct GlobZ {
  const zd* p;
  int64_t ld;
  GS_DEV zd operator()(int i, int j) const { return p[i + (int64_t)j * ld]; }
};

// ------------------------------------------------------------------------
// Shifted, clamped diagonal of the solve: T(j,j) - lam, replaced by smin
// when its abs1 is below smin (vectors.jl:87-90 / :412-415).
template <class TA> GS_DEV zd shifted_diag(const TA& Tc, int j, zd lam, double smin) {
  zd d = Tc(j, j) - lam;
  if (zabs1(d) < smin) d = zr(smin);
  return d;
}

// _usolve! on entries 0..m-1 of x (A = T shifted by lam, leading m x m).
template <int R, class TA, class VEC> GS_DEV double usolve_w(const TA& Tc, const double* cnorm, int m, zd lam,
                                                             double smin, VEC& x) {
  const int lane = threadIdx.x & 63;
  const double half = 0.5;
  const double smallnum = GS_FMIN / GS_EPS;
  const double bignum = 1.0 / smallnum;
  double xscale = 1.0;
  auto inrange = [&](int q) { return lane + 64 * q < m; };
  auto cabs1half = [](zd z) { return fabs(z.x / 2) + fabs(z.y / 2); };
  double tmax = 0.0;
  for (int j = lane; j < m; j += 64) tmax = fmax(tmax, fabs(cnorm[j]));
  tmax = wave_max(tmax);
  double tscale;
  if (tmax <= bignum * half) tscale = 1.0;
  else tscale = half / (smallnum * tmax);
  double xmax = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (inrange(q)) xmax = fmax(xmax, cabs1half(x.v[q]));
  xmax = wave_max(xmax);
  double xbound = xmax;
  double grow;
  if (tscale != 1.0) {
    grow = 0.0;
  } else {
    grow = half / fmax(xbound, smallnum);
    xbound = grow;
    bool toosmall = false;
    for (int j = m - 1; j >= 0; --j) {
      toosmall = (grow <= smallnum);
      if (toosmall) break;
      const double tjj = zabs1(shifted_diag(Tc, j, lam, smin));
      if (tjj >= smallnum) xbound = fmin(xbound, fmin(1.0, tjj) * grow);
      else xbound = 0.0;
      const double cn = cnorm[j] * tscale;
      if (tjj + cn >= smallnum) grow *= (tjj / (tjj + cn));
      else grow = 0.0;
    }
    if (!toosmall) grow = xbound;
  }
  if (grow * tscale > smallnum) {
    for (int j = m - 1; j >= 0; --j) {
      const zd xj = x.get(j) / shifted_diag(Tc, j, lam, smin);
      x.set(j, xj);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = lane + 64 * q;
        if (i < j) x.v[q] = x.v[q] - Tc(i, j) * xj;
      }
    }
    return xscale;
  }
  // careful scaled solve
  if (xmax > bignum * half) {
    xscale = (bignum * half) / xmax;
    x.scale(xscale);
    xmax = bignum;
  } else {
    xmax *= 2.0;
  }
  for (int j = m - 1; j >= 0; --j) {
    zd xjv = x.get(j);
    double xj = zabs1(xjv);
    const zd tjjs = shifted_diag(Tc, j, lam, smin) * tscale;
    const double tjj = zabs1(tjjs);
    const double cnj = cnorm[j] * tscale;
    if (tjj > smallnum) {
      if (tjj < 1.0 && xj > tjj * bignum) {
        const double rec = 1.0 / xj;
        x.scale(rec);
        xscale *= rec;
        xmax *= rec;
        xjv = x.get(j);
      }
      xjv = xjv / tjjs;
      x.set(j, xjv);
      xj = zabs1(xjv);
    } else if (tjj > 0.0) {
      if (xj > tjj * bignum) {
        double rec = (tjj * bignum) / xj;
        if (cnj > 1.0) rec /= cnj;
        x.scale(rec);
        xscale *= rec;
        xmax *= rec;
        xjv = x.get(j);
      }
      xjv = xjv / tjjs;
      x.set(j, xjv);
      xj = zabs1(xjv);
    } else {
#pragma unroll
      for (int q = 0; q < R; ++q) x.v[q] = zr(0.0);
      x.set(j, zr(1.0));
      xjv = zr(1.0);
      xscale = 0.0;
      xmax = 0.0;
    }
    if (xj > 1.0) {
      double rec = 1.0 / xj;
      if (cnj > (bignum - xmax) * rec) {
        rec *= half;
        x.scale(rec);
        xscale *= rec;
      }
    } else if (xj * cnj > bignum - xmax) {
      x.scale(half);
      xscale *= half;
    }
    if (j > 0) {
      const zd xjt = x.get(j) * tscale;
      double xm = 0.0;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = lane + 64 * q;
        if (i < j) {
          x.v[q] = x.v[q] - xjt * Tc(i, j);
          xm = fmax(xm, zabs1(x.v[q]));
        }
      }
      xmax = wave_max(xm);
    }
  }
  xscale /= tscale;
  return xscale;
}

// _cusolve! on entries o..o+m-1 of x: (A)^H y = s b with
// A(i, j) = T(o+i, o+j) for i < j and the shifted, clamped diagonal; cnorm
// is indexed from 0 (the reference passes the full tnorms, vectors.jl:417-420).
template <int R, class TA, class VEC> GS_DEV double cusolve_w(const TA& Tc, const double* cnorm, int o, int m, zd lam,
                                                              double smin, VEC& x) {
  const int lane = threadIdx.x & 63;
  const double half = 0.5;
  const double smallnum = GS_FMIN / GS_EPS;
  const double bignum = 1.0 / smallnum;
  double xscale = 1.0;
  auto inrange = [&](int q) {
    const int i = lane + 64 * q;
    return i >= o && i < o + m;
  };
  auto cabs1half = [](zd z) { return fabs(z.x / 2) + fabs(z.y / 2); };
  double tmax = 0.0;
  for (int j = lane; j < m; j += 64) tmax = fmax(tmax, fabs(cnorm[j]));
  tmax = wave_max(tmax);
  double tscale;
  if (tmax <= bignum * half) tscale = 1.0;
  else tscale = half / (smallnum * tmax);
  double xmax = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (inrange(q)) xmax = fmax(xmax, cabs1half(x.v[q]));
  xmax = wave_max(xmax);
  double xbound = xmax;
  double grow;
  if (tscale != 1.0) {
    grow = 0.0;
  } else {
    grow = half / fmax(xbound, smallnum);
    xbound = grow;
    bool toosmall = false;
    for (int j = 0; j < m; ++j) {
      toosmall = (grow <= smallnum);
      if (toosmall) break;
      const double xj = 1.0 + cnorm[j] * tscale;
      grow = fmin(grow, xbound / xj);
      const double tjj = zabs1(shifted_diag(Tc, o + j, lam, smin));
      if (tjj >= smallnum) {
        if (xj > tjj) xbound *= (tjj / xj);
      } else {
        xbound = 0.0;
      }
    }
    if (!toosmall) grow = fmin(grow, xbound);
  }
  if (grow * tscale > smallnum) {
    // forward substitution in axpy form: after x[j] is final, remove its
    // contribution from the later right-hand sides
    for (int j = 0; j < m; ++j) {
      const zd xj = x.get(o + j) / zconj(shifted_diag(Tc, o + j, lam, smin));
      x.set(o + j, xj);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = lane + 64 * q;  // absolute entry index
        if (i > o + j && i < o + m) x.v[q] = x.v[q] - zconj(Tc(o + j, i)) * xj;
      }
    }
    return xscale;
  }
  if (xmax > bignum * half) {
    xscale = (bignum * half) / xmax;
    x.scale(xscale);
    xmax = bignum;
  } else {
    xmax *= 2.0;
  }
  for (int j = 0; j < m; ++j) {
    double xj = zabs1(x.get(o + j));
    zd uscale = zr(tscale);
    double rec = 1.0 / fmax(xmax, 1.0);
    zd tjjs = zconj(shifted_diag(Tc, o + j, lam, smin)) * tscale;
    const double cnj = cnorm[j] * tscale;
    if (cnj > (bignum - xj) * rec) {
      rec *= half;
      const double tjj = zabs1(tjjs);
      if (tjj > 1.0) {
        rec = fmin(1.0, rec * tjj);
        uscale = uscale / tjjs;
      }
      if (rec < 1.0) {
        x.scale(rec);
        xscale *= rec;
        xmax *= rec;
      }
    }
    // csumj = sum_{i<j} (conj(A(i,j)) * uscale) * x[i]
    zd part = zr(0.0);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      if (i >= o && i < o + j) part = part + (zconj(Tc(i, o + j)) * uscale) * x.v[q];
    }
    const zd csumj = wave_zsum(part);
    const bool unscaled = (uscale.x == tscale && uscale.y == 0.0) || isnan(uscale.x) || isnan(uscale.y);
    if (unscaled) {
      zd xjv = x.get(o + j) - csumj;
      xj = zabs1(xjv);
      tjjs = zconj(shifted_diag(Tc, o + j, lam, smin)) * tscale;
      const double tjj = zabs1(tjjs);
      if (tjj > smallnum) {
        if (tjj < 1.0 && xj > tjj * bignum) {
          const double rec2 = 1.0 / xj;
          x.scale(rec2);
          xscale *= rec2;
          xmax *= rec2;
          xjv = xjv * rec2;
        }
        xjv = xjv / tjjs;
        x.set(o + j, xjv);
      } else if (tjj > 0.0) {
        if (xj > tjj * bignum) {
          const double rec2 = (tjj * bignum) / xj;
          x.scale(rec2);
          xscale *= rec2;
          xmax *= rec2;
          xjv = xjv * rec2;
        }
        xjv = xjv / tjjs;
        x.set(o + j, xjv);
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (inrange(q)) x.v[q] = zr(0.0);
        x.set(o + j, zr(1.0));
        xscale = 0.0;
        xmax = 0.0;
      }
    } else {
      x.set(o + j, x.get(o + j) / tjjs - csumj);
    }
    xmax = fmax(xmax, zabs1(x.get(o + j)));
  }
  xscale /= tscale;
  return xscale;
}

// max abs1 scaling then _enormalize! of one column held in a WaveVecZ
// (entries 0..n-1); normalize = 0 stops after the max-abs1 scaling.
template <int R, class VEC> GS_DEV void normalize_col(VEC& v, int n, bool enorm) {
  const int lane = threadIdx.x & 63;
  double t0 = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (lane + 64 * q < n) t0 = fmax(t0, zabs1(v.v[q]));
  t0 = wave_max(t0);
  const double remax = 1.0 / t0;
#pragma unroll
  for (int q = 0; q < R; ++q) v.v[q] = v.v[q] * remax;
  if (!enorm) return;
  // 2-norm without over/underflow, and the first index of the largest |v|^2
  double am = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (lane + 64 * q < n) am = fmax(am, fmax(fabs(v.v[q].x), fabs(v.v[q].y)));
  am = wave_max(am);
  double ss = 0.0;
  double best = -1.0;
  int bi = 0x7fffffff;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = lane + 64 * q;
    if (i < n) {
      const double a = v.v[q].x / am, b = v.v[q].y / am;
      ss += a * a + b * b;
      const double u = zabs2(v.v[q]);
      if (u > best) {
        best = u;
        bi = i;
      }
    }
  }
  ss = wave_sum(ss);
  const double nrm = am * sqrt(ss);
  // largest |v|^2, first index among ties (util.jl:578-584 uses strict >)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  const zd vi0 = v.get(bi);
  const double s = 1.0 / nrm;
  const zd tt = (s * zconj(vi0)) / sqrt(zabs2(vi0));
#pragma unroll
  for (int q = 0; q < R; ++q) v.v[q] = v.v[q] * tt;
  const zd r0 = v.get(bi);
  v.set(bi, zr(r0.x));
}

template <int NT, int R, bool GLOB> __global__ void __launch_bounds__(NT) ztrevc_kernel(TrevcArgs a) {
  extern __shared__ double smem[];
  const int b = blockIdx.x;
  if (b >= a.count) return;
  const int m = a.order ? a.order[b] : b;
  const int n = a.n[m];
  if (n <= 0) return;
  const int64_t o = a.off[m];
  const int64_t ldt = a.ldt > 0 ? a.ldt : n;
  const int64_t ldz = a.ldz > 0 ? a.ldz : n;
  const int64_t ldv = a.ldv > 0 ? a.ldv : n;
  const zd* Tg = reinterpret_cast<const zd*>(a.T) + o;
  const zd* Zg = a.Z ? reinterpret_cast<const zd*>(a.Z) + o : nullptr;
  zd* Vg = reinterpret_cast<zd*>(a.V) + o;
  zd* Tl = reinterpret_cast<zd*>(smem);
  double* tnorms = GLOB ? smem : reinterpret_cast<double*>(Tl + (n * (n + 1)) / 2);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  constexpr int NW = NT / 64;
  const bool left = a.left != 0;
  if (!GLOB) {
    for (int j = 0; j < n; ++j)
      for (int i = tid; i <= j; i += NT) Tl[tri(i, j)] = Tg[i + j * ldt];
    __syncthreads();
  }
  typedef typename std::conditional<GLOB, GlobZ, TriLdsZ>::type TA;
  TA Tc;
  if constexpr (GLOB) Tc = GlobZ{Tg, ldt};
  else Tc = TriLdsZ{Tl};
  // strictly-upper column norms: abs for right (vectors.jl:70-75), abs1 for
  // left (vectors.jl:395-400)
  for (int j = tid; j < n; j += NT) {
    double s = 0.0;
    for (int i = 0; i < j; ++i) s += left ? zabs1(Tc(i, j)) : zabs(Tc(i, j));
    tnorms[j] = s;
  }
  __syncthreads();
  const double ulp = GS_EPS;
  const double smallnum = GS_FMIN * (double)n;  // safemin * n (vectors.jl:61,386)
  // solves, one eigenvector per wavefront at a time; the solution (with the
  // scale at the diagonal slot) is parked in V's column
  for (int kk = (left ? w : n - 1 - w); left ? (kk < n) : (kk >= 0); kk += (left ? NW : -NW)) {
    const zd lam = Tc(kk, kk);
    const double smin = fmax(ulp * zabs1(lam), smallnum);
    WaveVecZ<R> x;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      zd v = zr(0.0);
      if (!left) {
        if (i < kk) v = -Tc(i, kk);
      } else {
        if (i > kk && i < n) v = -zconj(Tc(kk, i));
      }
      x.v[q] = v;
    }
    double vscale;
    if (!left) {
      vscale = (kk > 0) ? usolve_w<R>(Tc, tnorms, kk, lam, smin, x) : 1.0;
    } else {
      vscale = (kk < n - 1) ? cusolve_w<R>(Tc, tnorms, kk + 1, n - 1 - kk, lam, smin, x) : 1.0;
    }
    x.set(kk, zr(vscale));
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      if (i < n) Vg[i + kk * ldv] = (left ? (i >= kk) : (i <= kk)) ? x.v[q] : zr(0.0);
    }
  }
  __syncthreads();
  // back-transform by Z and normalise, one column per wavefront
  for (int kk = w; kk < n; kk += NW) {
    WaveVecZ<R> x;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      x.v[q] = (i < n) ? Vg[i + kk * ldv] : zr(0.0);
    }
    WaveVecZ<R> y;
    if (Zg) {
      const zd vs = x.get(kk);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = lane + 64 * q;
        y.v[q] = (r < n) ? vs * Zg[r + kk * ldz] : zr(0.0);
      }
      const int i0 = left ? kk + 1 : 0, i1 = left ? n : kk;
      for (int i = i0; i < i1; ++i) {
        const zd xi = x.get(i);
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int r = lane + 64 * q;
          if (r < n) y.v[q] = y.v[q] + Zg[r + (int64_t)i * ldz] * xi;
        }
      }
    } else {
      y = x;
    }
    normalize_col<R>(y, n, a.normalize != 0);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = lane + 64 * q;
      if (r < n) Vg[r + kk * ldv] = y.v[q];
    }
  }
}

size_t ztrevc_lds_bytes(int n) { return 16 * ((size_t)n * (n + 1) / 2) + 8 * (size_t)n + 64; }

template <int NT, int R, bool GLOB> static void launch_ztrevc_t(const TrevcArgs& a, size_t lds, hipStream_t s) {
  (void)hipFuncSetAttribute((const void*)ztrevc_kernel<NT, R, GLOB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
  hipLaunchKernelGGL((ztrevc_kernel<NT, R, GLOB>), dim3(a.count), dim3(NT), lds, s, a);
}

// Orders up to 128 keep T's triangle in LDS; above, T is read from global
// memory (only the column norms in LDS), up to ztrevc_max_n().
hipError_t launch_ztrevc(const TrevcArgs& a, int nmax, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  if (nmax <= 64) launch_ztrevc_t<256, 1, false>(a, ztrevc_lds_bytes(nmax), s);
  else if (nmax <= 128) launch_ztrevc_t<256, 2, false>(a, ztrevc_lds_bytes(nmax), s);
  else if (nmax <= 256) launch_ztrevc_t<512, 4, true>(a, 8 * (size_t)nmax + 64, s);
  else if (nmax <= 512) launch_ztrevc_t<256, 8, true>(a, 8 * (size_t)nmax + 64, s);
  else if (nmax <= 1024) launch_ztrevc_t<256, 16, true>(a, 8 * (size_t)nmax + 64, s);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}
int ztrevc_max_n() { return 1024; }

// ------------------------------------------------------- larger orders
// One matrix of order 1024 < n <= trevc_big_max_n() over the whole GPU:
// column norms in global memory, one solve per workgroup (a single
// wavefront, solution vector in LDS), solutions in planar X = [Re | Im]
// (n x 2n), the back-transform Z X as three FP64 MFMA GEMMs on Z's planar
// copy, then one wavefront per column normalises and writes V.  The
// per-column steps are those of ztrevc_kernel.
struct ZBig {
  int n, left, normalize;
  const zd* T;
  int64_t ldt;
  const zd* Tt;  // T's transpose (left solves), ld n
  const zd* Z;   // may be null
  int64_t ldz;
  zd* V;
  int64_t ldv;
  double* tnorms;
  double* Zp;  // [Re Z | Im Z], n x 2n, ld n
  double* X;   // n x 2n, ld n
  double* Y;   // n x 2n, ld n
};

// workgroup t solves column n-1-t (right) or t (left), longest first
template <int R, class TA> __global__ void __launch_bounds__(64) ztrevc_big_solve(ZBig a, TA Tc) {
  extern __shared__ double smem[];
  const int n = a.n, lane = threadIdx.x & 63, t = blockIdx.x;
  const bool left = a.left != 0;
  const int kk = left ? t : n - 1 - t;
  const zd lam = Tc(kk, kk);
  const double smin = fmax(GS_EPS * zabs1(lam), GS_FMIN * (double)n);  // vectors.jl:61,386
  WaveVecL<R> x{{reinterpret_cast<zd*>(smem)}};
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = lane + 64 * q;
    zd v = zr(0.0);
    if (!left) {
      if (i < kk) v = -Tc(i, kk);
    } else {
      if (i > kk && i < n) v = -zconj(Tc(kk, i));
    }
    x.v[q] = v;
  }
  double vscale;
  if (!left) vscale = (kk > 0) ? usolve_w<R>(Tc, a.tnorms, kk, lam, smin, x) : 1.0;
  else vscale = (kk < n - 1) ? cusolve_w<R>(Tc, a.tnorms, kk + 1, n - 1 - kk, lam, smin, x) : 1.0;
  x.set(kk, zr(vscale));
  double* xr = a.X + (int64_t)kk * n;
  double* xi = xr + (int64_t)n * n;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = lane + 64 * q;
    if (i < n) {
      const zd z = (left ? (i >= kk) : (i <= kk)) ? x.v[q] : zr(0.0);
      xr[i] = z.x;
      xi[i] = z.y;
    }
  }
}

// Zp = [Re Z | Im Z]
__global__ void __launch_bounds__(256) ztrevc_big_planar(int n, const zd* Z, int64_t ldz, double* Zp) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nn; e += (int64_t)gridDim.x * 256) {
    const int64_t c = e / n, r = e - c * n;
    const zd z = Z[r + c * ldz];
    Zp[e] = z.x;
    Zp[e + nn] = z.y;
  }
}

template <int R> __global__ void __launch_bounds__(64) ztrevc_big_finish(ZBig a) {
  extern __shared__ double smem[];
  const int n = a.n, lane = threadIdx.x & 63, kk = blockIdx.x;
  WaveVecL<R> y{{reinterpret_cast<zd*>(smem)}};
  const double* src = a.Z ? a.Y : a.X;
  const double* yr = src + (int64_t)kk * n;
  const double* yi = yr + (int64_t)n * n;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = lane + 64 * q;
    y.v[q] = (r < n) ? zmk(yr[r], yi[r]) : zr(0.0);
  }
  normalize_col<R>(y, n, a.normalize != 0);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = lane + 64 * q;
    if (r < n) a.V[r + (int64_t)kk * a.ldv] = y.v[q];
  }
}

size_t ztrevc_big_ws_bytes(int n) {
  const size_t nn = (size_t)n * n;
  return trevc_seg(8 * (size_t)n) + 4 * trevc_seg(16 * nn);
}

template <int R> static hipError_t launch_zbig_r(const ZBig& b, hipStream_t s) {
  const int n = b.n;
  const size_t lds = 64 * R * sizeof(zd);
  if (b.left) {
    auto k = ztrevc_big_solve<R, TransG<zd>>;
    (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k, dim3(n), dim3(64), lds, s, b, TransG<zd>{b.Tt, n});
  } else {
    auto k = ztrevc_big_solve<R, GlobZ>;
    (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k, dim3(n), dim3(64), lds, s, b, GlobZ{b.T, b.ldt});
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (b.Z) {
    // [Yr | Yi] = Re Z [Xr | Xi]; Yr -= Im Z Xi; Yi += Im Z Xr
    const int64_t nn = (int64_t)n * n;
    const double *Zr = b.Zp, *Zi = b.Zp + nn, *Xr = b.X, *Xi = b.X + nn;
    double *Yr = b.Y, *Yi = b.Y + nn;
    if ((e = launch_dgemm(false, false, n, 2 * n, n, 1.0, Zr, n, Xr, n, 0.0, Yr, n, s)) != hipSuccess) return e;
    if ((e = launch_dgemm(false, false, n, n, n, -1.0, Zi, n, Xi, n, 1.0, Yr, n, s)) != hipSuccess) return e;
    if ((e = launch_dgemm(false, false, n, n, n, 1.0, Zi, n, Xr, n, 1.0, Yi, n, s)) != hipSuccess) return e;
  }
  (void)hipFuncSetAttribute((const void*)ztrevc_big_finish<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(ztrevc_big_finish<R>, dim3(n), dim3(64), lds, s, b);
  return hipGetLastError();
}

// a.T, a.Z, a.V point at the matrix (complex interleaved), ld from
// a.ldt/ldz/ldv; ws holds ztrevc_big_ws_bytes(n)
hipError_t launch_ztrevc_big(const TrevcArgs& a, int n, void* ws, hipStream_t s) {
  if (n <= ztrevc_max_n() || n > trevc_big_max_n() || !ws) return hipErrorInvalidValue;
  const size_t nn = (size_t)n * n;
  char* p = static_cast<char*>(ws);
  ZBig b;
  b.n = n;
  b.left = a.left;
  b.normalize = a.normalize;
  b.T = reinterpret_cast<const zd*>(a.T);
  b.ldt = a.ldt > 0 ? a.ldt : n;
  b.Z = reinterpret_cast<const zd*>(a.Z);
  b.ldz = a.ldz > 0 ? a.ldz : n;
  b.V = reinterpret_cast<zd*>(a.V);
  b.ldv = a.ldv > 0 ? a.ldv : n;
  b.tnorms = reinterpret_cast<double*>(p);
  p += trevc_seg(8 * (size_t)n);
  zd* tt = reinterpret_cast<zd*>(p);
  b.Tt = tt;
  p += trevc_seg(16 * nn);
  b.Zp = reinterpret_cast<double*>(p);
  p += trevc_seg(16 * nn);
  b.X = reinterpret_cast<double*>(p);
  p += trevc_seg(16 * nn);
  b.Y = reinterpret_cast<double*>(p);
  // strictly-upper column norms: abs for right (vectors.jl:70-75), abs1 for
  // left (vectors.jl:395-400)
  hipLaunchKernelGGL(trevc_colnorm_kernel<zd>, dim3((n + 3) / 4), dim3(256), 0, s, n, b.T, b.ldt, b.left != 0,
                     b.tnorms);
  if (b.left)
    hipLaunchKernelGGL(trevc_transpose_kernel<zd>, dim3((n + 31) / 32, (n + 31) / 32), dim3(256), 0, s, n, b.T,
                       b.ldt, tt);
  if (b.Z) hipLaunchKernelGGL(ztrevc_big_planar, dim3(2048), dim3(256), 0, s, n, b.Z, b.ldz, b.Zp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return n <= 2048 ? launch_zbig_r<32>(b, s) : launch_zbig_r<64>(b, s);
}

}  // namespace gs
