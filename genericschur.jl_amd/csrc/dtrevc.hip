// Eigenvectors of a real standard-form Schur factorization (T
// quasi-triangular, Z) on gfx950, following GenericSchur.jl:
//   right   vectors.jl:133-362 (dtrevc-style)
//   left    vectors.jl:462-698
//   1x1 / 2x2 solves  _xsolve util.jl:698-751, _safe_lu_solve! :603-687
//   unit 2-norm, largest entry real: _enormalize! util.jl:572-592
//
// Same organisation as the complex kernel: T (upper triangle plus the
// subdiagonal of the 2x2 blocks) in LDS, one eigenvector (or conjugate
// pair) per wavefront with the complex solution in registers, solutions
// parked in V, then back-transform by Z and normalisation.  The conjugate
// partner of a complex pair is written last as the conjugate of the
// finished column, which equals the reference's copy-then-normalise.
#include <type_traits>

#include "dev_common.h"
#include "internal.h"
#include "trevc_big.h"
#include "wavevec.h"

namespace gs {


GS_DEV int trd(int i, int j) { return (j * (j + 1)) / 2 + i; }

// Entry (i, j), i <= j, of T: from the compact copy in LDS, or from T in
// global memory for orders whose triangle does not fit a CU's LDS.
struct TriLdsD {
  const double* p;
  GS_DEV double operator()(int i, int j) const { return p[trd(i, j)]; }
};
struct GlobD {
  const double* p;
  int64_t ld;
  GS_DEV double operator()(int i, int j) const { return p[i + (int64_t)j * ld]; }
};

// T entry (0-based) from compact upper storage + subdiagonal
template <class TA> GS_DEV double Tat(const TA& Tc, const double* sub, int i, int j) {
  if (i <= j) return Tc(i, j);
  if (i == j + 1) return sub[j];
  return 0.0;
}

// _safe_lu_solve! (util.jl:603-687) for a complex 2x2 system; A column-major
// (a11, a21, a12, a22), b in/out.  Returns the scale.
GS_DEV double safe_lu2(zd* A, zd* b) {
  const double smallnum = GS_FMIN / GS_EPS;
  double smin = 0.0;
  for (int i = 0; i < 4; ++i) smin = fmax(smin, zabs(A[i]));
  smin = fmax(GS_EPS * smin, smallnum);
  // complete pivoting for i = 1: the last maximal entry in (ip, jp) order
  int ipsv = 0, jpsv = 0;
  double xmax = 0.0;
  for (int ip = 0; ip < 2; ++ip)
    for (int jp = 0; jp < 2; ++jp)
      if (zabs(A[ip + 2 * jp]) >= xmax) {
        xmax = zabs(A[ip + 2 * jp]);
        ipsv = ip;
        jpsv = jp;
      }
  if (ipsv != 0) {
    for (int j = 0; j < 2; ++j) {
      const zd t = A[0 + 2 * j];
      A[0 + 2 * j] = A[1 + 2 * j];
      A[1 + 2 * j] = t;
    }
    const zd t = b[0];
    b[0] = b[1];
    b[1] = t;
  }
  if (jpsv != 0) {
    for (int j = 0; j < 2; ++j) {
      const zd t = A[j + 0];
      A[j + 0] = A[j + 2];
      A[j + 2] = t;
    }
  }
  if (zabs(A[0]) < smin) A[0] = zr(smin);
  A[1] = A[1] / A[0];
  b[1] = b[1] - A[1] * b[0];
  A[3] = A[3] - A[1] * A[2];
  if (zabs(A[3]) < smin) A[3] = zr(smin);
  double scale = 1.0;
  const double a8 = 8.0 * smallnum;
  bool needscl = (a8 * zabs(b[0]) > zabs(A[0])) || (a8 * zabs(b[1]) > zabs(A[3]));
  if (needscl) {
    scale = (1.0 / 8.0) / fmax(zabs(b[0]), zabs(b[1]));
    b[0] = b[0] * scale;
    b[1] = b[1] * scale;
  }
  {
    const zd t = zdiv(zr(1.0), A[3]);
    b[1] = b[1] * t;
  }
  {
    const zd t = zdiv(zr(1.0), A[0]);
    b[0] = b[0] * t;
    b[0] = b[0] - (t * A[2]) * b[1];
  }
  if (jpsv != 0) {
    const zd t = b[0];
    b[0] = b[1];
    b[1] = t;
  }
  return scale;
}

// _xsolve (util.jl:698-751): (a A - bw I) x = s B, na = 1 or 2, A column-major.
GS_DEV double xsolve_dev(int na, const double* Ae, zd bw, zd* B, double smin, double& xnorm) {
  double s = 1.0;
  const double small = 2.0 * GS_FMIN;
  const double bignum = s / small;
  smin = fmax(smin, small);
  if (na == 1) {
    double csr = Ae[0] - bw.x;
    double csi = -bw.y;
    double cnorm = fabs(csr) + fabs(csi);
    if (cnorm < smin) {
      csr = smin;
      csi = 0.0;
      cnorm = smin;
    }
    const double bnorm = zabs1(B[0]);
    if (cnorm < 1.0 && bnorm > 1.0 && bnorm > bignum * cnorm) s = 1.0 / bnorm;
    const zd x = (B[0] * s) / zmk(csr, csi);
    xnorm = zabs1(x);
    B[0] = x;
    return s;
  }
  zd M[4];
  double cmax = 0.0;
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < 2; ++i) {
      zd z = zr(Ae[i + 2 * j]);
      if (i == j) z = z - bw;
      M[i + 2 * j] = z;
      cmax = fmax(cmax, zabs1(z));
    }
  if (cmax < smin) {
    const double bnorm = fmax(zabs1(B[0]), zabs1(B[1]));
    if (smin < 1.0 && bnorm > 1.0 && bnorm > bignum * smin) s = 1.0 / bnorm;
    const double t1 = s / smin;
    B[0] = B[0] * t1;
    B[1] = B[1] * t1;
  } else {
    s = safe_lu2(M, B);
  }
  xnorm = sqrt(zabs2(B[0]) + zabs2(B[1]));
  return s;
}

// ------------------------------------------------------------- right
// Solution for column ki (0-based) with ip = 0 (real wr) or -1 (pair,
// lambda = wr + i wi, ki the second row of the block).  Entries 0..ki hold
// the vector on exit.
template <int R, class TA, class VEC> GS_DEV void dsolve_right(const TA& Tc, const double* sub, const double* tnorms, int ki,
                                                     int ip, double smallnum, double bignum, VEC& v) {
  const int lane = threadIdx.x & 63;
  const double ulp = GS_EPS;
  const double wr = Tc(ki, ki);
  const double wi = (ip == 0) ? 0.0 : sqrt(fabs(sub[ki - 1])) * sqrt(fabs(Tc(ki - 1, ki)));
  const double smin = fmax(ulp * (fabs(wr) + fabs(wi)), smallnum);
  const zd lam = zmk(wr, wi);
  int jtop;
  if (ip == 0) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      v.v[q] = (i < ki) ? zr(-Tc(i, ki)) : (i == ki ? zr(1.0) : zr(0.0));
    }
    jtop = ki - 1;
  } else {
    zd a, c;
    if (fabs(Tc(ki - 1, ki)) >= fabs(sub[ki - 1])) {
      a = zr(1.0);
      c = zmk(0.0, wi / Tc(ki - 1, ki));
    } else {
      a = zr(-wi / sub[ki - 1]);
      c = zmk(0.0, 1.0);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      zd val = zr(0.0);
      if (i < ki - 1) val = zmk(-a.x * Tc(i, ki - 1), -(c.y * Tc(i, ki)));
      else if (i == ki - 1) val = a;
      else if (i == ki) val = c;
      v.v[q] = val;
    }
    jtop = ki - 2;
  }
  int j = jtop;
  while (j >= 0) {
    const bool two = (j > 0) && (sub[j - 1] != 0.0);
    if (!two) {
      zd B[1] = {v.get(j)};
      const double Ae[1] = {Tc(j, j)};
      double t1;
      double vscale = xsolve_dev(1, Ae, lam, B, smin, t1);
      if (t1 > 1.0 && tnorms[j] > bignum / t1) {
        B[0] = B[0] / t1;
        vscale /= t1;
      }
      if (vscale != 1.0) {
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (lane + 64 * q <= ki) v.v[q] = v.v[q] * vscale;
      }
      v.set(j, B[0]);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = lane + 64 * q;
        if (i < j) v.v[q] = v.v[q] - B[0] * Tc(i, j);
      }
      j -= 1;
    } else {
      zd B[2] = {v.get(j - 1), v.get(j)};
      const double Ae[4] = {Tc(j - 1, j - 1), sub[j - 1], Tc(j - 1, j), Tc(j, j)};
      double t1;
      double vscale = xsolve_dev(2, Ae, lam, B, smin, t1);
      if (t1 > 1.0) {
        const double beta = fmax(tnorms[j - 1], tnorms[j]);
        if (beta > bignum / t1) {
          B[0] = B[0] / t1;
          B[1] = B[1] / t1;
          vscale /= t1;
        }
      }
      if (vscale != 1.0) {
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (lane + 64 * q <= ki) v.v[q] = v.v[q] * vscale;
      }
      v.set(j - 1, B[0]);
      v.set(j, B[1]);
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = lane + 64 * q;
        if (i < j - 1) v.v[q] = (v.v[q] - B[0] * Tc(i, j - 1) - B[1] * Tc(i, j));
      }
      j -= 2;
    }
  }
}

// -------------------------------------------------------------- left
// Solution for column ki with ip = 0 or 1 (pair, ki the first row of the
// block, lambda = wr - i wi in the solves).  Entries ki..n-1 hold it.
// Forward substitution in axpy form; the reference's overflow guards
// (vmax, vcrit) and the j1 = j + 1 quirk of vectors.jl:627 are kept.
template <int R, class TA, class VEC> GS_DEV void dsolve_left(const TA& Tc, const double* sub, const double* tnorms, int n,
                                                    int ki, int ip, double smallnum, double bignum, VEC& v) {
  const int lane = threadIdx.x & 63;
  const double ulp = GS_EPS;
  const double wr = Tc(ki, ki);
  const double wi = (ip == 0) ? 0.0 : sqrt(fabs(Tc(ki, ki + 1))) * sqrt(fabs(sub[ki]));
  const double smin = fmax(ulp * (fabs(wr) + fabs(wi)), smallnum);
  int jstart;
  zd lam;
  if (ip == 0) {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      v.v[q] = (i > ki && i < n) ? zr(-Tc(ki, i)) : (i == ki ? zr(1.0) : zr(0.0));
    }
    jstart = ki + 1;
    lam = zr(wr);
  } else {
    zd a, c;
    if (fabs(Tc(ki, ki + 1)) >= fabs(sub[ki])) {
      a = zr(wi / Tc(ki, ki + 1));
      c = zmk(0.0, 1.0);
    } else {
      a = zr(1.0);
      c = zmk(0.0, -wi / sub[ki]);
    }
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      zd val = zr(0.0);
      if (i > ki + 1 && i < n) val = zmk(-a.x * Tc(ki, i), -(c.y * Tc(ki + 1, i)));
      else if (i == ki) val = a;
      else if (i == ki + 1) val = c;
      v.v[q] = val;
    }
    jstart = ki + 2;
    lam = zmk(wr, -wi);
  }
  const bool pair = (ip != 0);
  double vmax = 1.0, vcrit = bignum;
  int j = jstart;
  while (j < n) {
    const bool blk = (j < n - 1) && (sub[j] != 0.0);
    if (!blk) {
      if (tnorms[j] > vcrit) {
        const double t1 = 1.0 / vmax;
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (lane + 64 * q >= ki) v.v[q] = v.v[q] * t1;
        vmax = 1.0;
        vcrit = bignum;
      }
      zd B[1] = {v.get(j)};
      const double Ae[1] = {Tc(j, j)};
      double xn;
      const double vscale = xsolve_dev(1, Ae, lam, B, smin, xn);
      if (vscale != 1.0) {
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (lane + 64 * q >= ki) v.v[q] = v.v[q] * vscale;
      }
      v.set(j, B[0]);
      vmax = fmax(pair ? zabs1(B[0]) : zabs(B[0]), vmax);
      vcrit = bignum / vmax;
      // remove x from the later right-hand sides
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = lane + 64 * q;
        if (i > j && i < n) v.v[q] = v.v[q] - B[0] * Tc(j, i);
      }
      j += 1;
    } else {
      // 2x2 block at rows j, j+1; with a pair the guard reads tnorms[j-1]
      const double beta = pair ? fmax(tnorms[j - 1], tnorms[j]) : fmax(tnorms[j + 1], tnorms[j]);
      if (beta > vcrit) {
        const double t1 = 1.0 / vmax;
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (lane + 64 * q >= ki) v.v[q] = v.v[q] * t1;
        vmax = 1.0;
        vcrit = bignum;
      }
      zd B[2] = {v.get(j), v.get(j + 1)};
      const double Ae[4] = {Tc(j, j), Tc(j, j + 1), sub[j], Tc(j + 1, j + 1)};
      double xn;
      const double vscale = xsolve_dev(2, Ae, lam, B, smin, xn);
      if (vscale != 1.0) {
#pragma unroll
        for (int q = 0; q < R; ++q)
          if (lane + 64 * q >= ki) v.v[q] = v.v[q] * vscale;
      }
      v.set(j, B[0]);
      v.set(j + 1, B[1]);
      vmax = pair ? fmax(fmax(zabs1(B[0]), zabs1(B[1])), vmax) : fmax(fmax(zabs(B[0]), zabs(B[1])), vmax);
      vcrit = bignum / vmax;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = lane + 64 * q;
        if (i > j + 1 && i < n) v.v[q] = v.v[q] - B[0] * Tc(j, i) - B[1] * Tc(j + 1, i);
      }
      j += 2;
    }
  }
}

template <int R, class VEC> GS_DEV void enormalize_w(VEC& v, int n);

// max-|.| normalisation of a finished column (maximum(abs, ...) or abs1)
template <int R, class VEC> GS_DEV void scale_by_max(VEC& y, int lo, int hi, bool use_abs1) {
  const int lane = threadIdx.x & 63;
  double m = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = lane + 64 * q;
    if (i >= lo && i < hi) m = fmax(m, use_abs1 ? zabs1(y.v[q]) : zabs(y.v[q]));
  }
  m = wave_max(m);
  const double s = 1.0 / m;
#pragma unroll
  for (int q = 0; q < R; ++q) y.v[q] = y.v[q] * s;
}

template <int R, class VEC> GS_DEV void enormalize_w(VEC& v, int n) {
  const int lane = threadIdx.x & 63;
  double am = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q)
    if (lane + 64 * q < n) am = fmax(am, fmax(fabs(v.v[q].x), fabs(v.v[q].y)));
  am = wave_max(am);
  double ss = 0.0, best = -1.0;
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
  const zd tt = ((1.0 / nrm) * zconj(vi0)) / sqrt(zabs2(vi0));
#pragma unroll
  for (int q = 0; q < R; ++q) v.v[q] = v.v[q] * tt;
  const zd r0 = v.get(bi);
  v.set(bi, zr(r0.x));
}

template <int NT, int R, bool GLOB> __global__ void __launch_bounds__(NT) dtrevc_kernel(TrevcArgsD a) {
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
  const double* Tg = a.T + o;
  const double* Zg = a.Z ? a.Z + o : nullptr;
  zd* Vg = reinterpret_cast<zd*>(a.V) + o;
  double* Tl = smem;
  double* sub = GLOB ? smem : Tl + (n * (n + 1)) / 2;
  double* tnorms = sub + n;
  int* kind = reinterpret_cast<int*>(tnorms + n);  // per column: 0 real, -1/+1 pair rows
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  constexpr int NW = NT / 64;
  const bool left = a.left != 0;
  if (!GLOB)
    for (int j = 0; j < n; ++j)
      for (int i = tid; i <= j; i += NT) Tl[trd(i, j)] = Tg[i + j * ldt];
  for (int j = tid; j < n; j += NT) sub[j] = (j + 1 < n) ? Tg[(j + 1) + j * ldt] : 0.0;
  __syncthreads();
  typedef typename std::conditional<GLOB, GlobD, TriLdsD>::type TA;
  TA Tc;
  if constexpr (GLOB) Tc = GlobD{Tg, ldt};
  else Tc = TriLdsD{Tl};
  for (int j = tid; j < n; j += NT) {
    double s = 0.0;
    for (int i = 0; i < j; ++i) s += fabs(Tc(i, j));
    tnorms[j] = s;
    // right: the second row of a pair (T[j, j-1] != 0) solves; left: the first
    const bool pr = (j > 0) && (sub[j - 1] != 0.0);
    const bool pl = (j + 1 < n) && (sub[j] != 0.0);
    kind[j] = left ? (pl ? 1 : 0) : (pr ? -1 : 0);
  }
  __syncthreads();
  // the reference walks ki = n..1 (right) or 1..n (left) and skips the
  // partner of each pair; mark the partners (ip == 1 / ip == -1 rows)
  if (tid == 0) {
    if (!left) {
      for (int ki = n - 1; ki >= 0; --ki)
        if (kind[ki] == -1) {
          if (ki - 1 >= 0) kind[ki - 1] = 2;
          --ki;
        }
    } else {
      for (int ki = 0; ki < n; ++ki)
        if (kind[ki] == 1) {
          if (ki + 1 < n) kind[ki + 1] = 2;
          ++ki;
        }
    }
  }
  __syncthreads();
  const double ulp = GS_EPS;
  const double smallnum = GS_FMIN * ((double)n / ulp);  // vectors.jl:146,475
  const double bignum = (1.0 - ulp) / smallnum;
  for (int t = w; t < n; t += NW) {
    const int ki = left ? t : n - 1 - t;
    const int kd = kind[ki];
    if (kd == 2) continue;
    WaveVecZ<R> v;
    if (!left) dsolve_right<R>(Tc, sub, tnorms, ki, kd, smallnum, bignum, v);
    else dsolve_left<R>(Tc, sub, tnorms, n, ki, kd, smallnum, bignum, v);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      if (i < n) Vg[i + ki * ldv] = v.v[q];
    }
  }
  __syncthreads();
  for (int t = w; t < n; t += NW) {
    const int ki = t;
    const int kd = kind[ki];
    if (kd == 2) continue;
    WaveVecZ<R> x;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      x.v[q] = (i < n) ? Vg[i + ki * ldv] : zr(0.0);
    }
    // solution support: right 0..ki, left ki..n-1
    const int lo = left ? ki : 0, hi = left ? n : ki + 1;
    WaveVecZ<R> y;
    if (Zg) {
      if ((!left && ki == 0 && kd == 0) || (left && ki == n - 1 && kd == 0)) {
#pragma unroll
        for (int q = 0; q < R; ++q) {
          const int r = lane + 64 * q;
          y.v[q] = (r < n) ? zr(Zg[r + (int64_t)ki * ldz]) : zr(0.0);
        }
      } else {
#pragma unroll
        for (int q = 0; q < R; ++q) y.v[q] = zr(0.0);
        for (int i = lo; i < hi; ++i) {
          const zd xi = x.get(i);
#pragma unroll
          for (int q = 0; q < R; ++q) {
            const int r = lane + 64 * q;
            if (r < n) y.v[q] = y.v[q] + Zg[r + (int64_t)i * ldz] * xi;
          }
        }
        scale_by_max<R>(y, 0, n, false);
        if (!left && kd == -1) {
#pragma unroll
          for (int q = 0; q < R; ++q) y.v[q] = zconj(y.v[q]);
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int i = lane + 64 * q;
        y.v[q] = (i >= lo && i < hi) ? x.v[q] : zr(0.0);
      }
      if (kd == 0) {
        scale_by_max<R>(y, lo, hi, false);
      } else {
        if (!left) {
#pragma unroll
          for (int q = 0; q < R; ++q) y.v[q] = zconj(y.v[q]);
        }
        scale_by_max<R>(y, lo, hi, true);
      }
    }
    if (a.normalize) enormalize_w<R>(y, n);
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = lane + 64 * q;
      if (r < n) Vg[r + ki * ldv] = y.v[q];
    }
    // the conjugate partner of a pair
    if (kd != 0) {
      const int pk = left ? ki + 1 : ki - 1;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = lane + 64 * q;
        if (r < n) Vg[r + pk * ldv] = zconj(y.v[q]);
      }
    }
  }
}

size_t dtrevc_lds_bytes(int n) { return 8 * ((size_t)n * (n + 1) / 2 + 2 * (size_t)n) + 4 * (size_t)n + 64; }
static size_t dtrevc_glob_lds_bytes(int n) { return 8 * 2 * (size_t)n + 4 * (size_t)n + 64; }

template <int NT, int R, bool GLOB> static void launch_dtrevc_t(const TrevcArgsD& a, size_t lds, hipStream_t s) {
  (void)hipFuncSetAttribute((const void*)dtrevc_kernel<NT, R, GLOB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
  hipLaunchKernelGGL((dtrevc_kernel<NT, R, GLOB>), dim3(a.count), dim3(NT), lds, s, a);
}

// Orders up to 192 keep T's triangle in LDS; above, T is read from global
// memory (subdiagonal, column norms and block kinds in LDS), up to
// dtrevc_max_n().
hipError_t launch_dtrevc(const TrevcArgsD& a, int nmax, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  if (nmax <= 64) launch_dtrevc_t<256, 1, false>(a, dtrevc_lds_bytes(nmax), s);
  else if (nmax <= 128) launch_dtrevc_t<256, 2, false>(a, dtrevc_lds_bytes(nmax), s);
  else if (nmax <= 192) launch_dtrevc_t<256, 3, false>(a, dtrevc_lds_bytes(nmax), s);
  else if (nmax <= 256) launch_dtrevc_t<512, 4, true>(a, dtrevc_glob_lds_bytes(nmax), s);
  else if (nmax <= 512) launch_dtrevc_t<512, 8, true>(a, dtrevc_glob_lds_bytes(nmax), s);
  else if (nmax <= 1024) launch_dtrevc_t<256, 16, true>(a, dtrevc_glob_lds_bytes(nmax), s);
  else return hipErrorInvalidValue;
  return hipGetLastError();
}
int dtrevc_max_n() { return 1024; }

// ------------------------------------------------------- larger orders
// One matrix of order 1024 < n <= trevc_big_max_n() over the whole GPU:
// the block kinds and column norms in global memory, one solve per
// workgroup (a single wavefront whose solution vector lives in LDS,
// WaveVecL), the solutions in a planar workspace X = [Re | Im] (n x 2n),
// the back-transform [Re | Im] of Z X as one FP64 MFMA GEMM, then one
// wavefront per column scales, normalises and writes V.  The per-column
// steps are those of dtrevc_kernel.
struct DBig {
  int n, left, normalize;
  const double* T;
  int64_t ldt;
  const double* Tt;  // T's transpose (left solves), ld n
  const double* Z;   // may be null
  int64_t ldz;
  zd* V;
  int64_t ldv;
  double* sub;
  double* tnorms;
  int* kind;
  double* X;  // n x 2n, ld n
  double* Y;  // n x 2n, ld n
};

// subdiagonal and block kinds (as in dtrevc_kernel), one workgroup
__global__ void __launch_bounds__(1024) dtrevc_big_kinds(DBig a) {
  extern __shared__ double smem[];
  int* kl = reinterpret_cast<int*>(smem);
  const int n = a.n, tid = threadIdx.x;
  const bool left = a.left != 0;
  for (int j = tid; j < n; j += 1024) {
    const double sj = (j + 1 < n) ? a.T[(j + 1) + (int64_t)j * a.ldt] : 0.0;
    const double sp = (j > 0) ? a.T[j + (int64_t)(j - 1) * a.ldt] : 0.0;
    a.sub[j] = sj;
    kl[j] = left ? (sj != 0.0 ? 1 : 0) : (sp != 0.0 ? -1 : 0);
  }
  __syncthreads();
  if (tid == 0) {
    if (!left) {
      for (int ki = n - 1; ki >= 0; --ki)
        if (kl[ki] == -1) {
          if (ki - 1 >= 0) kl[ki - 1] = 2;
          --ki;
        }
    } else {
      for (int ki = 0; ki < n; ++ki)
        if (kl[ki] == 1) {
          if (ki + 1 < n) kl[ki + 1] = 2;
          ++ki;
        }
    }
  }
  __syncthreads();
  for (int j = tid; j < n; j += 1024) a.kind[j] = kl[j];
}

// workgroup t solves column n-1-t (right) or t (left): the longest solves
// are dispatched first
template <int R, class TA> __global__ void __launch_bounds__(64) dtrevc_big_solve(DBig a, TA Tc) {
  extern __shared__ double smem[];
  const int n = a.n, lane = threadIdx.x & 63, t = blockIdx.x;
  const bool left = a.left != 0;
  const int ki = left ? t : n - 1 - t;
  const int kd = a.kind[ki];
  double* xr = a.X + (int64_t)ki * n;
  double* xi = xr + (int64_t)n * n;
  if (kd == 2) {
    for (int i = lane; i < n; i += 64) {
      xr[i] = 0.0;
      xi[i] = 0.0;
    }
    return;
  }
  const double ulp = GS_EPS;
  const double smallnum = GS_FMIN * ((double)n / ulp);  // vectors.jl:146,475
  const double bignum = (1.0 - ulp) / smallnum;
  WaveVecL<R> v{{reinterpret_cast<zd*>(smem)}};
  if (!left) dsolve_right<R>(Tc, a.sub, a.tnorms, ki, kd, smallnum, bignum, v);
  else dsolve_left<R>(Tc, a.sub, a.tnorms, n, ki, kd, smallnum, bignum, v);
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int i = lane + 64 * q;
    if (i < n) {
      const zd z = v.v[q];
      xr[i] = z.x;
      xi[i] = z.y;
    }
  }
}

template <int R> __global__ void __launch_bounds__(64) dtrevc_big_finish(DBig a) {
  extern __shared__ double smem[];
  const int n = a.n, lane = threadIdx.x & 63, ki = blockIdx.x;
  const bool left = a.left != 0;
  const int kd = a.kind[ki];
  if (kd == 2) return;
  WaveVecL<R> y{{reinterpret_cast<zd*>(smem)}};
  const int lo = left ? ki : 0, hi = left ? n : ki + 1;
  if (a.Z) {
    if ((!left && ki == 0 && kd == 0) || (left && ki == n - 1 && kd == 0)) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = lane + 64 * q;
        y.v[q] = (r < n) ? zr(a.Z[r + (int64_t)ki * a.ldz]) : zr(0.0);
      }
    } else {
      const double* yr = a.Y + (int64_t)ki * n;
      const double* yi = yr + (int64_t)n * n;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = lane + 64 * q;
        y.v[q] = (r < n) ? zmk(yr[r], yi[r]) : zr(0.0);
      }
      scale_by_max<R>(y, 0, n, false);
      if (!left && kd == -1) {
#pragma unroll
        for (int q = 0; q < R; ++q) y.v[q] = zconj(y.v[q]);
      }
    }
  } else {
    const double* xr = a.X + (int64_t)ki * n;
    const double* xi = xr + (int64_t)n * n;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int i = lane + 64 * q;
      y.v[q] = (i >= lo && i < hi) ? zmk(xr[i], xi[i]) : zr(0.0);
    }
    if (kd == 0) {
      scale_by_max<R>(y, lo, hi, false);
    } else {
      if (!left) {
#pragma unroll
        for (int q = 0; q < R; ++q) y.v[q] = zconj(y.v[q]);
      }
      scale_by_max<R>(y, lo, hi, true);
    }
  }
  if (a.normalize) enormalize_w<R>(y, n);
  const int pk = left ? ki + 1 : ki - 1;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = lane + 64 * q;
    if (r < n) {
      const zd z = y.v[q];
      a.V[r + (int64_t)ki * a.ldv] = z;
      if (kd != 0) a.V[r + (int64_t)pk * a.ldv] = zconj(z);
    }
  }
}

size_t dtrevc_big_ws_bytes(int n) {
  const size_t nn = (size_t)n * n;
  return 2 * trevc_seg(8 * (size_t)n) + trevc_seg(4 * (size_t)n) + trevc_seg(8 * nn) + 2 * trevc_seg(16 * nn);
}

template <int R> static hipError_t launch_dbig_r(const DBig& b, hipStream_t s) {
  const int n = b.n;
  const size_t lds = 64 * R * sizeof(zd);
  if (b.left) {
    auto k = dtrevc_big_solve<R, TransG<double>>;
    (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k, dim3(n), dim3(64), lds, s, b, TransG<double>{b.Tt, n});
  } else {
    auto k = dtrevc_big_solve<R, GlobD>;
    (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(k, dim3(n), dim3(64), lds, s, b, GlobD{b.T, b.ldt});
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (b.Z) {
    e = launch_dgemm(false, false, n, 2 * n, n, 1.0, b.Z, b.ldz, b.X, n, 0.0, b.Y, n, s);
    if (e != hipSuccess) return e;
  }
  (void)hipFuncSetAttribute((const void*)dtrevc_big_finish<R>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(dtrevc_big_finish<R>, dim3(n), dim3(64), lds, s, b);
  return hipGetLastError();
}

// a.T, a.Z, a.V point at the matrix (no offsets), ld from a.ldt/ldz/ldv;
// ws holds dtrevc_big_ws_bytes(n)
hipError_t launch_dtrevc_big(const TrevcArgsD& a, int n, void* ws, hipStream_t s) {
  if (n <= dtrevc_max_n() || n > trevc_big_max_n() || !ws) return hipErrorInvalidValue;
  const size_t nn = (size_t)n * n;
  char* p = static_cast<char*>(ws);
  DBig b;
  b.n = n;
  b.left = a.left;
  b.normalize = a.normalize;
  b.T = a.T;
  b.ldt = a.ldt > 0 ? a.ldt : n;
  b.Z = a.Z;
  b.ldz = a.ldz > 0 ? a.ldz : n;
  b.V = reinterpret_cast<zd*>(a.V);
  b.ldv = a.ldv > 0 ? a.ldv : n;
  b.sub = reinterpret_cast<double*>(p);
  p += trevc_seg(8 * (size_t)n);
  b.tnorms = reinterpret_cast<double*>(p);
  p += trevc_seg(8 * (size_t)n);
  b.kind = reinterpret_cast<int*>(p);
  p += trevc_seg(4 * (size_t)n);
  double* tt = reinterpret_cast<double*>(p);
  b.Tt = tt;
  p += trevc_seg(8 * nn);
  b.X = reinterpret_cast<double*>(p);
  p += trevc_seg(16 * nn);
  b.Y = reinterpret_cast<double*>(p);
  hipLaunchKernelGGL(dtrevc_big_kinds, dim3(1), dim3(1024), 4 * (size_t)n, s, b);
  hipLaunchKernelGGL(trevc_colnorm_kernel<double>, dim3((n + 3) / 4), dim3(256), 0, s, n, b.T, b.ldt, false,
                     b.tnorms);
  if (b.left)
    hipLaunchKernelGGL(trevc_transpose_kernel<double>, dim3((n + 31) / 32, (n + 31) / 32), dim3(256), 0, s, n, b.T,
                       b.ldt, tt);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return n <= 2048 ? launch_dbig_r<32>(b, s) : launch_dbig_r<64>(b, s);
}

int trevc_big_max_n() { return 4096; }

}  // namespace gs
