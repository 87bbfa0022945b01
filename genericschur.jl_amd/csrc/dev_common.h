// Device-side scalar helpers shared by the gfx950 kernels: a ComplexF64
// type with the semantics the restated algorithm depends on (abs = hypot,
// abs1 = |re| + |im|, principal sqrt, Smith division), block reductions over
// 64-wide wavefronts, and the Householder reflector of xLARFG form used by
// GenericSchur.jl (householder.jl:12-102).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GS_DEV __device__ __forceinline__

struct zd {
  double x, y;
};

GS_DEV zd zmk(double a, double b) { zd r; r.x = a; r.y = b; return r; }
GS_DEV zd zr(double a) { return zmk(a, 0.0); }
GS_DEV zd operator+(zd a, zd b) { return zmk(a.x + b.x, a.y + b.y); }
GS_DEV zd operator-(zd a, zd b) { return zmk(a.x - b.x, a.y - b.y); }
GS_DEV zd operator-(zd a) { return zmk(-a.x, -a.y); }
GS_DEV zd operator*(zd a, zd b) { return zmk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
GS_DEV zd operator*(double s, zd a) { return zmk(s * a.x, s * a.y); }
GS_DEV zd operator*(zd a, double s) { return zmk(a.x * s, a.y * s); }
GS_DEV zd operator/(zd a, double s) { return zmk(a.x / s, a.y / s); }
GS_DEV zd zconj(zd a) { return zmk(a.x, -a.y); }
GS_DEV double zabs1(zd a) { return fabs(a.x) + fabs(a.y); }
GS_DEV double zabs(zd a) { return hypot(a.x, a.y); }
GS_DEV double zabs2(zd a) { return a.x * a.x + a.y * a.y; }
GS_DEV bool zeq(zd a, zd b) { return a.x == b.x && a.y == b.y; }

// Smith's algorithm (R. L. Smith, CACM 5(8), 1962).
GS_DEV zd zdiv(zd a, zd b) {
  const double ar = fabs(b.x), ai = fabs(b.y);
  if (!(ai > ar)) {
    if (b.x == 0.0 && b.y == 0.0) return zmk(a.x / b.x, a.y / b.x);
    const double r = b.y / b.x;
    const double d = b.x + b.y * r;
    return zmk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
  } else {
    const double r = b.x / b.y;
    const double d = b.x * r + b.y;
    return zmk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
  }
}
GS_DEV zd operator/(zd a, zd b) { return zdiv(a, b); }

// principal square root, with power-of-four scaling near the thresholds
GS_DEV zd zsqrt(zd z) {
  const double x = z.x, y = z.y;
  if (x == 0.0 && y == 0.0) return zmk(0.0, y);
  const double s = fmax(fabs(x), fabs(y));
  double scale = 1.0, unscale = 1.0;
  const double big = 1.7976931348623157e308 / 16.0;
  const double tiny = 2.2250738585072014e-308 * 16.0 / 2.220446049250313e-16;
  if (s > big) { scale = 0.25; unscale = 2.0; }
  else if (s < tiny) { scale = 1.0 / (2.220446049250313e-16 * 2.220446049250313e-16); unscale = 2.220446049250313e-16; }
  const double xs = x * scale, ys = y * scale;
  const double t = sqrt((fabs(xs) + hypot(xs, ys)) * 0.5);
  if (!(xs < 0.0)) return zmk(t * unscale, (ys / (2.0 * t)) * unscale);
  return zmk((fabs(ys) / (2.0 * t)) * unscale, copysign(t, ys) * unscale);
}

#define GS_EPS 2.220446049250313e-16
#define GS_FMIN 2.2250738585072014e-308
#define GS_FMAX 1.7976931348623157e308

// util.jl:562-570 (dlapy3 scaling)
GS_DEV double hypot3(double x, double y, double z) {
  const double xa = fabs(x), ya = fabs(y), za = fabs(z);
  const double w = fmax(fmax(xa, ya), za);
  const double rw = 1.0 / w;
  const double a = rw * xa, b = rw * ya, c = rw * za;
  return w * sqrt(a * a + b * b + c * c);
}

// Julia's Base.Math hypot(x, y) for Float64: the fused multiply-add branch,
// correctly rounded.
GS_DEV double jl_hypot(double x, double y) {
#pragma clang fp contract(off)
  double ax = fabs(x), ay = fabs(y);
  if (__builtin_isinf(ax) || __builtin_isinf(ay)) return __builtin_huge_val();
  if (ay > ax) {
    const double t = ax;
    ax = ay;
    ay = t;
  }
  if (ay <= ax * sqrt(GS_EPS / 2.0)) return ax;
  double scale = GS_EPS * sqrt(GS_FMIN);
  if (ax > sqrt(GS_FMAX / 2.0)) {
    ax = ax * scale;
    ay = ay * scale;
    scale = 1.0 / scale;
  } else if (ay < sqrt(GS_FMIN)) {
    ax = ax / scale;
    ay = ay / scale;
  } else {
    scale = 1.0;
  }
  double h = sqrt(__builtin_fma(ax, ax, ay * ay));
  const double hsq = h * h, axsq = ax * ax;
  h -= (__builtin_fma(-ay, ay, hsq - axsq) + __builtin_fma(h, h, -hsq) - __builtin_fma(ax, ax, -axsq)) / (2.0 * h);
  return h * scale;
}

// ------------------------------------------------------------ reductions
// Butterfly reductions leave a bitwise identical result in every lane, so
// control flow that depends on them stays uniform.
GS_DEV double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
GS_DEV double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
  return v;
}
GS_DEV zd wave_zsum(zd v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    v.x = v.x + __shfl_xor(v.x, o, 64);
    v.y = v.y + __shfl_xor(v.y, o, 64);
  }
  return v;
}

// Block-wide reductions for NT threads; `red` is LDS scratch of NT/64
// doubles.  Every thread must call them (they contain barriers when NT > 64).
template <int NT> GS_DEV double block_max(double v, double* red) {
  v = wave_max(v);
  if (NT > 64) {
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    v = red[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) v = fmax(v, red[i]);
  }
  return v;
}
template <int NT> GS_DEV double block_sum(double v, double* red) {
  v = wave_sum(v);
  if (NT > 64) {
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    v = red[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) v = v + red[i];
  }
  return v;
}

// Largest int over the block (every thread must call; barriers when NT > 64).
template <int NT> GS_DEV int block_max_int(int v, int* redi) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int u = __shfl_xor(v, o, 64);
    v = u > v ? u : v;
  }
  if (NT > 64) {
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) redi[w] = v;
    __syncthreads();
    v = redi[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) v = redi[i] > v ? redi[i] : v;
  }
  return v;
}

template <int NT> GS_DEV void bsync() {
  if (NT > 64) __syncthreads();
  else __builtin_amdgcn_wave_barrier();
}

// ------------------------------------------- split dot products (4 lanes)
// out[i] = sum_{k0 <= k < k1} f(k, i) for i0 <= i < i1: four consecutive
// lanes share an output, each summing a quarter of the k range, and the
// partial sums are combined with lane shuffles.  Every thread must call
// (the loop is uniform); out is written by one lane per output.
template <int NT, class F> GS_DEV void zdots4(int i0, int i1, int k0, int k1, F f, zd* out) {
  const int s = threadIdx.x & 3, il = threadIdx.x >> 2;
  const int seg = (k1 - k0 + 3) >> 2;
  const int ks = k0 + s * seg, ke = min(ks + seg, k1);
  for (int ib = i0; ib < i1; ib += NT / 4) {
    const int i = ib + il;
    zd acc = zr(0.0);
    if (i < i1) {
#pragma unroll 4
      for (int k = ks; k < ke; ++k) acc = acc + f(k, i);
    }
    acc.x += __shfl_xor(acc.x, 1, 64);
    acc.y += __shfl_xor(acc.y, 1, 64);
    acc.x += __shfl_xor(acc.x, 2, 64);
    acc.y += __shfl_xor(acc.y, 2, 64);
    if (s == 0 && i < i1) out[i] = acc;
  }
}
// As zdots4, with the output handed to epi(i, sum) on the lane that holds it.
template <int NT, class F, class E> GS_DEV void zdots4e(int i0, int i1, int k0, int k1, F f, E epi) {
  const int s = threadIdx.x & 3, il = threadIdx.x >> 2;
  const int seg = (k1 - k0 + 3) >> 2;
  const int ks = k0 + s * seg, ke = min(ks + seg, k1);
  for (int ib = i0; ib < i1; ib += NT / 4) {
    const int i = ib + il;
    zd acc = zr(0.0);
    if (i < i1) {
#pragma unroll 4
      for (int k = ks; k < ke; ++k) acc = acc + f(k, i);
    }
    acc.x += __shfl_xor(acc.x, 1, 64);
    acc.y += __shfl_xor(acc.y, 1, 64);
    acc.x += __shfl_xor(acc.x, 2, 64);
    acc.y += __shfl_xor(acc.y, 2, 64);
    if (s == 0 && i < i1) epi(i, acc);
  }
}
template <int NT, class F, class E> GS_DEV void ddots4e(int i0, int i1, int k0, int k1, F f, E epi) {
  const int s = threadIdx.x & 3, il = threadIdx.x >> 2;
  const int seg = (k1 - k0 + 3) >> 2;
  const int ks = k0 + s * seg, ke = min(ks + seg, k1);
  for (int ib = i0; ib < i1; ib += NT / 4) {
    const int i = ib + il;
    double acc = 0.0;
    if (i < i1) {
#pragma unroll 4
      for (int k = ks; k < ke; ++k) acc += f(k, i);
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    if (s == 0 && i < i1) epi(i, acc);
  }
}
template <int NT, class F> GS_DEV void ddots4(int i0, int i1, int k0, int k1, F f, double* out) {
  const int s = threadIdx.x & 3, il = threadIdx.x >> 2;
  const int seg = (k1 - k0 + 3) >> 2;
  const int ks = k0 + s * seg, ke = min(ks + seg, k1);
  for (int ib = i0; ib < i1; ib += NT / 4) {
    const int i = ib + il;
    double acc = 0.0;
    if (i < i1) {
#pragma unroll 4
      for (int k = ks; k < ke; ++k) acc += f(k, i);
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    if (s == 0 && i < i1) out[i] = acc;
  }
}

// util.jl:41-80 (xLASCL-style multiply by cto/cfrom), the multipliers in the
// order the reference applies them.  Returns how many were written (<= 4).
GS_DEV int safescale_muls(double cfrom, double cto, double* mul) {
  const double smlnum = GS_FMIN;
  const double bignum = 1.0 / smlnum;
  double cfromc = cfrom, ctoc = cto;
  bool done = false;
  int k = 0;
  while (!done && k < 64) {
    double m;
    const double cfrom1 = cfromc * smlnum;
    if (cfrom1 == cfromc) {
      m = ctoc / cfromc;
      done = true;
    } else {
      const double cto1 = ctoc / bignum;
      if (cto1 == ctoc) {
        m = cto;
        done = true;
        cfromc = 1.0;
      } else if (fabs(cfrom1) > fabs(ctoc) && ctoc != 0.0) {
        m = smlnum;
        cfromc = 1.0;
      } else if (fabs(cto1) > fabs(cfromc)) {
        m = bignum;
        cfromc = cfrom1;
      } else {
        m = ctoc / cfromc;
        done = true;
      }
    }
    if (k < 4) mul[k] = m;
    ++k;
  }
  return k < 4 ? k : 4;
}

// ------------------------------------------------------------ reflectors
// Complex xLARFG on the 2-vector (a, x), householder.jl:56-102 with n = 2.
// On exit a = beta (real), x = v2; returns tau.
GS_DEV zd zreflector2(zd& a, zd& x) {
  const double sfmin = GS_FMIN / GS_EPS;
  double ar = a.x, ai = a.y;
  double xnorm = zabs(x);
  if (xnorm == 0.0 && ai == 0.0) return zr(0.0);
  double beta = -copysign(hypot3(ar, ai, xnorm), ar);
  int kount = 0;
  zd alpha = a;
  if (fabs(beta) < sfmin) {
    const double rsfmin = 1.0 / sfmin;
    bool smallb = true;
    while (smallb) {
      kount += 1;
      x = x * rsfmin;
      beta *= rsfmin;
      ar *= rsfmin;
      ai *= rsfmin;
      smallb = (fabs(beta) < sfmin) && (kount < 20);
    }
    xnorm = zabs(x);
    alpha = zmk(ar, ai);
    beta = -copysign(hypot3(ar, ai, xnorm), ar);
  }
  const zd tau = zmk((beta - ar) / beta, -ai / beta);
  const zd t = zdiv(zr(1.0), alpha - zr(beta));
  x = x * t;
  for (int j = 0; j < kount; ++j) beta *= sfmin;
  a = zr(beta);
  return tau;
}

// Real xLARFG on (a, x1[, x2]) of length nr = 2 or 3, householder.jl:12-54.
// The norm of the tail follows _norm2 (util.jl:506-525) for 1 or 2 entries.
GS_DEV double norm2_small(double x1, double x2, int m) {
  if (m == 1) return fabs(x1);
  double scale = 0.0, ssq = 0.0;
  const double xs[2] = {x1, x2};
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    if (xs[i] != 0.0) {
      const double a = fabs(xs[i]);
      if (scale < a) {
        const double q = scale / a;
        ssq = 1.0 + ssq * (q * q);
        scale = a;
      } else {
        const double q = a / scale;
        ssq = ssq + q * q;
      }
    }
  }
  return scale * sqrt(ssq);
}

GS_DEV double dreflector(int nr, double& a, double& x1, double& x2) {
  if (nr <= 1) return 0.0;
  const double sfmin = 2.0 * GS_FMIN / GS_EPS;
  const int m = nr - 1;
  double alpha = a;
  double xnorm = norm2_small(x1, x2, m);
  if (xnorm == 0.0) return 0.0;
  double beta = -copysign(hypot(alpha, xnorm), alpha);
  int kount = 0;
  if (fabs(beta) < sfmin) {
    const double rsfmin = 1.0 / sfmin;
    bool smallb = true;
    while (smallb) {
      kount += 1;
      x1 *= rsfmin;
      if (m > 1) x2 *= rsfmin;
      beta *= rsfmin;
      alpha *= rsfmin;
      smallb = (fabs(beta) < sfmin) && (kount < 20);
    }
    xnorm = norm2_small(x1, x2, m);
    beta = -copysign(hypot(alpha, xnorm), alpha);
  }
  const double tau = (beta - alpha) / beta;
  const double t = 1.0 / (alpha - beta);
  x1 *= t;
  if (m > 1) x2 *= t;
  for (int j = 0; j < kount; ++j) beta *= sfmin;
  a = beta;
  return tau;
}

// safemin(Float64) (util.jl:5-12) is floatmin for IEEE double.
#define GS_SAFEMIN GS_FMIN

// Square root of x in [2^-4, 2^4] (the reflectors' scaled sums of squares):
// the hardware reciprocal square root estimate, one coupled Newton step for
// sqrt(x) and 1 / (2 sqrt(x)), and one residual correction; within an ulp
// or so, no range or special-value handling.
GS_DEV double sqrt_scaled(double x) {
  const double r = __builtin_amdgcn_rsq(x);
  double s = x * r, h = 0.5 * r;
  const double e = fma(-s, h, 0.5);
  s = fma(s, e, s);
  h = fma(h, e, h);
  return fma(fma(-s, s, x), h, s);
}

// ------------------------------------------------ low-latency reflectors
// The same reflectors (householder.jl:12-102) for the bulge chase, with the
// norms folded into one scaled square root and the divisions merged into a
// single reciprocal (entries scaled by a power of two, the reciprocal from
// the hardware estimate and two Newton steps).  Same result
// up to rounding.  Outside |beta| in [2^-300, 2^300] (where the merged
// reciprocal could over/underflow) they fall back to the reference formulas
// above.
GS_DEV zd zreflector2_fast(zd& a, zd& x) {
  const double ar = a.x, ai = a.y;
  if (x.x == 0.0 && x.y == 0.0 && ai == 0.0) return zr(0.0);
  // the norm of (ar, ai, x) with the entries scaled by a power of two (exact)
  const double w = fmax(fmax(fabs(ar), fabs(ai)), fmax(fabs(x.x), fabs(x.y)));
  int e;
  (void)frexp(w, &e);
  const double p = ldexp(ar, -e), q = ldexp(ai, -e), s = ldexp(x.x, -e), t = ldexp(x.y, -e);
  const double beta = -copysign(ldexp(sqrt_scaled(p * p + q * q + s * s + t * t), e), ar);
  const double ab = fabs(beta);
  if (!(ab >= 0x1p-300 && ab <= 0x1p300)) return zreflector2(a, x);
  // tau = ((beta - ar) / beta, -ai / beta);  1 / (alpha - beta) = conj(d) / |d|^2
  const double dr = ar - beta;
  const double dd = dr * dr + ai * ai;
  // 1 / (beta dd), in [2^-905, 2^900] here: reciprocal estimate and two
  // Newton steps (no range fix-ups needed)
  const double bd = beta * dd;
  double r = __builtin_amdgcn_rcp(bd);
  r = fma(r, fma(-bd, r, 1.0), r);
  r = fma(r, fma(-bd, r, 1.0), r);
  const double rb = dd * r;     // 1 / beta
  const double rdd = beta * r;  // 1 / |d|^2
  const zd tau = zmk((beta - ar) * rb, -ai * rb);
  const zd inv = zmk(dr * rdd, -ai * rdd);
  x = x * inv;
  a = zr(beta);
  return tau;
}

// zreflector2_fast for a wavefront whose lanes all hold the same a and x:
// the formulas run straight through and one branch on the scalar unit
// replaces them by zreflector2's when x = 0 with a real (tau = 0) or |beta|
// lies outside [2^-300, 2^300].  Inside that range the sum of squares
// needs no scaling: it lies in [2^-600, 2^600], so no square overflows and
// squares that underflow are below its rounding; the square root is the
// same iteration as sqrt_scaled's.
GS_DEV bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }
GS_DEV zd zreflector2_wave(zd& a, zd& x) {
  const double ar = a.x, ai = a.y;
  const double beta = -copysign(sqrt_scaled(ar * ar + ai * ai + x.x * x.x + x.y * x.y), ar);
  const double ab = fabs(beta);
  const bool rare = ((x.x == 0.0) & (x.y == 0.0) & (ai == 0.0)) | !((ab >= 0x1p-300) & (ab <= 0x1p300));
  const double dr = ar - beta;
  const double dd = dr * dr + ai * ai;
  const double bd = beta * dd;
  double r = __builtin_amdgcn_rcp(bd);
  r = fma(r, fma(-bd, r, 1.0), r);
  r = fma(r, fma(-bd, r, 1.0), r);
  const double rb = dd * r;
  const double rdd = beta * r;
  zd tau = zmk((beta - ar) * rb, -ai * rb);
  const zd inv = zmk(dr * rdd, -ai * rdd);
  zd xo = x * inv, ao = zr(beta);
  if (wave_any(rare)) {
    ao = a;
    xo = x;
    tau = zreflector2(ao, xo);
  }
  a = ao;
  x = xo;
  return tau;
}

// dreflector_fast for three-element reflectors whose input every lane holds
// alike (the chase's main steps): the unscaled sum of squares (within the
// range checked below it neither overflows nor loses anything to underflow,
// as for zreflector2_wave), the formulas straight through and one scalar
// branch to the reference formulas when x = 0 or |beta| lies outside
// [2^-300, 2^300].
GS_DEV double dreflector3_wave(double& a, double& x1, double& x2) {
  const double alpha = a;
  const double beta = -copysign(sqrt_scaled(alpha * alpha + x1 * x1 + x2 * x2), alpha);
  const double ab = fabs(beta);
  const bool rare = ((x1 == 0.0) & (x2 == 0.0)) | !((ab >= 0x1p-300) & (ab <= 0x1p300));
  const double d = alpha - beta;
  const double bd = beta * d;
  double r = __builtin_amdgcn_rcp(bd);
  r = fma(r, fma(-bd, r, 1.0), r);
  r = fma(r, fma(-bd, r, 1.0), r);
  double tau = -(d * d) * r;
  const double sc = beta * r;
  double ao = beta, y1 = x1 * sc, y2 = x2 * sc;
  if (wave_any(rare)) {
    ao = a;
    y1 = x1;
    y2 = x2;
    tau = dreflector(3, ao, y1, y2);
  }
  a = ao;
  x1 = y1;
  x2 = y2;
  return tau;
}

GS_DEV double dreflector_fast(int nr, double& a, double& x1, double& x2) {
  if (nr <= 1) return 0.0;
  const double alpha = a;
  const double x2v = (nr == 3) ? x2 : 0.0;
  if (x1 == 0.0 && x2v == 0.0) return 0.0;
  const double w = fmax(fabs(alpha), fmax(fabs(x1), fabs(x2v)));
  int e;
  (void)frexp(w, &e);
  const double p = ldexp(alpha, -e), q = ldexp(x1, -e), s = ldexp(x2v, -e);
  const double beta = -copysign(ldexp(sqrt_scaled(p * p + q * q + s * s), e), alpha);
  const double ab = fabs(beta);
  if (!(ab >= 0x1p-300 && ab <= 0x1p300)) return dreflector(nr, a, x1, x2);
  // tau = (beta - alpha) / beta, scale = 1 / (alpha - beta); beta d lies in
  // [2^-600, 2^602] here: reciprocal estimate and two Newton steps
  const double d = alpha - beta;
  const double bd = beta * d;
  double r = __builtin_amdgcn_rcp(bd);
  r = fma(r, fma(-bd, r, 1.0), r);
  r = fma(r, fma(-bd, r, 1.0), r);
  const double tau = -(d * d) * r;
  const double sc = beta * r;
  x1 *= sc;
  if (nr == 3) x2 *= sc;
  a = beta;
  return tau;
}

// uniform broadcast of lane l's double (l must be wave-uniform)
GS_DEV double rdfirst(double v) {
  const unsigned long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(u & 0xffffffffu));
  const int hi = __builtin_amdgcn_readfirstlane((int)(u >> 32));
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

GS_DEV double rdlane(double v, int l) {
  const unsigned long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(u & 0xffffffffu), l);
  const int hi = __builtin_amdgcn_readlane((int)(u >> 32), l);
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// ordering point for LDS traffic among the lanes of one wavefront: the
// hardware processes a wave's LDS operations in order, so only compiler
// reordering has to be prevented
// Quad permutation of a double across lanes (DPP quad_perm, CTRL = s0 |
// s1 << 2 | s2 << 4 | s3 << 6: lane 4q + i takes lane 4q + s_i's value).
template <int CTRL> GS_DEV double qperm(double v) {
  const unsigned long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffu), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Workgroup barrier ordering LDS accesses only: global memory operations
// still in flight stay in flight.
GS_DEV void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

GS_DEV void wsync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// block synchronisation that degrades to a wavefront ordering point when
// the block is a single wavefront
template <int NT> GS_DEV void bsync_full() {
  if (NT > 64) __syncthreads();
  else wsync();
}

// The right update of a pair (x0, x1) by a logged complex reflector,
// ss = tau1 x0 + tau2 x1, x0 -= ss, x1 -= ss conj(v2) with tau2 = Re(tau1 v2)
// (GenericSchur.jl:447-460), and a complex product, with a fixed rounding:
// no contraction beyond the fma calls written here.  Every path that applies
// a reflector or a scaling to Z (in place during the sweeps, or replayed
// from the Z log) goes through these, so the paths give the same bits.
GS_DEV void zrot_right_fixed(zd& x0, zd& x1, zd tau1, zd v2) {
#pragma clang fp contract(off)
  const double tau2 = fma(tau1.x, v2.x, -(tau1.y * v2.y));
  const double sr = fma(tau1.x, x0.x, fma(-tau1.y, x0.y, tau2 * x1.x));
  const double si = fma(tau1.x, x0.y, fma(tau1.y, x0.x, tau2 * x1.y));
  x0 = zmk(x0.x - sr, x0.y - si);
  x1 = zmk(fma(-sr, v2.x, fma(-si, v2.y, x1.x)), fma(-si, v2.x, fma(sr, v2.y, x1.y)));
}
GS_DEV zd zmul_fixed(zd a, zd b) {
#pragma clang fp contract(off)
  return zmk(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x));
}

// Copy the W x W block at src (column-major, leading dimension ld, W <= WMAX)
// into LDS with column pitch WL (win_load), or back (win_store), by one
// wavefront.  A batch's global loads are all issued before the first of them
// is used, so the block costs one memory latency per batch, not one per
// column.  Lane l covers row l % R of columns l / R, l / R + 64 / R, ...;
// loads clamp the row and column into the block and are not masked, stores
// are.
// (Entries in rows < rskip of columns >= cskip are left as they are: the
// caller fills them from elsewhere.)
template <class T, int WMAX, int WL>
GS_DEV void win_load(const T* src, int64_t ld, int W, T* win, int lane, int rskip = 0, int cskip = 0) {
  constexpr int R = WMAX < 64 ? WMAX : 64;
  constexpr int CPI = 64 / R;
  constexpr int NI = WMAX / CPI;
  constexpr int B0 = 256 / (int)sizeof(T);
  constexpr int B = NI < B0 ? NI : B0;
  static_assert(64 % R == 0 && WMAX % CPI == 0 && NI % B == 0 && WL >= WMAX, "window shape");
  W = __builtin_amdgcn_readfirstlane(W);  // uniform: branches on SCC
  const int r = lane % R, ch = lane / R;
  const T* bp = src + min(r, W - 1);
  for (int i0 = 0; i0 < NI && CPI * i0 < W; i0 += B) {
    T t[B];
#pragma unroll
    for (int i = 0; i < B; ++i) t[i] = bp[(int64_t)min(CPI * (i0 + i) + ch, W - 1) * ld];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int c = CPI * (i0 + i) + ch;
      if (r < W && c < W && !(r < rskip && c >= cskip)) win[r + c * WL] = t[i];
    }
  }
}
// win_load in two halves, so that other work can run while the loads are in
// flight: win_fetch issues one batch of loads (B = WMAX / (64 / R) entries
// per lane) into t, win_put writes them to the window.
template <class T, int WMAX> struct WinRegs {
  static constexpr int R = WMAX < 64 ? WMAX : 64;
  static constexpr int NI = WMAX / (64 / R);
  T t[NI];
};
template <class T, int WMAX> GS_DEV void win_fetch(const T* src, int64_t ld, int W, WinRegs<T, WMAX>& w, int lane) {
  constexpr int R = WinRegs<T, WMAX>::R, CPI = 64 / R, NI = WinRegs<T, WMAX>::NI;
  W = __builtin_amdgcn_readfirstlane(W);
  const int r = lane % R, ch = lane / R;
  const T* bp = src + min(r, W - 1);
#pragma unroll
  for (int i = 0; i < NI; ++i) w.t[i] = bp[(int64_t)min(CPI * i + ch, W - 1) * ld];
}
template <class T, int WMAX, int WL>
GS_DEV void win_put(const WinRegs<T, WMAX>& w, int W, T* win, int lane, int rskip = 0, int cskip = 0) {
  constexpr int R = WinRegs<T, WMAX>::R, CPI = 64 / R, NI = WinRegs<T, WMAX>::NI;
  W = __builtin_amdgcn_readfirstlane(W);
  const int r = lane % R, ch = lane / R;
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int c = CPI * i + ch;
    if (r < W && c < W && !(r < rskip && c >= cskip)) win[r + c * WL] = w.t[i];
  }
}

template <class T, int WMAX, int WL> GS_DEV void win_store(T* dst, int64_t ld, int W, const T* win, int lane) {
  constexpr int R = WMAX < 64 ? WMAX : 64;
  constexpr int CPI = 64 / R;
  constexpr int NI = WMAX / CPI;
  constexpr int B0 = 128 / (int)sizeof(T);  // LDS reads only: short batches suffice
  constexpr int B = NI < B0 ? NI : B0;
  static_assert(64 % R == 0 && WMAX % CPI == 0 && NI % B == 0 && WL >= WMAX, "window shape");
  W = __builtin_amdgcn_readfirstlane(W);  // uniform: branches on SCC
  const int r = lane % R, ch = lane / R;
  T* bp = dst + r + (int64_t)ch * ld;
  // opaque offset: keeps the compiler from computing the LDS addresses once
  // and holding them in registers for the whole kernel
  int o = 0;
  asm volatile("" : "+s"(o));
  win += o;
  for (int i0 = 0; i0 < NI && CPI * i0 < W; i0 += B) {
    T t[B];
#pragma unroll
    for (int i = 0; i < B; ++i) t[i] = win[r + (CPI * (i0 + i) + ch) * WL];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int c0 = CPI * (i0 + i);
      if (r < W && c0 + ch < W) bp[(int64_t)c0 * ld] = t[i];
    }
  }
}
