// TEST INFRASTRUCTURE ONLY: scalar and complex arithmetic for the CPU oracle.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load the oracle; the product library never does.
//
// Real types supported: double, __float128 (113-bit, libquadmath) and
// qd_real (212-bit quad-double, qd.hpp).  The complex type `cx<R>` follows
// Julia's Complex semantics where the restated algorithm depends on them:
// abs(z) = hypot(re, im), abs1(z) = |re| + |im|, principal sqrt, and a
// scaled Smith (1962) division.
#pragma once
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <limits>
#include <quadmath.h>

#include "qd.hpp"

namespace orc {

// ---------------------------------------------------------------- real traits
template <class R> struct rt;

template <> struct rt<double> {
  static double eps() { return std::numeric_limits<double>::epsilon(); }
  static double fmin() { return std::numeric_limits<double>::min(); }
  static double fmax() { return std::numeric_limits<double>::max(); }
  static double sqrt(double x) { return std::sqrt(x); }
  static double abs(double x) { return std::fabs(x); }
  static double copysign(double x, double y) { return std::copysign(x, y); }
  static double hypot(double x, double y) { return std::hypot(x, y); }
  static bool isnan(double x) { return std::isnan(x); }
  static double from(double x) { return x; }
  static double to_double(double x) { return x; }
};

template <> struct rt<__float128> {
  static __float128 eps() { return FLT128_EPSILON; }
  static __float128 fmin() { return FLT128_MIN; }
  static __float128 fmax() { return FLT128_MAX; }
  static __float128 sqrt(__float128 x) { return sqrtq(x); }
  static __float128 abs(__float128 x) { return fabsq(x); }
  static __float128 copysign(__float128 x, __float128 y) { return copysignq(x, y); }
  static __float128 hypot(__float128 x, __float128 y) { return hypotq(x, y); }
  static bool isnan(__float128 x) { return isnanq(x); }
  static __float128 from(double x) { return (__float128)x; }
  static double to_double(__float128 x) { return (double)x; }
};

template <> struct rt<qd_real> {
  static qd_real eps() { return qd_real::epsilon(); }
  static qd_real fmin() { return qd_real::min_normal(); }
  static qd_real fmax() { return qd_real::max_value(); }
  static qd_real sqrt(const qd_real& x) { return qd_sqrt(x); }
  static qd_real abs(const qd_real& x) { return qd_abs(x); }
  static qd_real copysign(const qd_real& x, const qd_real& y) { return qd_copysign(x, y); }
  static qd_real hypot(const qd_real& x, const qd_real& y) { return qd_hypot(x, y); }
  static bool isnan(const qd_real& x) { return x.isnan(); }
  static qd_real from(double x) { return qd_real(x); }
  static double to_double(const qd_real& x) { return x.to_double(); }
};

template <class R> inline R rmax(const R& a, const R& b) { return (a < b) ? b : a; }
template <class R> inline R rmin(const R& a, const R& b) { return (b < a) ? b : a; }
template <class R> inline R rmax3(const R& a, const R& b, const R& c) { return rmax(rmax(a, b), c); }

// ------------------------------------------------------------- complex type
template <class R> struct cx {
  R re, im;
  cx() : re(0.0), im(0.0) {}
  cx(const R& r) : re(r), im(0.0) {}
  cx(const R& r, const R& i) : re(r), im(i) {}
  cx& operator+=(const cx& b) { re = re + b.re; im = im + b.im; return *this; }
  cx& operator-=(const cx& b) { re = re - b.re; im = im - b.im; return *this; }
  cx& operator*=(const cx& b) { *this = *this * b; return *this; }
  cx& operator*=(const R& s) { re = re * s; im = im * s; return *this; }
  friend cx operator+(const cx& a, const cx& b) { return cx(a.re + b.re, a.im + b.im); }
  friend cx operator-(const cx& a, const cx& b) { return cx(a.re - b.re, a.im - b.im); }
  friend cx operator-(const cx& a) { return cx(-a.re, -a.im); }
  friend cx operator*(const cx& a, const cx& b) {
    return cx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
  }
  friend cx operator*(const R& s, const cx& a) { return cx(s * a.re, s * a.im); }
  friend cx operator*(const cx& a, const R& s) { return cx(a.re * s, a.im * s); }
  friend cx operator/(const cx& a, const R& s) { return cx(a.re / s, a.im / s); }
  friend cx operator/(const cx& a, const cx& b) { return cdiv(a, b); }
  friend bool operator==(const cx& a, const cx& b) { return a.re == b.re && a.im == b.im; }
  friend bool operator!=(const cx& a, const cx& b) { return !(a == b); }

  // Smith's algorithm (R. L. Smith, CACM 5(8), 1962) with pre-scaling of the
  // divisor so that neither the ratio nor the denominator overflows.
  static cx cdiv(const cx& a, const cx& b) {
    const R ar = rt<R>::abs(b.re), ai = rt<R>::abs(b.im);
    if (!(ai > ar) ) {
      if (b.re == R(0.0) && b.im == R(0.0)) {
        return cx(a.re / b.re, a.im / b.re);  // IEEE inf/nan propagation
      }
      const R r = b.im / b.re;
      const R d = b.re + b.im * r;
      return cx((a.re + a.im * r) / d, (a.im - a.re * r) / d);
    } else {
      const R r = b.re / b.im;
      const R d = b.re * r + b.im;
      return cx((a.re * r + a.im) / d, (a.im * r - a.re) / d);
    }
  }
};

template <class R> inline R real(const R& x) { return x; }
template <class R> inline R imag(const R&) { return R(0.0); }
template <class R> inline R real(const cx<R>& z) { return z.re; }
template <class R> inline R imag(const cx<R>& z) { return z.im; }
template <class R> inline R conj(const R& x) { return x; }
template <class R> inline cx<R> conj(const cx<R>& z) { return cx<R>(z.re, -z.im); }
template <class R> inline R absv(const R& x) { return rt<R>::abs(x); }
template <class R> inline R absv(const cx<R>& z) { return rt<R>::hypot(z.re, z.im); }
// abs1 (util.jl:31-32)
template <class R> inline R abs1(const R& x) { return rt<R>::abs(x); }
template <class R> inline R abs1(const cx<R>& z) { return rt<R>::abs(z.re) + rt<R>::abs(z.im); }
template <class R> inline R abs2(const R& x) { return x * x; }
template <class R> inline R abs2(const cx<R>& z) { return z.re * z.re + z.im * z.im; }
template <class R> inline bool isreal(const R&) { return true; }
template <class R> inline bool isreal(const cx<R>& z) { return z.im == R(0.0); }

// principal square root, scaled by a power of 4 when |z| is near the
// over/underflow thresholds
template <class R> inline cx<R> csqrt(const cx<R>& z) {
  const R x = z.re, y = z.im;
  if (x == R(0.0) && y == R(0.0)) return cx<R>(R(0.0), y);
  R ax = rt<R>::abs(x), ay = rt<R>::abs(y);
  R s = rmax(ax, ay);
  R scale = R(1.0), unscale = R(1.0);
  const R big = rt<R>::fmax() / R(16.0);
  const R tiny = rt<R>::fmin() * R(16.0) / rt<R>::eps();
  if (s > big) { scale = R(0.25); unscale = R(2.0); }
  else if (s < tiny) { scale = R(1.0) / (rt<R>::eps() * rt<R>::eps()); unscale = rt<R>::eps(); }
  const R xs = x * scale, ys = y * scale;
  const R t = rt<R>::sqrt((rt<R>::abs(xs) + rt<R>::hypot(xs, ys)) * R(0.5));
  if (!(xs < R(0.0))) {
    return cx<R>(t * unscale, (ys / (R(2.0) * t)) * unscale);
  } else {
    return cx<R>((rt<R>::abs(ys) / (R(2.0) * t)) * unscale, rt<R>::copysign(t, ys) * unscale);
  }
}

template <class T> struct real_of { typedef T type; };
template <class R> struct real_of<cx<R>> { typedef R type; };
template <class T> struct is_cx { static const bool value = false; };
template <class R> struct is_cx<cx<R>> { static const bool value = true; };

}  // namespace orc
