// TEST INFRASTRUCTURE ONLY: quad-double (about 212-bit) real type for the
// high-precision oracle that states the error of each double-precision result.
//
// Method: Y. Hida, X. S. Li, D. H. Bailey, "Algorithms for quad-double
// precision floating point arithmetic", ARITH-15 (2001).  A value is an
// unevaluated sum of four non-overlapping doubles x0 + x1 + x2 + x3 with
// |x0| > |x1| > ... .  Addition and multiplication are the "sloppy"
// variants of that paper (error bounded relative to the operands), division
// is long division with a final renormalisation, and the square root is three
// Newton steps on 1/sqrt(a) starting from the double-precision value.
#pragma once
#include <cmath>

namespace orc {

struct qd_real {
  double x[4];

  qd_real() { x[0] = x[1] = x[2] = x[3] = 0.0; }
  qd_real(double a) { x[0] = a; x[1] = x[2] = x[3] = 0.0; }
  qd_real(double a, double b, double c, double d) { x[0] = a; x[1] = b; x[2] = c; x[3] = d; }

  static qd_real epsilon() { return qd_real(1.21543267145725e-63); }  // 2^-209
  static qd_real min_normal() { return qd_real(1.6259745436952323e-260); }
  static qd_real max_value() { return qd_real(1.79769313486231570815e+308, 9.97920154767359795037e+291,
                                              5.53956966280111259858e+275, 3.07507889307840487279e+259); }
  bool isnan() const { return std::isnan(x[0]) || std::isnan(x[1]) || std::isnan(x[2]) || std::isnan(x[3]); }
  double to_double() const { return x[0] + (x[1] + (x[2] + x[3])); }

  // ------------------------------------------------------------ error-free ops
  static inline double quick_two_sum(double a, double b, double& e) {
    double s = a + b;
    e = b - (s - a);
    return s;
  }
  static inline double two_sum(double a, double b, double& e) {
    double s = a + b;
    double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
    return s;
  }
  static inline double two_prod(double a, double b, double& e) {
    double p = a * b;
    e = std::fma(a, b, -p);
    return p;
  }
  static inline void three_sum(double& a, double& b, double& c) {
    double t1, t2, t3;
    t1 = two_sum(a, b, t2);
    a = two_sum(c, t1, t3);
    b = two_sum(t2, t3, c);
  }
  static inline void three_sum2(double& a, double& b, double& c) {
    double t1, t2, t3;
    t1 = two_sum(a, b, t2);
    a = two_sum(c, t1, t3);
    b = t2 + t3;
  }
  static inline void renorm(double& c0, double& c1, double& c2, double& c3, double& c4) {
    double s0, s1, s2 = 0.0, s3 = 0.0;
    if (std::isinf(c0)) return;
    s0 = quick_two_sum(c3, c4, c4);
    s0 = quick_two_sum(c2, s0, c3);
    s0 = quick_two_sum(c1, s0, c2);
    c0 = quick_two_sum(c0, s0, c1);
    s0 = c0;
    s1 = c1;
    if (s1 != 0.0) {
      s1 = quick_two_sum(s1, c2, s2);
      if (s2 != 0.0) {
        s2 = quick_two_sum(s2, c3, s3);
        if (s3 != 0.0) s3 += c4; else s2 += c4;
      } else {
        s1 = quick_two_sum(s1, c3, s2);
        if (s2 != 0.0) s2 = quick_two_sum(s2, c4, s3); else s1 = quick_two_sum(s1, c4, s2);
      }
    } else {
      s0 = quick_two_sum(s0, c2, s1);
      if (s1 != 0.0) {
        s1 = quick_two_sum(s1, c3, s2);
        if (s2 != 0.0) s2 = quick_two_sum(s2, c4, s3); else s1 = quick_two_sum(s1, c4, s2);
      } else {
        s0 = quick_two_sum(s0, c3, s1);
        if (s1 != 0.0) s1 = quick_two_sum(s1, c4, s2); else s0 = quick_two_sum(s0, c4, s1);
      }
    }
    c0 = s0; c1 = s1; c2 = s2; c3 = s3;
  }

  // --------------------------------------------------------------- arithmetic
  friend qd_real operator+(const qd_real& a, const qd_real& b) {
    double s0, s1, s2, s3, t0, t1, t2, t3;
    s0 = two_sum(a.x[0], b.x[0], t0);
    s1 = two_sum(a.x[1], b.x[1], t1);
    s2 = two_sum(a.x[2], b.x[2], t2);
    s3 = two_sum(a.x[3], b.x[3], t3);
    s1 = two_sum(s1, t0, t0);
    three_sum(s2, t0, t1);
    three_sum2(s3, t0, t2);
    t0 = t0 + t1 + t3;
    renorm(s0, s1, s2, s3, t0);
    return qd_real(s0, s1, s2, s3);
  }
  friend qd_real operator-(const qd_real& a) { return qd_real(-a.x[0], -a.x[1], -a.x[2], -a.x[3]); }
  friend qd_real operator-(const qd_real& a, const qd_real& b) { return a + (-b); }

  friend qd_real operator*(const qd_real& a, const qd_real& b) {
    double p0, p1, p2, p3, p4, p5, q0, q1, q2, q3, q4, q5, t0, t1, s0, s1, s2;
    p0 = two_prod(a.x[0], b.x[0], q0);
    p1 = two_prod(a.x[0], b.x[1], q1);
    p2 = two_prod(a.x[1], b.x[0], q2);
    p3 = two_prod(a.x[0], b.x[2], q3);
    p4 = two_prod(a.x[1], b.x[1], q4);
    p5 = two_prod(a.x[2], b.x[0], q5);
    three_sum(p1, p2, q0);           // order-1 terms
    three_sum(p2, q1, q2);           // order-2 terms
    three_sum(p3, p4, p5);
    s0 = two_sum(p2, p3, t0);        // combine order-2 sums
    s1 = two_sum(q1, p4, t1);        // order-3 terms
    s2 = q2 + p5;
    s1 = two_sum(s1, t0, t0);
    s2 += (t0 + t1);
    s1 += a.x[0] * b.x[3] + a.x[1] * b.x[2] + a.x[2] * b.x[1] + a.x[3] * b.x[0] + q0 + q3 + q4 + q5;
    renorm(p0, p1, s0, s1, s2);
    return qd_real(p0, p1, s0, s1);
  }

  friend qd_real operator/(const qd_real& a, const qd_real& b) {
    double q0, q1, q2, q3, q4 = 0.0;
    qd_real r;
    q0 = a.x[0] / b.x[0];
    r = a - b * qd_real(q0);
    q1 = r.x[0] / b.x[0];
    r = r - b * qd_real(q1);
    q2 = r.x[0] / b.x[0];
    r = r - b * qd_real(q2);
    q3 = r.x[0] / b.x[0];
    if (!std::isfinite(q0)) return qd_real(q0);
    renorm(q0, q1, q2, q3, q4);
    return qd_real(q0, q1, q2, q3);
  }

  qd_real& operator+=(const qd_real& b) { *this = *this + b; return *this; }
  qd_real& operator-=(const qd_real& b) { *this = *this - b; return *this; }
  qd_real& operator*=(const qd_real& b) { *this = *this * b; return *this; }
  qd_real& operator/=(const qd_real& b) { *this = *this / b; return *this; }

  // lexicographic comparison of normalised expansions
  friend bool operator<(const qd_real& a, const qd_real& b) {
    for (int i = 0; i < 4; ++i) {
      if (a.x[i] < b.x[i]) return true;
      if (a.x[i] > b.x[i]) return false;
    }
    return false;
  }
  friend bool operator>(const qd_real& a, const qd_real& b) { return b < a; }
  friend bool operator<=(const qd_real& a, const qd_real& b) { return !(b < a); }
  friend bool operator>=(const qd_real& a, const qd_real& b) { return !(a < b); }
  friend bool operator==(const qd_real& a, const qd_real& b) {
    return a.x[0] == b.x[0] && a.x[1] == b.x[1] && a.x[2] == b.x[2] && a.x[3] == b.x[3];
  }
  friend bool operator!=(const qd_real& a, const qd_real& b) { return !(a == b); }
};

inline qd_real qd_abs(const qd_real& a) { return (a.x[0] < 0.0) ? -a : a; }
inline qd_real qd_copysign(const qd_real& a, const qd_real& b) {
  qd_real m = qd_abs(a);
  return std::signbit(b.x[0]) ? -m : m;
}
inline qd_real qd_sqrt(const qd_real& a) {
  if (a.x[0] == 0.0) return qd_real(0.0);
  if (a.x[0] < 0.0 || a.isnan()) return qd_real(std::nan(""));
  if (std::isinf(a.x[0])) return a;
  qd_real r(1.0 / std::sqrt(a.x[0]));
  qd_real h = a * qd_real(0.5);
  for (int i = 0; i < 3; ++i) r = r + (qd_real(0.5) - h * (r * r)) * r;
  return r * a;
}
inline qd_real qd_hypot(const qd_real& a, const qd_real& b) {
  qd_real x = qd_abs(a), y = qd_abs(b);
  qd_real m = (x < y) ? y : x;
  qd_real s = (x < y) ? x : y;
  if (m.x[0] == 0.0) return qd_real(0.0);
  if (std::isinf(m.x[0])) return m;
  qd_real r = s / m;
  return m * qd_sqrt(qd_real(1.0) + r * r);
}

}  // namespace orc
