// Standard-form 2x2 real Schur block, shared by the real QR kernels.
#pragma once
#include "dev_common.h"

namespace gs {

// _gs2x2! (GenericSchur.jl:716-803); hypot is Julia's (jl_hypot)
GS_DEV void gs2x2_dev(double& a, double& b, double& c, double& d, double& cs, double& sn, zd& w1, zd& w2) {
  auto sgn = [](double x) { return (x < 0.0) ? -1.0 : 1.0; };
  const double half = 0.5;
  const double small = 4.0 * GS_EPS;
  if (c == 0.0) {
    cs = 1.0;
    sn = 0.0;
  } else if (b == 0.0) {
    cs = 0.0;
    sn = 1.0;
    const double a0 = a, c0 = c, d0 = d;
    a = d0;
    b = -c0;
    c = 0.0;
    d = a0;
  } else if ((a - d) == 0.0 && (b * c < 0.0)) {
    cs = 1.0;
    sn = 0.0;
  } else {
    const double asubd = a - d;
    double p = half * asubd;
    const double bcmax = fmax(fabs(b), fabs(c));
    const double bcmis = fmin(fabs(b), fabs(c)) * sgn(b) * sgn(c);
    const double scale = fmax(fabs(p), bcmax);
    double z = (p / scale) * p + (bcmax / scale) * bcmis;
    if (z >= small) {
      z = p + sqrt(scale) * sqrt(z) * sgn(p);
      a = d + z;
      d = d - (bcmax / z) * bcmis;
      const double tau = jl_hypot(c, z);
      cs = z / tau;
      sn = c / tau;
      b = b - c;
      c = 0.0;
    } else {
      const double sigma = b + c;
      double tau = jl_hypot(sigma, asubd);
      cs = sqrt(half * (1.0 + fabs(sigma) / tau));
      sn = -(p / (tau * cs)) * sgn(sigma);
      const double aa = a * cs + b * sn;
      const double bb = -a * sn + b * cs;
      const double cc = c * cs + d * sn;
      const double dd = -c * sn + d * cs;
      a = aa * cs + cc * sn;
      b = bb * cs + dd * sn;
      c = -aa * sn + cc * cs;
      d = -bb * sn + dd * cs;
      const double midad = half * (a + d);
      a = midad;
      d = a;
      if (c != 0.0) {
        if (b != 0.0) {
          if (b * c >= 0.0) {
            const double sab = sqrt(fabs(b));
            const double sac = sqrt(fabs(c));
            p = sab * sac * sgn(c);
            tau = 1.0 / sqrt(fabs(b + c));
            a = midad + p;
            d = midad - p;
            b = b - c;
            c = 0.0;
            const double cs1 = sab * tau;
            const double sn1 = sac * tau;
            const double csn = cs * cs1 - sn * sn1;
            const double snn = cs * sn1 + sn * cs1;
            cs = csn;
            sn = snn;
          }
        } else {
          b = -c;
          c = 0.0;
          const double tmp = cs;
          cs = -sn;
          sn = tmp;
        }
      }
    }
  }
  if (c == 0.0) {
    w1 = zr(a);
    w2 = zr(d);
  } else {
    const double rti = sqrt(fabs(b)) * sqrt(fabs(c));
    w1 = zmk(a, rti);
    w2 = zmk(d, -rti);
  }
}

}  // namespace gs
