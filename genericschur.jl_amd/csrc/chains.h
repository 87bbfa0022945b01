// Deferred application of a double-shift sweep's logged reflectors to lines
// (rows or columns) of H and Z, shared by the windowed batched QR kernel and
// the large-matrix path.
#pragma once
#include "dev_common.h"

namespace gs {

// The logged reflectors of bulge steps ka .. kb (log entry k - mx: tau1, v2,
// v3) applied in order to one line x(t) = p[(t - 1) * st] of H or Z: the
// right update of rows (GenericSchur.jl:916-935, 943-948) or the left
// update of columns (:900-915), which have the same form.  The line's
// values x(k), x(k+1) ride in registers; x(k+2) is read PF steps ahead.
GS_DEV void chain_line(double* p, int64_t st, const double* lg, int mx, int ka, int kb, int iend) {
  double c0 = p[(int64_t)(ka - 1) * st], c1 = p[(int64_t)ka * st];
  const int k3 = min(kb, iend - 2);  // steps with three-element reflectors
  constexpr int PF = 8;
  double nb[PF];
#pragma unroll
  for (int q = 0; q < PF; ++q)
    if (ka + q <= k3) nb[q] = p[(int64_t)(ka + q + 1) * st];
  int k = ka;
  for (; k + PF - 1 <= k3; k += PF) {
    double cur[PF];
#pragma unroll
    for (int q = 0; q < PF; ++q) cur[q] = nb[q];
#pragma unroll
    for (int q = 0; q < PF; ++q)
      if (k + PF + q <= k3) nb[q] = p[(int64_t)(k + PF + q + 1) * st];
#pragma unroll
    for (int q = 0; q < PF; ++q) {
      const double* e = lg + 3 * (k + q - mx);
      const double tau1 = e[0], v2 = e[1], v3 = e[2];
      const double tau2 = tau1 * v2, tau3 = tau1 * v3;
      const double b = cur[q];
      const double ss = c0 + v2 * c1 + v3 * b;
      p[(int64_t)(k + q - 1) * st] = c0 - ss * tau1;
      c0 = c1 - ss * tau2;
      c1 = b - ss * tau3;
    }
  }
  for (; k <= k3; ++k) {
    const double* e = lg + 3 * (k - mx);
    const double tau1 = e[0], v2 = e[1], v3 = e[2];
    const double tau2 = tau1 * v2, tau3 = tau1 * v3;
    const double b = p[(int64_t)(k + 1) * st];
    const double ss = c0 + v2 * c1 + v3 * b;
    p[(int64_t)(k - 1) * st] = c0 - ss * tau1;
    c0 = c1 - ss * tau2;
    c1 = b - ss * tau3;
  }
  if (k <= kb) {  // k == iend - 1: two-element reflector
    const double* e = lg + 3 * (k - mx);
    const double tau1 = e[0], v2 = e[1];
    const double tau2 = tau1 * v2;
    const double ss = c0 + v2 * c1;
    p[(int64_t)(k - 1) * st] = c0 - ss * tau1;
    c0 = c1 - ss * tau2;
    p[(int64_t)k * st] = c0;
    return;
  }
  p[(int64_t)kb * st] = c0;
  p[(int64_t)(kb + 1) * st] = c1;
}

}  // namespace gs
