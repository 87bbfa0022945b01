// Deferred application of a double-shift sweep's logged reflectors to lines
// (rows or columns) of H and Z, shared by the windowed batched QR kernel and
// the large-matrix path.
#pragma once
#include "dev_common.h"

namespace gs {

// The logged reflectors of bulge steps ka .. kb (log entry k - mx: tau1, v2,
// v3; kb - ka < MAXS) applied in order to one line x(t) = p[(t - 1) * st] of
// H or Z: the right update of rows (GenericSchur.jl:916-935, 943-948) or the
// left update of columns (:900-915), which have the same form.  A step
// k = iend - 1 has a two-element reflector.  The line's segment moves
// through registers in groups of G elements, two groups in flight: a group's
// loads are issued once the group two places before it has been stored, so a
// chain waits on memory about once per group, not once per element.
template <int MAXS, int G>
GS_DEV void chain_line(double* p, int64_t st, const double* lg, int mx, int ka, int kb, int iend) {
  constexpr int NE = MAXS + 2, NG = (NE + G - 1) / G;
  // uniform counts (SCC branches)
  const int s = __builtin_amdgcn_readfirstlane(kb - ka + 1);
  const int k3 = __builtin_amdgcn_readfirstlane(min(kb, iend - 2));  // last step with a three-element reflector
  ka = __builtin_amdgcn_readfirstlane(ka);
  const int ne = (k3 == ka + s - 1) ? s + 2 : s + 1;                   // elements x(ka) .. x(ka + ne - 1)
  // log offset through an opaque (volatile asm) SGPR: the log is re-read per
  // line (LDS broadcasts, or cached loads) rather than every step's
  // reflector being held in registers across the callers' loops over lines
  int lb = 3 * (ka - __builtin_amdgcn_readfirstlane(mx));
  asm volatile("" : "+s"(lb));
  const double* lv = lg + lb;
  double* q0 = p + (int64_t)(ka - 1) * st;
  double x[NG * G];
#pragma unroll
  for (int g = 0; g < 2 && g < NG; ++g)
    if (g * G < ne) {
#pragma unroll
      for (int i = 0; i < G; ++i) x[g * G + i] = q0[(int64_t)min(g * G + i, ne - 1) * st];
    }
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    if (g * G >= ne) continue;
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int q = g * G + i;
      if (q < s) {
        const double tau1 = lv[3 * q], v2 = lv[3 * q + 1];
        const double tau2 = tau1 * v2;
        if (q <= k3 - ka) {
          const double v3 = lv[3 * q + 2];
          const double tau3 = tau1 * v3;
          const double ss = x[q] + v2 * x[q + 1] + v3 * x[q + 2];
          x[q] = x[q] - ss * tau1;
          x[q + 1] = x[q + 1] - ss * tau2;
          x[q + 2] = x[q + 2] - ss * tau3;
        } else {
          const double ss = x[q] + v2 * x[q + 1];
          x[q] = x[q] - ss * tau1;
          x[q + 1] = x[q + 1] - ss * tau2;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int q = g * G + i;
      if (q < ne) q0[(int64_t)q * st] = x[q];
    }
    __builtin_amdgcn_sched_barrier(0);
    if (g + 2 < NG && (g + 2) * G < ne) {
#pragma unroll
      for (int i = 0; i < G; ++i) x[(g + 2) * G + i] = q0[(int64_t)min((g + 2) * G + i, ne - 1) * st];
    }
  }
}

}  // namespace gs
