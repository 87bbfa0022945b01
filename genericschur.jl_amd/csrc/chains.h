// Deferred application of a double-shift sweep's logged reflectors to lines
// (rows or columns) of H and Z, shared by the windowed batched QR kernel and
// the large-matrix path.
#pragma once
#include "dev_common.h"

namespace gs {

// A step of a logged real reflector on a line's entries, x -= (tau1 ss)
// (1, v2, v3) with ss = x0 + v2 x1 + v3 x2 (GenericSchur.jl:900-948), its
// two-element form, and a Givens rotation of two entries (:674-687), with a
// fixed rounding: no contraction beyond the fma calls written here.  Every
// path that applies them to Z (in place during the sweeps, or replayed from
// the Z log) goes through these, so the paths give the same bits.
GS_DEV void drefl3_fixed(double& x0, double& x1, double& x2, double tau1, double v2, double v3) {
#pragma clang fp contract(off)
  const double tau2 = tau1 * v2, tau3 = tau1 * v3;
  const double ss = fma(v3, x2, fma(v2, x1, x0));
  x0 = fma(-ss, tau1, x0);
  x1 = fma(-ss, tau2, x1);
  x2 = fma(-ss, tau3, x2);
}
GS_DEV void drefl2_fixed(double& x0, double& x1, double tau1, double v2) {
#pragma clang fp contract(off)
  const double tau2 = tau1 * v2;
  const double ss = fma(v2, x1, x0);
  x0 = fma(-ss, tau1, x0);
  x1 = fma(-ss, tau2, x1);
}
GS_DEV void drot_fixed(double& a1, double& a2, double cs, double sn) {
#pragma clang fp contract(off)
  const double b1 = fma(a1, cs, a2 * sn);
  const double b2 = fma(-a1, sn, a2 * cs);
  a1 = b1;
  a2 = b2;
}

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
        if (q <= k3 - ka)
          drefl3_fixed(x[q], x[q + 1], x[q + 2], tau1, v2, lv[3 * q + 2]);
        else
          drefl2_fixed(x[q], x[q + 1], tau1, v2);
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

// chain_line for a full chunk of three-element reflectors (kb = ka + MAXS - 1,
// kb <= iend - 2): every count and index is a compile-time constant, so the
// compiler waits on memory once per group rather than before every store.
// Same operations in the same order as chain_line.
template <int MAXS, int G>
GS_DEV void chain_line_full(double* p, int64_t st, const double* lg, int mx, int ka) {
  constexpr int S = MAXS, NE = MAXS + 2, NG = (NE + G - 1) / G;
  ka = __builtin_amdgcn_readfirstlane(ka);
  int lb = 3 * (ka - __builtin_amdgcn_readfirstlane(mx));
  asm volatile("" : "+s"(lb));
  const double* lv = lg + lb;
  double* q0 = p + (int64_t)(ka - 1) * st;
  double x[NG * G];
#pragma unroll
  for (int e = 0; e < 2 * G; ++e)
    if (e < NE) x[e] = q0[(int64_t)e * st];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int q = g * G + i;
      if (q < S) drefl3_fixed(x[q], x[q + 1], x[q + 2], lv[3 * q], lv[3 * q + 1], lv[3 * q + 2]);
    }
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int q = g * G + i;
      if (q < NE) q0[(int64_t)q * st] = x[q];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < G; ++i) {
      const int e = (g + 2) * G + i;
      if (e < NE) x[e] = q0[(int64_t)e * st];
    }
  }
}

// Real Z log (the windowed batched QR kernel, bins n = 65 .. 128): records
// of doubles, a sweep (mx, iend) then (tau1, v2, v3) per step k = mx ..
// iend-1 (the last step's reflector has two elements), or a Givens rotation
// (-j, cs, sn) of columns j, j+1 from a 2 x 2 block's standardization.
// Applies records [0, len) to one row of Z, x(c) = row[(c - 1) st], with the
// operations and order of the in-place updates.
GS_DEV void dlog_row(double* row, int64_t st, const double* __restrict__ lg, int64_t len) {
  int64_t pos = 0;
  while (pos < len) {
    const double h0 = lg[pos];
    if (h0 < 0.0) {
      const int j = (int)(-h0);
      double a1 = row[(int64_t)(j - 1) * st], a2 = row[(int64_t)j * st];
      drot_fixed(a1, a2, lg[pos + 1], lg[pos + 2]);
      row[(int64_t)(j - 1) * st] = a1;
      row[(int64_t)j * st] = a2;
      pos += 3;
    } else {
      const int mx = (int)h0, ie = (int)lg[pos + 1];
      const double* e = lg + pos + 2;
      double c0 = row[(int64_t)(mx - 1) * st], c1 = row[(int64_t)mx * st];
      for (int k = mx; k < ie - 1; ++k) {
        double b = row[(int64_t)(k + 1) * st];
        drefl3_fixed(c0, c1, b, e[3 * (k - mx)], e[3 * (k - mx) + 1], e[3 * (k - mx) + 2]);
        row[(int64_t)(k - 1) * st] = c0;
        c0 = c1;
        c1 = b;
      }
      drefl2_fixed(c0, c1, e[3 * (ie - 1 - mx)], e[3 * (ie - 1 - mx) + 1]);
      row[(int64_t)(ie - 2) * st] = c0;
      row[(int64_t)(ie - 1) * st] = c1;
      pos += 2 + 3 * (int64_t)(ie - mx);
    }
  }
}

}  // namespace gs
