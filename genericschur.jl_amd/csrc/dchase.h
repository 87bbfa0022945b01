// The bulge chase of one chunk of a real double-shift sweep
// (GenericSchur.jl:877-949 restricted to an LDS window), shared by the
// batched windowed QR kernel and the large path's chunk kernel.  One
// wavefront; the window holds H(wb:we, wb:we) column-major with pitch WP.
//
// Step k in lanes:
//   row pass     lane c: column k + c, rows k .. k + nr - 1 (left update);
//   column pass  lanes 0..3: rows k .. k + 3, lanes 4 and up: rows
//                wb .. k - 1, columns k .. k + nr - 1 (right update).
// Every operand except the nr x nr block at rows/columns k .. k + nr - 1 is
// final before the step's reflector, so all loads are issued before it;
// the block's rows come to the column pass from the row pass's lanes by
// quad permutations, and the next reflector's inputs (column k, rows
// k + 1 .. k + 3) from the column pass's lanes 1..3 by v_readlane, so no
// LDS round trip lies between one reflector and the next.  The annihilated
// column's (beta, 0, 0) are stored once per chunk (no later step of the
// chunk reads them).  Each element receives the reference's updates with
// the same expressions, in the same order, as the one-lane-per-column form.
// The sweep's first step and its last (two-element) step run in the general
// form; the steps between run a loop specialised to three-element
// reflectors, whose reflector takes the wave-uniform form (unscaled sum of
// squares, one scalar branch to the reference formulas).
#pragma once
#include "dev_common.h"

namespace gs {

// lg receives (tau1, v2, v3) per step at 3 (k - k0).  v0, v1, v2s are the
// first step's reflector input when k0 == mx (the sweep's first column);
// otherwise the chunk's first input is read from the window.
// One step of the chunk in its general form (any reflector length, the
// sweep's first column included): the sweep's first and last steps.
template <int WP, bool FAST>
GS_DEV void dchase_step(double* win, int wb, int we, int k0, int k, int mx, int istart, int iend, double& v0,
                        double& v1, double& v2s, double& nv0, double& nv1, double& nv2, double& betal, double* lg) {
  const int lane = threadIdx.x & 63;
  auto W = [&](int r, int c) -> double* { return win + (r - wb) + (c - wb) * WP; };
  const bool three = iend - k + 1 >= 3;
  const int nr = three ? 3 : 2;
  // operands that are final before this step's reflector; the third
  // entries are loaded in any case (from row k when nr = 2, and then
  // unused), so that no load sits under a branch
  const bool rv = k + lane <= we;
  double* const rp = W(k, rv ? k + lane : k);
  double p0 = rp[0], p1 = rp[1], p2 = rp[three ? 2 : 0];
  const int jm = three ? min(k + 3, iend) : iend;
  const int crow = lane < 4 ? k + lane : wb + lane - 4;
  const bool cv = lane < 4 ? (crow <= jm) : (crow <= k - 1);
  double* const cp = W(cv ? crow : k, k);
  double c0 = cp[0], c1 = cp[WP], c2 = cp[three ? 2 * WP : 0];
  if (k > mx) {
    v0 = nv0;
    v1 = nv1;
    v2s = three ? nv2 : 0.0;
  }
  double x1 = v1, x2 = three ? v2s : 0.0;
  const double tau1 = FAST ? dreflector_fast(nr, v0, x1, x2) : dreflector(nr, v0, x1, x2);
  const double vv2 = x1;
  const double tau2 = tau1 * vv2;
  const double vv3 = three ? x2 : 0.0;
  const double tau3 = tau1 * vv3;
  if (k == mx && mx > istart && lane == 0) *W(k, k - 1) = *W(k, k - 1) * (1.0 - tau1);
  // left update of rows k .. k + nr - 1
  if (three) {
    const double ss = p0 + vv2 * p1 + vv3 * p2;
    p0 = p0 - ss * tau1;
    p1 = p1 - ss * tau2;
    p2 = p2 - ss * tau3;
  } else {
    const double ss = p0 + vv2 * p1;
    p0 = p0 - ss * tau1;
    p1 = p1 - ss * tau2;
  }
  // the nr x nr block: lane r (row k + r) takes the row pass's lane c
  // (column k + c) entry r
  {
    const double b01 = qperm<0x00>(p1), b02 = qperm<0x00>(p2);
    const double b10 = qperm<0x55>(p0), b12 = qperm<0x55>(p2);
    const double b20 = qperm<0xAA>(p0), b21 = qperm<0xAA>(p1);
    const bool l0 = lane == 0, l1 = lane == 1, l2 = lane == 2 && three;
    c0 = l0 ? p0 : l1 ? b01 : l2 ? b02 : c0;
    c1 = l0 ? b10 : l1 ? p1 : l2 ? b12 : c1;
    c2 = l0 ? b20 : l1 ? b21 : l2 ? p2 : c2;
  }
  // right update of columns k .. k + nr - 1
  if (three) {
    const double ss = c0 + vv2 * c1 + vv3 * c2;
    c0 = c0 - ss * tau1;
    c1 = c1 - ss * tau2;
    c2 = c2 - ss * tau3;
  } else {
    const double ss = c0 + vv2 * c1;
    c0 = c0 - ss * tau1;
    c1 = c1 - ss * tau2;
  }
  nv0 = rdlane(c0, 1);
  nv1 = rdlane(c0, 2);
  nv2 = rdlane(c0, 3);
  // the block's final values come from the column pass
  if (rv && lane >= nr) {
    rp[0] = p0;
    rp[1] = p1;
    if (three) rp[2] = p2;
  }
  if (cv) {
    cp[0] = c0;
    cp[WP] = c1;
    if (three) cp[2 * WP] = c2;
  }
  if (lane == 0) {
    lg[3 * (k - k0)] = tau1;
    lg[3 * (k - k0) + 1] = vv2;
    lg[3 * (k - k0) + 2] = vv3;
  }
  if (k > mx && lane == k - k0) betal = v0;
  wsync();
}

// lg receives (tau1, v2, v3) per step at 3 (k - k0).  v0, v1, v2s are the
// first step's reflector input when k0 == mx (the sweep's first column);
// otherwise the chunk's first input is read from the window.
template <int WP, bool FAST>
GS_DEV void dchase_chunk(double* win, int wb, int we, int k0, int kend, int mx, int istart, int iend, double v0,
                         double v1, double v2s, double* lg) {
  const int lane = threadIdx.x & 63;
  auto W = [&](int r, int c) -> double* { return win + (r - wb) + (c - wb) * WP; };
  // the chunk's first reflector input when it continues a sweep (the bulge
  // left in column k0 - 1 by the previous chunk), read before the loop so
  // that every step takes its input from registers and no step waits for
  // its own operand loads before the reflector
  double nv0 = 0.0, nv1 = 0.0, nv2 = 0.0;
  if (k0 > mx && k0 <= kend) {
    nv0 = *W(k0, k0 - 1);
    nv1 = *W(k0 + 1, k0 - 1);
    nv2 = (iend - k0 + 1 >= 3) ? *W(k0 + 2, k0 - 1) : 0.0;
  }
  // every memory operation issued before the loop has completed, so the
  // loop's waits follow its own loads only
  __builtin_amdgcn_s_waitcnt(0);
  double betal = 0.0;  // lane s: beta of step k0 + s
  int k = k0;
  // the sweep's first step
  if (k == mx && k <= kend) {
    dchase_step<WP, FAST>(win, wb, we, k0, k, mx, istart, iend, v0, v1, v2s, nv0, nv1, nv2, betal, lg);
    ++k;
  }
  // steps k > mx with three-element reflectors (k <= iend - 2): the block
  // rows k .. k + 2, the column pass rows k .. k + 3 (k + 3 <= iend when
  // k <= iend - 3) and wb .. k - 1
  const int k3 = min(kend, iend - 2);
  for (; k <= k3; ++k) {
    const int rc = min(lane, we - k);
    double* const rp = W(k, k + rc);
    const double q0 = rp[0], q1 = rp[1], q2 = rp[2];
    const int jm = min(k + 3, iend);
    const bool cv = lane < 4 ? (k + lane <= jm) : (wb + lane - 4 <= k - 1);
    double* const cp = W(cv ? (lane < 4 ? k + lane : wb + lane - 4) : k, k);
    double c0 = cp[0], c1 = cp[WP], c2 = cp[2 * WP];
    double a0 = nv0, x1 = nv1, x2 = nv2;
    const double tau1 = FAST ? dreflector3_wave(a0, x1, x2) : dreflector(3, a0, x1, x2);
    const double vv2 = x1, vv3 = x2;
    const double tau2 = tau1 * vv2, tau3 = tau1 * vv3;
    const double ss = q0 + vv2 * q1 + vv3 * q2;
    const double p0 = q0 - ss * tau1, p1 = q1 - ss * tau2, p2 = q2 - ss * tau3;
    {
      const double b01 = qperm<0x00>(p1), b02 = qperm<0x00>(p2);
      const double b10 = qperm<0x55>(p0), b12 = qperm<0x55>(p2);
      const double b20 = qperm<0xAA>(p0), b21 = qperm<0xAA>(p1);
      const bool l0 = lane == 0, l1 = lane == 1, l2 = lane == 2;
      c0 = l0 ? p0 : l1 ? b01 : l2 ? b02 : c0;
      c1 = l0 ? b10 : l1 ? p1 : l2 ? b12 : c1;
      c2 = l0 ? b20 : l1 ? b21 : l2 ? p2 : c2;
    }
    const double st = c0 + vv2 * c1 + vv3 * c2;
    c0 = c0 - st * tau1;
    c1 = c1 - st * tau2;
    c2 = c2 - st * tau3;
    nv0 = rdlane(c0, 1);
    nv1 = rdlane(c0, 2);
    nv2 = rdlane(c0, 3);
    if (lane >= 3 && lane == rc) {
      rp[0] = p0;
      rp[1] = p1;
      rp[2] = p2;
    }
    if (cv) {
      cp[0] = c0;
      cp[WP] = c1;
      cp[2 * WP] = c2;
    }
    if (lane == 0) {
      lg[3 * (k - k0)] = tau1;
      lg[3 * (k - k0) + 1] = vv2;
      lg[3 * (k - k0) + 2] = vv3;
    }
    if (lane == k - k0) betal = a0;
    wsync();
  }
  // the sweep's last step (two-element reflector at k = iend - 1)
  if (k <= kend) dchase_step<WP, FAST>(win, wb, we, k0, k, mx, istart, iend, v0, v1, v2s, nv0, nv1, nv2, betal, lg);
  // the annihilated columns: (k, k - 1) = beta, (k + 1, k - 1) = 0 and
  // (k + 2, k - 1) = 0
  {
    const int kk = k0 + lane;
    if (kk <= kend && kk > mx) {
      *W(kk, kk - 1) = betal;
      *W(kk + 1, kk - 1) = 0.0;
      if (kk < iend - 1) *W(kk + 2, kk - 1) = 0.0;
    }
    wsync();
  }
}

}  // namespace gs
