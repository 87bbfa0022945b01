// Reordering of a real Schur form on gfx950: gordschur!(T, Z, select) for
// real T (rordschur.jl:3-45) with _trexchange! (:54-165, LAPACK xTREXC),
// _swap1or2 (:168-297, xLAEXC) and _syl1or2 (:300-466, xLASY2).
//
// One wavefront per matrix, T and Z in global memory.  The block logic, the
// small Sylvester solve, the 3-vector reflectors and the trial transformation
// of the copied diagonal block D are scalar: every lane evaluates them on the
// same data (so every branch is uniform).  The updates of T's rows and
// columns and of Z's columns are spread over the lanes, one row or column
// per lane, with a wavefront barrier after each.  Everything is evaluated
// without contraction and in the CPU restatement's operation order (the
// reference's BLAS calls on length-2 and -3 vectors in the reference BLAS
// order), so T, Z and the eigenvalues match it bit for bit.
#include <hip/hip_runtime.h>

#pragma clang fp contract(off)

#include "dev_common.h"
#include "gs2x2.h"
#include "internal.h"
#include "steqr.h"

namespace gs {

namespace {

// Julia's NaN-propagating max
GS_DEV double jmaxd(double a, double b) {
  if (a != a || b != b) return a + b;
  return (a < b) ? b : a;
}

// LinearAlgebra.generic_norm2 of a short vector
GS_DEV double norm2_short(int n, const double* x) {
  double maxabs = fabs(x[0]);
  for (int i = 1; i < n; ++i) maxabs = jmaxd(maxabs, fabs(x[i]));
  const bool inf = !(maxabs != maxabs) && !(maxabs <= GS_FMAX);
  if (maxabs == 0.0 || inf) return maxabs;
  const double t = ((double)n * maxabs) * maxabs;
  const bool finite = !(t != t) && t <= GS_FMAX;
  if (finite && !(maxabs * maxabs == 0.0)) {
    double sum = x[0] * x[0];
    for (int i = 1; i < n; ++i) sum = sum + x[i] * x[i];
    return sqrt(sum);
  }
  double y = fabs(x[0]) / maxabs;
  double sum = y * y;
  for (int i = 1; i < n; ++i) {
    y = fabs(x[i]) / maxabs;
    sum = sum + y * y;
  }
  return maxabs * sqrt(sum);
}

// LinearAlgebra.reflector! (stdlib) on a 3-vector
GS_DEV double reflector3(double* x) {
  double xi1 = x[0];
  const double normu = norm2_short(3, x);
  if (normu == 0.0) return 0.0;
  const double nu = copysign(normu, xi1);
  xi1 = xi1 + nu;
  x[0] = -nu;
  x[1] = x[1] / xi1;
  x[2] = x[2] / xi1;
  return xi1 / nu;
}

// _syl1or2 (xLASY2, isgn = -1): Tl X - X Tr = scale B; 2 x 2 column-major
// arrays (element (i, j) at [i + 2 j], 0-based).  Returns scale.
GS_DEV double syl1or2(int n1, int n2, const double* tl, const double* tr, const double* b, double* x) {
  const double epst = GS_EPS;
  const double smallnum = GS_FMIN / epst;
  double scale = 1.0;
  if (n1 == 2 && n2 == 2) {
    double smin = jmaxd(jmaxd(jmaxd(fabs(tr[0]), fabs(tr[1])), jmaxd(fabs(tr[2]), fabs(tr[3]))),
                        jmaxd(jmaxd(fabs(tl[0]), fabs(tl[1])), jmaxd(fabs(tl[2]), fabs(tl[3]))));
    smin = jmaxd(epst * smin, smallnum);
    double TT[4][4];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) TT[i][j] = 0.0;
    TT[0][0] = tl[0] - tr[0];
    TT[1][1] = tl[3] - tr[0];
    TT[2][2] = tl[0] - tr[3];
    TT[3][3] = tl[3] - tr[3];
    TT[0][1] = tl[2];
    TT[1][0] = tl[1];
    TT[2][3] = tl[2];
    TT[3][2] = tl[1];
    TT[0][2] = -tr[1];
    TT[1][3] = -tr[1];
    TT[2][0] = -tr[2];
    TT[3][1] = -tr[2];
    double bv[4] = {b[0], b[1], b[2], b[3]};
    int ipsv = 0, jpsv = 0;
    int jpiv[4] = {0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
      double xmax = 0.0;
      for (int ip = i; ip < 4; ++ip)
        for (int jp = i; jp < 4; ++jp)
          if (fabs(TT[ip][jp]) > xmax) {
            xmax = fabs(TT[ip][jp]);
            ipsv = ip;
            jpsv = jp;
          }
      if (ipsv != i) {
        for (int k = 0; k < 4; ++k) {
          const double t = TT[ipsv][k];
          TT[ipsv][k] = TT[i][k];
          TT[i][k] = t;
        }
        const double t = bv[i];
        bv[i] = bv[ipsv];
        bv[ipsv] = t;
      }
      if (jpsv != i)
        for (int k = 0; k < 4; ++k) {
          const double t = TT[k][jpsv];
          TT[k][jpsv] = TT[k][i];
          TT[k][i] = t;
        }
      jpiv[i] = jpsv;
      if (fabs(TT[i][i]) < smin) TT[i][i] = smin;
      for (int j = i + 1; j < 4; ++j) {
        TT[j][i] = TT[j][i] / TT[i][i];
        bv[j] = bv[j] - TT[j][i] * bv[i];
        for (int k = i + 1; k < 4; ++k) TT[j][k] = TT[j][k] - TT[j][i] * TT[i][k];
      }
    }
    if (fabs(TT[3][3]) < smin) TT[3][3] = smin;
    const double a = 8.0 * smallnum;
    if ((a * fabs(bv[0]) > TT[0][0]) || (a * fabs(bv[1]) > TT[1][1]) || (a * fabs(bv[2]) > TT[2][2]) ||
        (a * fabs(bv[3]) > TT[3][3])) {
      const double bmax = jmaxd(jmaxd(fabs(bv[0]), fabs(bv[1])), jmaxd(fabs(bv[2]), fabs(bv[3])));
      scale = (1.0 / 8.0) / bmax;
      for (int k = 0; k < 4; ++k) bv[k] = bv[k] * scale;
    }
    double xt[4];
    for (int i = 1; i <= 4; ++i) {
      const int k = 4 - i;
      const double t = 1.0 / TT[k][k];
      xt[k] = bv[k] * t;
      for (int j = k + 1; j < 4; ++j) xt[k] = xt[k] - (t * TT[k][j]) * xt[j];
    }
    for (int i = 1; i <= 3; ++i) {
      const int jj = jpiv[3 - i];
      if (jj != 3 - i) {
        const double t = xt[3 - i];
        xt[3 - i] = xt[jj];
        xt[jj] = t;
      }
    }
    for (int k = 0; k < 4; ++k) x[k] = xt[k];
    return scale;
  }
  double smin, tmp[4], b1, b2;
  if (n1 == 1) {
    smin = jmaxd(epst * jmaxd(fabs(tl[0]), jmaxd(jmaxd(fabs(tr[0]), fabs(tr[1])), jmaxd(fabs(tr[2]), fabs(tr[3])))),
                 smallnum);
    tmp[0] = tl[0] - tr[0];
    tmp[1] = -tr[2];
    tmp[2] = -tr[1];
    tmp[3] = tl[0] - tr[3];
    b1 = b[0];
    b2 = b[2];
  } else {
    smin = jmaxd(epst * jmaxd(fabs(tr[0]), jmaxd(jmaxd(fabs(tl[0]), fabs(tl[1])), jmaxd(fabs(tl[2]), fabs(tl[3])))),
                 smallnum);
    tmp[0] = tl[0] - tr[0];
    tmp[1] = tl[1];
    tmp[2] = tl[2];
    tmp[3] = tl[3] - tr[0];
    b1 = b[0];
    b2 = b[1];
  }
  int ipiv = 0;
  double amax = fabs(tmp[0]);
  for (int k = 1; k < 4; ++k)
    if (fabs(tmp[k]) > amax) {
      amax = fabs(tmp[k]);
      ipiv = k;
    }
  double u11 = tmp[ipiv];
  if (amax <= smin) u11 = smin;
  // _syl2x2_inds / _syl2x2_swaps (rordschur.jl:467-483), 0-based
  const int iu12 = (ipiv == 0) ? 2 : (ipiv == 1) ? 3 : (ipiv == 2) ? 0 : 1;
  const int il21 = (ipiv == 0) ? 1 : (ipiv == 1) ? 0 : (ipiv == 2) ? 3 : 2;
  const int iu22 = (ipiv == 0) ? 3 : (ipiv == 1) ? 2 : (ipiv == 2) ? 1 : 0;
  const bool xswap = ipiv >= 2, bswap = (ipiv & 1) != 0;
  const double u12 = tmp[iu12];
  const double l21 = tmp[il21] / u11;
  double u22 = tmp[iu22] - u12 * l21;
  if (fabs(u22) <= smin) u22 = smin;
  if (bswap) {
    const double t = b1;
    b1 = b2;
    b2 = t - l21 * b2;
  } else {
    b2 = b2 - l21 * b1;
  }
  if (((2.0 * smallnum) * fabs(b2) > fabs(u22)) || ((2.0 * smallnum) * fabs(b1) > fabs(u11))) {
    scale = (1.0 / 2.0) / jmaxd(fabs(b1), fabs(b2));
    b1 = scale * b1;
    b2 = scale * b2;
  }
  const double x22 = b2 / u22;
  double x1 = b1 / u11 - (u12 / u11) * x22, x2 = x22;
  if (xswap) {
    const double t = x1;
    x1 = x2;
    x2 = t;
  }
  x[0] = x1;
  if (n1 == 1) x[2] = x2;
  else x[1] = x2;
  return scale;
}

// 1-based accessor of a column-major matrix
struct RMat {
  double* p;
  int64_t ld;
  GS_DEV double& operator()(int i, int j) const { return p[(i - 1) + (int64_t)(j - 1) * ld]; }
};

// The row / column transformations of the swaps; PAR: one row or column per
// lane (T, Z in global memory), otherwise all of it in this lane (the
// private copy D).  Operation order as the CPU restatement.
template <bool PAR> GS_DEV int l0() { return PAR ? (int)threadIdx.x : 0; }
template <bool PAR> constexpr int lstep() { return PAR ? 64 : 1; }

template <bool PAR, class M> GS_DEV void g2_lmul(M A, int i, int c0, int c1, double c, double s) {
  for (int j = c0 + l0<PAR>(); j <= c1; j += lstep<PAR>()) {
    const double a1 = A(i, j), a2 = A(i + 1, j);
    A(i, j) = c * a1 + s * a2;
    A(i + 1, j) = -s * a1 + c * a2;
  }
}
template <bool PAR, class M> GS_DEV void g2_rmul(M A, int r0, int r1, int i, double c, double s) {
  for (int r = r0 + l0<PAR>(); r <= r1; r += lstep<PAR>()) {
    const double a1 = A(r, i), a2 = A(r, i + 1);
    A(r, i) = a1 * c + a2 * s;
    A(r, i + 1) = -a1 * s + a2 * c;
  }
}
template <bool PAR, class M> GS_DEV void hh3_lmul(M A, int r0, int c0, int c1, double v1, double v2, double tau) {
  for (int j = c0 + l0<PAR>(); j <= c1; j += lstep<PAR>()) {
    double va = A(r0, j);
    va = va + (v1 * A(r0 + 1, j) + v2 * A(r0 + 2, j));
    va = tau * va;
    A(r0, j) = A(r0, j) - va;
    A(r0 + 1, j) = A(r0 + 1, j) + (-va) * v1;
    A(r0 + 2, j) = A(r0 + 2, j) + (-va) * v2;
  }
}
template <bool PAR, class M> GS_DEV void hh3_rmul(M A, int r0, int r1, int c0, double v1, double v2, double tau) {
  for (int i = r0 + l0<PAR>(); i <= r1; i += lstep<PAR>()) {
    double x = A(i, c0 + 1) * v1 + A(i, c0 + 2) * v2;
    x = x + A(i, c0);
    A(i, c0) = A(i, c0) + (-tau) * x;
    A(i, c0 + 1) = A(i, c0 + 1) + x * ((-tau) * v1);
    A(i, c0 + 2) = A(i, c0 + 2) + x * ((-tau) * v2);
  }
}
template <bool PAR, class M> GS_DEV void hy3_lmul(M A, int r0, int c0, int c1, double v1, double v2, double tau) {
  for (int j = c0 + l0<PAR>(); j <= c1; j += lstep<PAR>()) {
    double va = A(r0 + 2, j);
    va = va + v1 * A(r0, j);
    va = va + v2 * A(r0 + 1, j);
    va = tau * va;
    A(r0 + 2, j) = A(r0 + 2, j) - va;
    A(r0, j) = A(r0, j) - v1 * va;
    A(r0 + 1, j) = A(r0 + 1, j) - v2 * va;
  }
}
template <bool PAR, class M> GS_DEV void hy3_rmul(M A, int r0, int r1, int c0, double v1, double v2, double tau) {
  for (int i = r0 + l0<PAR>(); i <= r1; i += lstep<PAR>()) {
    double x = A(i, c0 + 2);
    x = x + A(i, c0) * v1;
    x = x + A(i, c0 + 1) * v2;
    x = tau * x;
    A(i, c0 + 2) = A(i, c0 + 2) - x;
    A(i, c0) = A(i, c0) - x * v1;
    A(i, c0 + 1) = A(i, c0 + 1) - x * v2;
  }
}

// the private 4 x 4 copy of the diagonal block (1-based)
struct DMat {
  double* p;
  GS_DEV double& operator()(int i, int j) const { return p[(i - 1) + 4 * (j - 1)]; }
};

// _swap1or2(T, Z, j1, n1, n2) (xLAEXC); uniform across the wavefront
GS_DEV bool swap1or2(int n, RMat T, RMat Z, bool wantz, int j1, int n1, int n2) {
  const bool l0w = threadIdx.x == 0;
  if (j1 + n1 > n) return true;
  const int j2 = j1 + 1, j3 = j1 + 2, j4 = j1 + 3;
  if (n1 == 1 && n2 == 1) {
    const double T11 = T(j1, j1), T22 = T(j2, j2);
    double c, s, r;
    givens_re(T(j1, j2), T22 - T11, c, s, r);
    __syncthreads();
    if (j3 <= n) g2_lmul<true>(T, j1, j3, n, c, s);
    g2_rmul<true>(T, 1, j1 - 1, j1, c, s);
    if (wantz) g2_rmul<true>(Z, 1, n, j1, c, s);
    __syncthreads();
    if (l0w) {
      T(j1, j1) = T22;
      T(j2, j2) = T11;
    }
    __syncthreads();
    return true;
  }
  const int nd = n1 + n2;
  double Dm[16];
  DMat D{Dm};
  double dnorm = 0.0;
  for (int j = 1; j <= nd; ++j)
    for (int i = 1; i <= nd; ++i) {
      D(i, j) = T(j1 + i - 1, j1 + j - 1);
      dnorm = (i == 1 && j == 1) ? fabs(D(i, j)) : jmaxd(dnorm, fabs(D(i, j)));
    }
  const double smallnum = GS_FMIN / GS_EPS;
  const double thresh = jmaxd(10.0 * GS_EPS * dnorm, smallnum);
  double tl[4] = {0.0, 0.0, 0.0, 0.0}, tr[4] = {0.0, 0.0, 0.0, 0.0}, bb[4] = {0.0, 0.0, 0.0, 0.0};
  double X[4] = {0.0, 0.0, 0.0, 0.0};
  for (int j = 0; j < n1; ++j)
    for (int i = 0; i < n1; ++i) tl[i + 2 * j] = D(i + 1, j + 1);
  for (int j = 0; j < n2; ++j)
    for (int i = 0; i < n2; ++i) tr[i + 2 * j] = D(n1 + i + 1, n1 + j + 1);
  for (int j = 0; j < n2; ++j)
    for (int i = 0; i < n1; ++i) bb[i + 2 * j] = D(i + 1, n1 + j + 1);
  const double scale = syl1or2(n1, n2, tl, tr, bb, X);
  if (n1 == 1 && n2 == 2) {
    double u[3] = {X[2], scale, X[0]};
    const double tau = reflector3(u);
    const double T11 = T(j1, j1);
    hy3_lmul<false>(D, 1, 1, 3, u[1], u[2], tau);
    hy3_rmul<false>(D, 1, 3, 1, u[1], u[2], tau);
    if (jmaxd(jmaxd(fabs(D(3, 1)), fabs(D(3, 2))), fabs(D(3, 3) - T11)) > thresh) return false;
    __syncthreads();
    hy3_lmul<true>(T, j1, j1, n, u[1], u[2], tau);
    __syncthreads();
    hy3_rmul<true>(T, 1, j2, j1, u[1], u[2], tau);
    if (wantz) hy3_rmul<true>(Z, 1, n, j1, u[1], u[2], tau);
    __syncthreads();
    if (l0w) {
      T(j3, j1) = 0.0;
      T(j3, j2) = 0.0;
      T(j3, j3) = T11;
    }
    __syncthreads();
  } else if (n1 == 2 && n2 == 1) {
    double u[3] = {-X[0], -X[1], scale};
    const double tau = reflector3(u);
    const double T33 = T(j3, j3);
    hh3_lmul<false>(D, 1, 1, 3, u[1], u[2], tau);
    hh3_rmul<false>(D, 1, 3, 1, u[1], u[2], tau);
    if (jmaxd(jmaxd(fabs(D(1, 1) - T33), fabs(D(2, 1))), fabs(D(3, 1))) > thresh) return false;
    __syncthreads();
    hh3_rmul<true>(T, 1, j3, j1, u[1], u[2], tau);
    __syncthreads();
    hh3_lmul<true>(T, j1, j2, n, u[1], u[2], tau);
    if (wantz) hh3_rmul<true>(Z, 1, n, j1, u[1], u[2], tau);
    __syncthreads();
    if (l0w) {
      T(j1, j1) = T33;
      T(j2, j1) = 0.0;
      T(j3, j1) = 0.0;
    }
    __syncthreads();
  } else {
    double u1[3] = {-X[0], -X[1], scale};
    const double tau1 = reflector3(u1);
    const double tt = -tau1 * (X[2] + u1[1] * X[3]);
    double u2[3] = {-tt * u1[1] - X[3], -tt * u1[2], scale};
    const double tau2 = reflector3(u2);
    hh3_lmul<false>(D, 1, 1, 4, u1[1], u1[2], tau1);
    hh3_rmul<false>(D, 1, 4, 1, u1[1], u1[2], tau1);
    hh3_lmul<false>(D, 2, 1, 4, u2[1], u2[2], tau2);
    hh3_rmul<false>(D, 1, 4, 2, u2[1], u2[2], tau2);
    if (jmaxd(jmaxd(fabs(D(3, 1)), fabs(D(3, 2))), jmaxd(fabs(D(4, 1)), fabs(D(4, 2)))) > thresh) return false;
    __syncthreads();
    hh3_lmul<true>(T, j1, j1, n, u1[1], u1[2], tau1);
    __syncthreads();
    hh3_rmul<true>(T, 1, j4, j1, u1[1], u1[2], tau1);
    __syncthreads();
    hh3_lmul<true>(T, j2, j1, n, u2[1], u2[2], tau2);
    __syncthreads();
    hh3_rmul<true>(T, 1, j4, j2, u2[1], u2[2], tau2);
    if (wantz) {
      hh3_rmul<true>(Z, 1, n, j1, u1[1], u1[2], tau1);
      __syncthreads();
      hh3_rmul<true>(Z, 1, n, j2, u2[1], u2[2], tau2);
    }
    __syncthreads();
    if (l0w) {
      T(j3, j1) = 0.0;
      T(j3, j2) = 0.0;
      T(j4, j1) = 0.0;
      T(j4, j2) = 0.0;
    }
    __syncthreads();
  }
  // standardize a new 2 x 2 block at k (_gs2x2!)
  auto standardize = [&](int k) {
    double a = T(k, k), b = T(k, k + 1), c = T(k + 1, k), d = T(k + 1, k + 1), cs, sn;
    zd w1, w2;
    gs2x2_dev(a, b, c, d, cs, sn, w1, w2);
    __syncthreads();
    g2_lmul<true>(T, k, k + 2, n, cs, sn);
    g2_rmul<true>(T, 1, k - 1, k, cs, sn);
    if (wantz) g2_rmul<true>(Z, 1, n, k, cs, sn);
    __syncthreads();
    if (l0w) {
      T(k, k) = a;
      T(k, k + 1) = b;
      T(k + 1, k) = c;
      T(k + 1, k + 1) = d;
    }
    __syncthreads();
  };
  if (n2 == 2) standardize(j1);
  if (n1 == 2) standardize(j1 + n2);
  return true;
}

// _trexchange!(T, Z, iold, inew) (xTREXC): false when a swap was refused
GS_DEV bool trexchange(int n, RMat T, RMat Z, bool wantz, int iold, int& inew) {
  if (n <= 1 || iold == inew) return true;
  if (iold > 1 && T(iold, iold - 1) != 0.0) iold -= 1;
  auto nbsize = [&](int j) { return (j < n && T(j + 1, j) != 0.0) ? 2 : 1; };
  int nbold = nbsize(iold);
  if (inew > 1 && T(inew, inew - 1) != 0.0) inew -= 1;
  const int nbnew = nbsize(inew);
  if (iold == inew) return true;
  int ihere = iold;
  bool ok = true;
  if (iold < inew) {
    if (nbold == 2 && nbnew == 1) inew -= 1;
    else if (nbold == 1 && nbnew == 2) inew += 1;
    while (ihere < inew) {
      if (nbold == 1 || nbold == 2) {
        const int nbnext = nbsize(ihere + nbold);
        ok = swap1or2(n, T, Z, wantz, ihere, nbold, nbnext);
        if (!ok) {
          inew = ihere;
          return false;
        }
        ihere += nbnext;
        if (nbold == 2 && T(ihere + 1, ihere) == 0.0) nbold = 3;
      } else {
        int nbnext = nbsize(ihere + 1);
        ok = swap1or2(n, T, Z, wantz, ihere + 1, 1, nbnext);
        if (!ok) {
          inew = ihere;
          return false;
        }
        if (nbnext == 1) {
          ok = swap1or2(n, T, Z, wantz, ihere, 1, nbnext);
          ihere += 1;
        } else {
          if (T(ihere + 2, ihere + 1) == 0.0) nbnext = 1;
          if (nbnext == 2) {
            ok = swap1or2(n, T, Z, wantz, ihere, 1, nbnext);
            if (!ok) {
              inew = ihere;
              return false;
            }
            ihere += 2;
          } else {
            ok = swap1or2(n, T, Z, wantz, ihere, 1, 1);
            ok = swap1or2(n, T, Z, wantz, ihere + 1, 1, 1);
            ihere += 2;
          }
        }
      }
    }
  } else {
    while (ihere > inew) {
      if (nbold == 1 || nbold == 2) {
        const int nbnext = (ihere >= 3 && T(ihere - 1, ihere - 2) != 0.0) ? 2 : 1;
        ok = swap1or2(n, T, Z, wantz, ihere - nbnext, nbnext, nbold);
        if (!ok) {
          inew = ihere;
          return false;
        }
        ihere -= nbnext;
        if (nbold == 2 && T(ihere + 1, ihere) == 0.0) nbold = 3;
      } else {
        int nbnext = (ihere >= 3 && T(ihere - 1, ihere - 2) != 0.0) ? 2 : 1;
        ok = swap1or2(n, T, Z, wantz, ihere - nbnext, nbnext, 1);
        if (!ok) {
          inew = ihere;
          return false;
        }
        if (nbnext == 1) {
          ok = swap1or2(n, T, Z, wantz, ihere, nbnext, 1);
          ihere -= 1;
        } else {
          if (T(ihere, ihere - 1) == 0.0) nbnext = 1;
          if (nbnext == 2) {
            ok = swap1or2(n, T, Z, wantz, ihere - 1, 2, 1);
            if (!ok) {
              inew = ihere;
              return false;
            }
            ihere -= 2;
          } else {
            ok = swap1or2(n, T, Z, wantz, ihere, 1, 1);
            ok = swap1or2(n, T, Z, wantz, ihere - 1, 1, 1);
            ihere -= 2;
          }
        }
      }
    }
  }
  inew = ihere;
  return ok;
}

}  // namespace

// gordschur!(T, Z, select) for real T: one wavefront; info[0] = 0 or the
// 1-based k of IllConditionException(k); W (n complex) the eigenvalues.
__global__ void __launch_bounds__(64) dordschur_kernel(int n, double* Tp, int64_t ldt, double* Zp, int64_t ldz,
                                                       int wantz, const int32_t* select, double* W, int32_t* info) {
  RMat T{Tp, ldt};
  RMat Z{Zp, ldz};
  int ks = 0;
  bool pair = false;
  int status = 0;
  for (int k = 1; k <= n; ++k) {
    if (pair) {
      pair = false;
      continue;
    }
    bool swap = select[k - 1] != 0;
    if (k < n && T(k + 1, k) != 0.0) {
      pair = true;
      swap = swap || select[k] != 0;
    }
    if (swap) {
      ks += 1;
      if (k != ks) {
        int inew = ks;
        const bool ok = trexchange(n, T, Z, wantz != 0, k, inew);
        ks = inew;
        if (!ok) {
          status = k;
          break;
        }
      }
      if (pair) ks += 1;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    info[0] = status;
    if (status == 0) {
      for (int k = 1; k <= n; ++k) {
        W[2 * (k - 1)] = T(k, k);
        W[2 * (k - 1) + 1] = 0.0;
      }
      for (int k = 1; k <= n - 1; ++k)
        if (T(k + 1, k) != 0.0) {
          // vals[k] += y im; vals[k + 1] -= y im (complex: the real parts
          // get + 0.0 and - 0.0)
          const double y = sqrt(fabs(T(k, k + 1))) * sqrt(fabs(T(k + 1, k)));
          W[2 * (k - 1)] = W[2 * (k - 1)] + 0.0;
          W[2 * (k - 1) + 1] = W[2 * (k - 1) + 1] + y;
          W[2 * k] = W[2 * k] - 0.0;
          W[2 * k + 1] = W[2 * k + 1] - y;
        }
    }
  }
}

hipError_t launch_ordschur_d(int n, double* T, int64_t ldt, double* Z, int64_t ldz, bool wantz, const int32_t* select,
                             double* W, int32_t* info, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(dordschur_kernel, dim3(1), dim3(64), 0, s, n, T, ldt, Z, ldz, wantz ? 1 : 0, select, W, info);
  return hipGetLastError();
}

}  // namespace gs
