// TEST INFRASTRUCTURE ONLY — the CPU oracle for the Schur hot path.
//
// A restatement, in C++ templated on the real type, of GenericSchur.jl's
// complex and real Schur path and its eigenvector back-substitution.  Only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may call
// it, and only as the checker.  Each function cites the reference lines it
// follows.  Indices inside the restated routines are 1-based, column-major,
// as in the reference, through the accessor `Mat` below.
#pragma once
#include <algorithm>
#include <vector>

#include "scalar.hpp"

namespace orc {

enum {
  ORC_OK = 0,
  ORC_UNCONVERGED = 1,   // UnconvergedException (GenericSchur.jl:234-236,554-556)
  ORC_DIM = -1,          // DimensionMismatch
  ORC_NONREAL_SUB = -2,  // ArgumentError("algorithm assumes real subdiagonal")
};

template <class T> struct Mat {
  T* p;
  long ld;
  T& operator()(long i, long j) const { return p[(i - 1) + (j - 1) * ld]; }
};

// ------------------------------------------------------------ util.jl:5-12
template <class R> R safemin() {
  R sfmin = rt<R>::fmin();
  R small = R(1.0) / rt<R>::fmax();
  if (small >= sfmin) sfmin = small * (R(1.0) + rt<R>::eps());
  return sfmin;
}

// ------------------------------------------------- util.jl:506-525 (real)
template <class R> R norm2(long n, const R* x, long inc) {
  if (n < 1) return R(0.0);
  if (n == 1) return rt<R>::abs(x[0]);
  R scale(0.0), ssq(0.0);
  for (long i = 0; i < n; ++i) {
    const R xi = x[i * inc];
    if (xi != R(0.0)) {
      R a = rt<R>::abs(xi);
      if (scale < a) {
        R q = scale / a;
        ssq = R(1.0) + ssq * (q * q);
        scale = a;
      } else {
        R q = a / scale;
        ssq = ssq + q * q;
      }
    }
  }
  return scale * rt<R>::sqrt(ssq);
}

// ---------------------------------------------- util.jl:527-557 (complex)
template <class R> R norm2(long n, const cx<R>* x, long inc) {
  if (n < 1) return R(0.0);
  if (n == 1) return absv(x[0]);
  R scale(0.0), ssq(0.0);
  for (long i = 0; i < n; ++i) {
    const R parts[2] = {x[i * inc].re, x[i * inc].im};
    for (int p = 0; p < 2; ++p) {
      if (parts[p] != R(0.0)) {
        R a = rt<R>::abs(parts[p]);
        if (scale < a) {
          R q = scale / a;
          ssq = R(1.0) + ssq * (q * q);
          scale = a;
        } else {
          R q = a / scale;
          ssq = ssq + q * q;
        }
      }
    }
  }
  return scale * rt<R>::sqrt(ssq);
}

// ---------------------------------------------------------- util.jl:562-570
template <class R> R hypot3(R x, R y, R z) {
  R xa = rt<R>::abs(x), ya = rt<R>::abs(y), za = rt<R>::abs(z);
  R w = rmax3(xa, ya, za);
  R rw = R(1.0) / w;
  R a = rw * xa, b = rw * ya, c = rw * za;
  return w * rt<R>::sqrt(a * a + b * b + c * c);
}

// Julia's Base.Math hypot(x, y) for Float64 (the fused multiply-add branch,
// correctly rounded); other real types use their library hypot.
template <class R> R jhypot(R x, R y) { return rt<R>::hypot(x, y); }
template <> inline double jhypot<double>(double x, double y) {
  double ax = std::fabs(x), ay = std::fabs(y);
  if (std::isinf(ax) || std::isinf(ay)) return HUGE_VAL;
  if (ay > ax) std::swap(ax, ay);
  if (ay <= ax * std::sqrt(std::numeric_limits<double>::epsilon() / 2.0)) return ax;
  double scale = std::numeric_limits<double>::epsilon() * std::sqrt(std::numeric_limits<double>::min());
  if (ax > std::sqrt(std::numeric_limits<double>::max() / 2.0)) {
    ax = ax * scale;
    ay = ay * scale;
    scale = 1.0 / scale;
  } else if (ay < std::sqrt(std::numeric_limits<double>::min())) {
    ax = ax / scale;
    ay = ay / scale;
  } else {
    scale = 1.0;
  }
  double h = std::sqrt(std::fma(ax, ax, ay * ay));
  const double hsq = h * h, axsq = ax * ax;
  h -= (std::fma(-ay, ay, hsq - axsq) + std::fma(h, h, -hsq) - std::fma(ax, ax, -axsq)) / (2.0 * h);
  return h * scale;
}

// --------------------------------------- householder.jl:12-54 (real xLARFG)
template <class R> R reflector(long n, R* x, long inc) {
  if (n <= 1) return R(0.0);
  const R sfmin = R(2.0) * rt<R>::fmin() / rt<R>::eps();
  R alpha = x[0];
  R xnorm = norm2(n - 1, x + inc, inc);
  if (xnorm == R(0.0)) return R(0.0);
  R beta = -rt<R>::copysign(rt<R>::hypot(alpha, xnorm), alpha);
  int kount = 0;
  bool smallb = rt<R>::abs(beta) < sfmin;
  if (smallb) {
    const R rsfmin = R(1.0) / sfmin;
    while (smallb) {
      kount += 1;
      for (long j = 1; j < n; ++j) x[j * inc] = x[j * inc] * rsfmin;
      beta = beta * rsfmin;
      alpha = alpha * rsfmin;
      smallb = (rt<R>::abs(beta) < sfmin) && (kount < 20);
    }
    xnorm = norm2(n - 1, x + inc, inc);
    beta = -rt<R>::copysign(rt<R>::hypot(alpha, xnorm), alpha);
  }
  R tau = (beta - alpha) / beta;
  R t = R(1.0) / (alpha - beta);
  for (long j = 1; j < n; ++j) x[j * inc] = x[j * inc] * t;
  for (int j = 0; j < kount; ++j) beta = beta * sfmin;
  x[0] = beta;
  return tau;
}

// ------------------------------------- householder.jl:56-102 (complex xLARFG)
template <class R> cx<R> reflector(long n, cx<R>* x, long inc) {
  if (n < 1) return cx<R>();
  const R sfmin = rt<R>::fmin() / rt<R>::eps();
  cx<R> alpha = x[0];
  R ar = alpha.re, ai = alpha.im;
  R xnorm = norm2(n - 1, x + inc, inc);
  if (xnorm == R(0.0) && ai == R(0.0)) return cx<R>();
  R beta = -rt<R>::copysign(hypot3(ar, ai, xnorm), ar);
  int kount = 0;
  bool smallb = rt<R>::abs(beta) < sfmin;
  if (smallb) {
    const R rsfmin = R(1.0) / sfmin;
    while (smallb) {
      kount += 1;
      for (long j = 1; j < n; ++j) x[j * inc] = x[j * inc] * rsfmin;
      beta = beta * rsfmin;
      ar = ar * rsfmin;
      ai = ai * rsfmin;
      smallb = (rt<R>::abs(beta) < sfmin) && (kount < 20);
    }
    xnorm = norm2(n - 1, x + inc, inc);
    alpha = cx<R>(ar, ai);
    beta = -rt<R>::copysign(hypot3(ar, ai, xnorm), ar);
  }
  cx<R> tau((beta - ar) / beta, -ai / beta);
  cx<R> t = cx<R>(R(1.0)) / (alpha - cx<R>(beta));
  for (long j = 1; j < n; ++j) x[j * inc] = x[j * inc] * t;
  for (int j = 0; j < kount; ++j) beta = beta * sfmin;
  x[0] = cx<R>(beta);
  return tau;
}

// --------------------- hessenberg.jl:3-17 with householder.jl:140-172 and
// util.jl:100-114: unblocked Householder reduction A <- Q^H A Q.
// On exit A holds H on and above the subdiagonal and reflector i in
// A[i+2:n, i]; tau[i-1] holds its scalar.
template <class T> void hessenberg(long n, T* a, long lda, T* tau) {
  Mat<T> A{a, lda};
  std::vector<T> x(n > 0 ? n : 1);
  for (long i = 1; i <= n - 1; ++i) {
    T t = reflector(n - i, &A(i + 1, i), 1);
    tau[i - 1] = t;
    const long m = n - i;  // order of the reflector block
    // lmul!(H', A[i+1:n, i+1:n])   (householder.jl:157-172)
    for (long j = i + 1; j <= n; ++j) {
      T va = A(i + 1, j);
      for (long r = 2; r <= m; ++r) va = va + conj(A(i + r, i)) * A(i + r, j);
      va = conj(t) * va;
      A(i + 1, j) = A(i + 1, j) - va;
      for (long r = 2; r <= m; ++r) A(i + r, j) = A(i + r, j) - va * A(i + r, i);
    }
    // rmul!(A[:, i+1:n], H, x)     (householder.jl:140-155)
    for (long r = 1; r <= n; ++r) {
      T s = T();
      for (long c = 2; c <= m; ++c) s = s + A(r, i + c) * A(i + c, i);
      x[r - 1] = s + A(r, i + 1);
    }
    for (long r = 1; r <= n; ++r) A(r, i + 1) = A(r, i + 1) - t * x[r - 1];
    for (long c = 2; c <= m; ++c) {
      const T yc = conj(A(i + c, i));
      for (long r = 1; r <= n; ++r) A(r, i + c) = A(r, i + c) + x[r - 1] * (-t) * yc;
    }
  }
}

// ---------------- hessenberg.jl:150-166: Q = diag(1, H_1 H_2 ... H_{n-1}),
// formed by backward accumulation of the reflectors onto the identity.
template <class T> void materialize_q(long n, const T* a, long lda, const T* tau, T* q, long ldq) {
  Mat<const T> A{a, lda};
  Mat<T> Q{q, ldq};
  for (long j = 1; j <= n; ++j)
    for (long i = 1; i <= n; ++i) Q(i, j) = (i == j) ? T(1.0) : T();
  for (long i = n - 1; i >= 1; --i) {
    const T t = tau[i - 1];
    for (long c = i + 1; c <= n; ++c) {
      T s = Q(i + 1, c);
      for (long r = i + 2; r <= n; ++r) s = s + conj(A(r, i)) * Q(r, c);
      s = t * s;
      Q(i + 1, c) = Q(i + 1, c) - s;
      for (long r = i + 2; r <= n; ++r) Q(r, c) = Q(r, c) - s * A(r, i);
    }
  }
}

// ---------------------------------------------------------- util.jl:41-80
template <class T> void safescale(long len, T* a, typename real_of<T>::type cfrom,
                                  typename real_of<T>::type cto) {
  typedef typename real_of<T>::type R;
  const R smlnum = safemin<R>();
  const R bignum = R(1.0) / smlnum;
  R cfromc = cfrom, ctoc = cto, mul(1.0);
  bool done = false;
  while (!done) {
    R cfrom1 = cfromc * smlnum;
    if (cfrom1 == cfromc) {
      mul = ctoc / cfromc;
      done = true;
    } else {
      R cto1 = ctoc / bignum;
      if (cto1 == ctoc) {
        mul = cto;
        done = true;
        cfromc = R(1.0);
      } else if (rt<R>::abs(cfrom1) > rt<R>::abs(ctoc) && ctoc != R(0.0)) {
        mul = smlnum;
        done = false;
        cfromc = R(1.0);
      } else if (rt<R>::abs(cto1) > rt<R>::abs(cfromc)) {
        mul = bignum;
        done = false;
        cfromc = cfrom1;
      } else {
        mul = ctoc / cfromc;
        done = true;
      }
    }
    for (long i = 0; i < len; ++i) a[i] = a[i] * mul;
  }
}

// ---------------------------------------------------------- util.jl:14-29
// Returns whether A was scaled; cscale and anrm receive the factors.
template <class T> bool scale_matrix(long n, T* a, long lda, typename real_of<T>::type& cscale,
                                     typename real_of<T>::type& anrm) {
  typedef typename real_of<T>::type R;
  const R smlnum = rt<R>::sqrt(safemin<R>()) / rt<R>::eps();
  const R bignum = R(1.0) / smlnum;
  anrm = R(0.0);  // norm(A, Inf): largest entry modulus
  for (long j = 0; j < n; ++j)
    for (long i = 0; i < n; ++i) anrm = rmax(anrm, absv(a[i + j * lda]));
  cscale = R(1.0);
  bool scaled = false;
  if (anrm > R(0.0) && anrm < smlnum) {
    scaled = true;
    cscale = smlnum;
  } else if (anrm > bignum) {
    scaled = true;
    cscale = bignum;
  }
  if (scaled) {
    for (long j = 0; j < n; ++j) safescale(n, a + j * lda, anrm, cscale);
  }
  return scaled;
}

// --------------------------------------------- GenericSchur.jl:374-504
template <class R> void single_shift_qr(long n, Mat<cx<R>> HH, cx<R>* zp, long ldz, cx<R> shift,
                                        long istart, long iend) {
  typedef cx<R> T;
  Mat<T> Z{zp, ldz};
  const R ulp = rt<R>::eps();
  long istart1 = -1;
  T v[2];
  T h11s;
  R h21(0.0);
  bool flag = false;
  for (long mm = iend - 1; mm >= istart + 1; --mm) {
    T h11 = HH(mm, mm);
    T h22 = HH(mm + 1, mm + 1);
    h11s = h11 - shift;
    h21 = HH(mm + 1, mm).re;
    R s = abs1(h11s) + rt<R>::abs(h21);
    h11s = h11s / s;
    h21 = h21 / s;
    v[0] = h11s;
    v[1] = T(h21);
    R h10 = HH(mm, mm - 1).re;
    if (rt<R>::abs(h10) * rt<R>::abs(h21) <= ulp * (abs1(h11s) * (abs1(h11) + abs1(h22)))) {
      istart1 = mm;
      flag = true;
      break;
    }
  }
  if (!flag) {
    istart1 = istart;
    T h11 = HH(istart, istart);
    h11s = h11 - shift;
    h21 = HH(istart + 1, istart).re;
    R s = abs1(h11s) + rt<R>::abs(h21);
    h11s = h11s / s;
    h21 = h21 / s;
    v[0] = h11s;
    v[1] = T(h21);
  }
  for (long k = istart1; k <= iend - 1; ++k) {
    if (k > istart1) {
      v[0] = HH(k, k - 1);
      v[1] = HH(k + 1, k - 1);
    }
    T tau1 = reflector(2, v, 1);
    if (k > istart1) {
      HH(k, k - 1) = v[0];
      HH(k + 1, k - 1) = T();
    }
    T v2 = v[1];
    R tau2 = (tau1 * v2).re;
    for (long j = k; j <= n; ++j) {
      T ss = conj(tau1) * HH(k, j) + tau2 * HH(k + 1, j);
      HH(k, j) = HH(k, j) - ss;
      HH(k + 1, j) = HH(k + 1, j) - ss * v2;
    }
    const long jmax = std::min(k + 2, iend);
    for (long j = 1; j <= jmax; ++j) {
      T ss = tau1 * HH(j, k) + tau2 * HH(j, k + 1);
      HH(j, k) = HH(j, k) - ss;
      HH(j, k + 1) = HH(j, k + 1) - ss * conj(v2);
    }
    if (zp) {
      for (long j = 1; j <= n; ++j) {
        T ss = tau1 * Z(j, k) + tau2 * Z(j, k + 1);
        Z(j, k) = Z(j, k) - ss;
        Z(j, k + 1) = Z(j, k + 1) - ss * conj(v2);
      }
    }
    if (k == istart1 && istart1 > istart) {
      T t = T(R(1.0)) - tau1;
      t = t / absv(t);
      HH(istart1 + 1, istart1) = HH(istart1 + 1, istart1) * conj(t);
      if (istart1 + 2 <= iend) HH(istart1 + 2, istart1 + 1) = HH(istart1 + 2, istart1 + 1) * t;
      for (long j = istart1; j <= iend; ++j) {
        if (j != istart1 + 1) {
          if (n > j)
            for (long c = j + 1; c <= n; ++c) HH(j, c) = HH(j, c) * t;
          for (long r = 1; r <= j - 1; ++r) HH(r, j) = HH(r, j) * conj(t);
          if (zp)
            for (long r = 1; r <= n; ++r) Z(r, j) = Z(r, j) * conj(t);
        }
      }
    }
  }
  T t = HH(iend, iend - 1);
  if (t.im != R(0.0)) {
    R rtv = absv(t);
    HH(iend, iend - 1) = T(rtv);
    t = t / rtv;
    if (n > iend)
      for (long c = iend + 1; c <= n; ++c) HH(iend, c) = HH(iend, c) * conj(t);
    for (long r = 1; r <= iend - 1; ++r) HH(r, iend) = HH(r, iend) * t;
    if (zp)
      for (long r = 1; r <= n; ++r) Z(r, iend) = Z(r, iend) * t;
  }
}

// ----------------- GenericSchur.jl:194-335: complex Schur of a Hessenberg
// matrix (zlahqr logic).  `stats` (optional) receives [iterations, sweeps,
// bulge steps].
template <class R> int schur_hess_complex(long n, cx<R>* h, long ldh, cx<R>* zp, long ldz,
                                          long maxiter, long maxinner, bool checksd, cx<R>* w,
                                          long* stats = nullptr) {
  typedef cx<R> T;
  Mat<T> HH{h, ldh};
  if (checksd) {
    for (long j = 1; j <= n - 1; ++j)
      if (!isreal(HH(j + 1, j))) return ORC_NONREAL_SUB;
  }
  const R ulp = rt<R>::eps();
  const R smallnum = safemin<R>() * (R(double(n)) / ulp);
  const R half = R(1.0) / R(2.0);
  const R threeq = R(3.0) / R(4.0);
  for (long j = 1; j <= n - 1; ++j)
    for (long i = j + 2; i <= n; ++i) HH(i, j) = T();
  long it = 0, nsweep = 0, nstep = 0;
  long iend = n;
  long istart = 1;
  while (iend >= 1) {
    istart = 1;
    for (long its = 0; its <= maxinner; ++its) {
      it += 1;
      if (it > maxiter) {
        if (stats) { stats[0] = it; stats[1] = nsweep; stats[2] = nstep; }
        return ORC_UNCONVERGED;
      }
      for (long is = iend - 1; is >= istart; --is) {
        if (abs1(HH(is + 1, is)) <= smallnum) {
          istart = is + 1;
          break;
        }
        R tst = abs1(HH(is, is)) + abs1(HH(is + 1, is + 1));
        if (tst == R(0.0)) {
          if (is - 1 >= 1) tst = tst + rt<R>::abs(HH(is, is - 1).re);
          if (is + 2 <= n) tst = tst + rt<R>::abs(HH(is + 2, is + 1).re);
        }
        if (rt<R>::abs(HH(is + 1, is).re) <= ulp * tst) {
          R ab = rmax(abs1(HH(is + 1, is)), abs1(HH(is, is + 1)));
          R ba = rmin(abs1(HH(is + 1, is)), abs1(HH(is, is + 1)));
          R aa = rmax(abs1(HH(is + 1, is + 1)), abs1(HH(is, is) - HH(is + 1, is + 1)));
          R bb = rmin(abs1(HH(is + 1, is + 1)), abs1(HH(is, is) - HH(is + 1, is + 1)));
          R s = aa + ab;
          if (ba * (ab / s) <= rmax(smallnum, ulp * (bb * (aa / s)))) {
            istart = is + 1;
            break;
          }
        }
      }
      if (istart > 1) HH(istart, istart - 1) = T();
      if (istart >= iend) {
        iend -= 1;
        break;
      }
      T t;
      if (its % 30 == 10) {
        R s = threeq * rt<R>::abs(HH(istart + 1, istart).re);
        t = T(s) + HH(istart, istart);
      } else if (its % 30 == 20) {
        R s = threeq * rt<R>::abs(HH(iend, iend - 1).re);
        t = T(s) + HH(iend, iend);
      } else {
        t = HH(iend, iend);
        T u = csqrt(HH(iend - 1, iend)) * csqrt(HH(iend, iend - 1));
        R s = abs1(u);
        if (s != R(0.0)) {
          T x = half * (HH(iend - 1, iend - 1) - t);
          R sx = abs1(x);
          s = rmax(s, abs1(x));
          T xs = x / s, us = u / s;
          T y = s * csqrt(xs * xs + us * us);
          if (sx > R(0.0)) {
            T xx = x / sx;
            if (xx.re * y.re + xx.im * y.im < R(0.0)) y = -y;
          }
          t = t - u * (u / (x + y));
        }
      }
      nsweep += 1;
      nstep += iend - istart;  // upper bound on the bulge steps of this sweep
      single_shift_qr<R>(n, HH, zp, ldz, t, istart, iend);
    }
  }
  for (long j = 1; j <= n; ++j) w[j - 1] = HH(j, j);
  if (stats) { stats[0] = it; stats[1] = nsweep; stats[2] = nstep; }
  return ORC_OK;
}

// --------------------------------------- GenericSchur.jl:716-803 (dlanv2)
// In/out: a,b,c,d.  Out: rotation (cs, sn) and eigenvalues w1, w2.
template <class R> void gs2x2(R& a, R& b, R& c, R& d, R& cs, R& sn, cx<R>& w1, cx<R>& w2) {
  auto sgn = [](R x) { return (x < R(0.0)) ? R(-1.0) : R(1.0); };
  const R half = R(1.0) / R(2.0);
  const R small = R(4.0) * rt<R>::eps();
  if (c == R(0.0)) {
    cs = R(1.0);
    sn = R(0.0);
  } else if (b == R(0.0)) {
    cs = R(0.0);
    sn = R(1.0);
    R a0 = a, c0 = c, d0 = d;
    a = d0;
    b = -c0;
    c = R(0.0);
    d = a0;
  } else if ((a - d) == R(0.0) && (b * c < R(0.0))) {
    cs = R(1.0);
    sn = R(0.0);
  } else {
    R asubd = a - d;
    R p = half * asubd;
    R bcmax = rmax(rt<R>::abs(b), rt<R>::abs(c));
    R bcmis = rmin(rt<R>::abs(b), rt<R>::abs(c)) * sgn(b) * sgn(c);
    R scale = rmax(rt<R>::abs(p), bcmax);
    R z = (p / scale) * p + (bcmax / scale) * bcmis;
    if (z >= small) {
      z = p + rt<R>::sqrt(scale) * rt<R>::sqrt(z) * sgn(p);
      a = d + z;
      d = d - (bcmax / z) * bcmis;
      R tau = jhypot<R>(c, z);
      cs = z / tau;
      sn = c / tau;
      b = b - c;
      c = R(0.0);
    } else {
      R sigma = b + c;
      R tau = jhypot<R>(sigma, asubd);
      cs = rt<R>::sqrt(half * (R(1.0) + rt<R>::abs(sigma) / tau));
      sn = -(p / (tau * cs)) * sgn(sigma);
      R aa = a * cs + b * sn;
      R bb = -a * sn + b * cs;
      R cc = c * cs + d * sn;
      R dd = -c * sn + d * cs;
      a = aa * cs + cc * sn;
      b = bb * cs + dd * sn;
      c = -aa * sn + cc * cs;
      d = -bb * sn + dd * cs;
      R midad = half * (a + d);
      a = midad;
      d = a;
      if (c != R(0.0)) {
        if (b != R(0.0)) {
          if (b * c >= R(0.0)) {
            R sab = rt<R>::sqrt(rt<R>::abs(b));
            R sac = rt<R>::sqrt(rt<R>::abs(c));
            p = sab * sac * sgn(c);
            tau = R(1.0) / rt<R>::sqrt(rt<R>::abs(b + c));
            a = midad + p;
            d = midad - p;
            b = b - c;
            c = R(0.0);
            R cs1 = sab * tau;
            R sn1 = sac * tau;
            R csn = cs * cs1 - sn * sn1;
            R snn = cs * sn1 + sn * cs1;
            cs = csn;
            sn = snn;
          }
        } else {
          b = -c;
          c = R(0.0);
          R tmp = cs;
          cs = -sn;
          sn = tmp;
        }
      }
    }
  }
  if (c == R(0.0)) {
    w1 = cx<R>(a);
    w2 = cx<R>(d);
  } else {
    R rti = rt<R>::sqrt(rt<R>::abs(b)) * rt<R>::sqrt(rt<R>::abs(c));
    w1 = cx<R>(a, rti);
    w2 = cx<R>(d, -rti);
  }
}

// --------------------------------------------- GenericSchur.jl:837-952
template <class R> void double_shift_qr(long n, Mat<R> H, R* zp, long ldz, cx<R> shift1, cx<R> shift2,
                                        long istart, long iend) {
  Mat<R> Z{zp, ldz};
  const long i1 = 1, i2 = n;
  const R r1r = shift1.re, r1i = shift1.im, r2r = shift2.re, r2i = shift2.im;
  R v[3] = {R(0.0), R(0.0), R(0.0)};
  const R eps = rt<R>::eps();
  long mx = istart;
  for (long m = iend - 2; m >= istart; --m) {
    R H21s = H(m + 1, m);
    R s = rt<R>::abs(H(m, m) - r2r) + rt<R>::abs(r2i) + rt<R>::abs(H21s);
    H21s = H21s / s;
    v[0] = H21s * H(m, m + 1) + (H(m, m) - r1r) * ((H(m, m) - r2r) / s) - r1i * (r2i / s);
    v[1] = H21s * (H(m, m) + H(m + 1, m + 1) - r1r - r2r);
    v[2] = H21s * H(m + 2, m + 1);
    s = rt<R>::abs(v[0]) + rt<R>::abs(v[1]) + rt<R>::abs(v[2]);
    v[0] = v[0] / s;
    v[1] = v[1] / s;
    v[2] = v[2] / s;
    if (m > istart &&
        rt<R>::abs(H(m, m - 1)) * (rt<R>::abs(v[1]) + rt<R>::abs(v[2])) <=
            eps * rt<R>::abs(v[0]) * (rt<R>::abs(H(m - 1, m - 1)) + rt<R>::abs(H(m, m)) + rt<R>::abs(H(m + 1, m + 1)))) {
      mx = m;
      break;
    }
  }
  for (long k = mx; k <= iend - 1; ++k) {
    const long nr = std::min(3L, iend - k + 1);
    if (k > mx)
      for (long ii = 0; ii < nr; ++ii) v[ii] = H(k + ii, k - 1);
    R tau1 = reflector(nr, v, 1);
    if (k > mx) {
      H(k, k - 1) = v[0];
      H(k + 1, k - 1) = R(0.0);
      if (k < iend - 1) H(k + 2, k - 1) = R(0.0);
    } else if (mx > istart) {
      H(k, k - 1) = H(k, k - 1) * (R(1.0) - tau1);
    }
    const R v2 = v[1];
    const R tau2 = tau1 * v2;
    if (nr == 3) {
      const R v3 = v[2];
      const R tau3 = tau1 * v3;
      for (long j = k; j <= i2; ++j) {
        R ss = H(k, j) + v2 * H(k + 1, j) + v3 * H(k + 2, j);
        H(k, j) = H(k, j) - ss * tau1;
        H(k + 1, j) = H(k + 1, j) - ss * tau2;
        H(k + 2, j) = H(k + 2, j) - ss * tau3;
      }
      for (long j = i1; j <= std::min(k + 3, iend); ++j) {
        R ss = H(j, k) + v2 * H(j, k + 1) + v3 * H(j, k + 2);
        H(j, k) = H(j, k) - ss * tau1;
        H(j, k + 1) = H(j, k + 1) - ss * tau2;
        H(j, k + 2) = H(j, k + 2) - ss * tau3;
      }
      if (zp) {
        for (long j = 1; j <= n; ++j) {
          R ss = Z(j, k) + v2 * Z(j, k + 1) + v3 * Z(j, k + 2);
          Z(j, k) = Z(j, k) - ss * tau1;
          Z(j, k + 1) = Z(j, k + 1) - ss * tau2;
          Z(j, k + 2) = Z(j, k + 2) - ss * tau3;
        }
      }
    } else if (nr == 2) {
      for (long j = k; j <= i2; ++j) {
        R ss = H(k, j) + v2 * H(k + 1, j);
        H(k, j) = H(k, j) - ss * tau1;
        H(k + 1, j) = H(k + 1, j) - ss * tau2;
      }
      for (long j = i1; j <= iend; ++j) {
        R ss = H(j, k) + v2 * H(j, k + 1);
        H(j, k) = H(j, k) - ss * tau1;
        H(j, k + 1) = H(j, k + 1) - ss * tau2;
      }
      if (zp) {
        for (long j = 1; j <= n; ++j) {
          R ss = Z(j, k) + v2 * Z(j, k + 1);
          Z(j, k) = Z(j, k) - ss * tau1;
          Z(j, k + 1) = Z(j, k + 1) - ss * tau2;
        }
      }
    }
  }
}

// ------------ GenericSchur.jl:513-699: real Schur of a Hessenberg matrix
// (dlahqr logic) to standard form.  w receives the eigenvalues by position.
template <class R> int schur_hess_real(long n, R* h, long ldh, R* zp, long ldz, long maxiter, cx<R>* w,
                                       long* stats = nullptr) {
  Mat<R> HH{h, ldh};
  Mat<R> Z{zp, ldz};
  for (long j = 1; j <= n; ++j)
    for (long i = j + 2; i <= n; ++i) HH(i, j) = R(0.0);
  long iwcur = n;
  const R eps = rt<R>::eps();
  const R smallnum = rt<R>::fmin() * (R(double(n)) / eps);
  const R threeq = R(3.0) / R(4.0);
  const R m7_16 = R(-7.0) / R(16.0);
  long iter = 0, nsweep = 0, nstep = 0;
  long iend = n, istart = 1;
  while (iend >= 1) {
    istart = 1;
    long iterqr = 0;
    bool deflate = false;
    while (true) {
      iter += 1;
      if (iter > maxiter) {
        if (stats) { stats[0] = iter; stats[1] = nsweep; stats[2] = nstep; }
        return ORC_UNCONVERGED;
      }
      bool split = false;
      for (long k = iend; k >= istart + 1; --k) {
        if (rt<R>::abs(HH(k, k - 1)) < smallnum) {
          split = true;
        } else {
          R Hkk = HH(k, k);
          R Hkm1km1 = HH(k - 1, k - 1);
          R t = rt<R>::abs(Hkm1km1) + rt<R>::abs(Hkk);
          if (t == R(0.0)) {
            if (k > 2) t = t + rt<R>::abs(HH(k - 1, k - 2));
            if (k + 1 <= n) t = t + rt<R>::abs(HH(k + 1, k));
          }
          R aHkkm1 = rt<R>::abs(HH(k, k - 1));
          if (aHkkm1 <= t * eps) {
            R aHkm1k = rt<R>::abs(HH(k - 1, k));
            R ab = rmax(aHkkm1, aHkm1k);
            R ba = rmin(aHkkm1, aHkm1k);
            R aa = rmax(rt<R>::abs(Hkk), rt<R>::abs(Hkm1km1 - Hkk));
            R bb = rmin(rt<R>::abs(Hkk), rt<R>::abs(Hkm1km1 - Hkk));
            R s = aa + bb;
            if (ba * (ab / s) <= rmax(smallnum, eps * (bb * (aa / s)))) split = true;
          }
        }
        if (split) {
          istart = k;
          break;
        }
      }
      if (!split) istart = 1;
      if (istart > 1) HH(istart, istart - 1) = R(0.0);
      if (istart >= iend - 1) {
        deflate = true;
        break;
      }
      iterqr += 1;
      R H11, H12, H21, H22;
      if (iterqr == 10) {
        R s = rt<R>::abs(HH(istart + 1, istart)) + rt<R>::abs(HH(istart + 2, istart + 1));
        H11 = threeq * s + HH(istart, istart);
        H12 = m7_16 * s;
        H21 = s;
        H22 = H11;
      } else if (iterqr == 20) {
        R s = rt<R>::abs(HH(iend, iend - 1)) + rt<R>::abs(HH(iend - 1, iend - 2));
        H11 = threeq * s + HH(iend, iend);
        H12 = m7_16 * s;
        H21 = s;
        H22 = H11;
      } else {
        H11 = HH(iend - 1, iend - 1);
        H21 = HH(iend, iend - 1);
        H12 = HH(iend - 1, iend);
        H22 = HH(iend, iend);
      }
      R s = rt<R>::abs(H11) + rt<R>::abs(H12) + rt<R>::abs(H21) + rt<R>::abs(H22);
      R r1r, r2r, r1i, r2i;
      if (s == R(0.0)) {
        r1r = r2r = r1i = r2i = R(0.0);
      } else {
        H11 = H11 / s;
        H12 = H12 / s;
        H21 = H21 / s;
        H22 = H22 / s;
        R tr = (H11 + H22) / R(2.0);
        R d = (H11 - tr) * (H22 - tr) - H12 * H21;
        R rtd = rt<R>::sqrt(rt<R>::abs(d));
        if (d >= R(0.0)) {
          r1r = tr * s;
          r2r = r1r;
          r1i = rtd * s;
          r2i = -r1i;
        } else {
          r1r = tr + rtd;
          r2r = tr - rtd;
          if (rt<R>::abs(r1r - H22) <= rt<R>::abs(r2r - H22)) {
            r1r = r1r * s;
            r2r = r1r;
          } else {
            r2r = r2r * s;
            r1r = r2r;
          }
          r1i = r2i = R(0.0);
        }
      }
      nsweep += 1;
      nstep += iend - istart;
      double_shift_qr<R>(n, HH, zp, ldz, cx<R>(r1r, r1i), cx<R>(r2r, r2i), istart, iend);
    }
    if (deflate && istart >= iend) {
      w[iwcur - 1] = cx<R>(HH(iend, iend));
      iwcur -= 1;
      iend = istart - 1;
    } else if (deflate && istart + 1 == iend) {
      R a = HH(iend - 1, iend - 1), b = HH(iend - 1, iend), c = HH(iend, iend - 1), d = HH(iend, iend);
      R cs, sn;
      cx<R> w1, w2;
      gs2x2(a, b, c, d, cs, sn, w1, w2);
      w[iwcur - 1] = w2;
      iwcur -= 1;
      w[iwcur - 1] = w1;
      iwcur -= 1;
      // lmul!(G, HH[:, istart:n]) with G = Givens(iend-1, iend, cs, sn)
      for (long j = istart; j <= n; ++j) {
        R a1 = HH(iend - 1, j), a2 = HH(iend, j);
        HH(iend - 1, j) = cs * a1 + sn * a2;
        HH(iend, j) = -sn * a1 + cs * a2;
      }
      // rmul!(HH[1:iend, :], G')
      for (long i = 1; i <= iend; ++i) {
        R a1 = HH(i, iend - 1), a2 = HH(i, iend);
        HH(i, iend - 1) = a1 * cs + a2 * sn;
        HH(i, iend) = -a1 * sn + a2 * cs;
      }
      HH(iend - 1, iend - 1) = a;
      HH(iend - 1, iend) = b;
      HH(iend, iend - 1) = c;
      HH(iend, iend) = d;
      if (iend > 2) HH(iend - 1, iend - 2) = R(0.0);
      if (zp) {
        for (long i = 1; i <= n; ++i) {
          R a1 = Z(i, iend - 1), a2 = Z(i, iend);
          Z(i, iend - 1) = a1 * cs + a2 * sn;
          Z(i, iend) = -a1 * sn + a2 * cs;
        }
      }
    }
    iend = istart - 1;
  }
  for (long j = 1; j <= n; ++j)
    for (long i = j + 2; i <= n; ++i) HH(i, j) = R(0.0);
  if (stats) { stats[0] = iter; stats[1] = nsweep; stats[2] = nstep; }
  return ORC_OK;
}

// ------ GenericSchur.jl:350-372 (complex) and :805-835 (real) drivers.
// On entry `a` holds A (overwritten).  On exit t holds T, z holds Z (when
// wantz) and w the eigenvalues.
template <class R> int gschur(long n, cx<R>* a, long lda, cx<R>* t, long ldt, cx<R>* z, long ldz, cx<R>* w,
                              bool wantz, bool scale, long maxiter, long maxinner, long* stats = nullptr) {
  typedef cx<R> T;
  R cscale(1.0), anrm(1.0);
  bool scaled = false;
  if (scale) scaled = scale_matrix(n, a, lda, cscale, anrm);
  std::vector<T> tau(n > 1 ? n - 1 : 1);
  hessenberg(n, a, lda, tau.data());
  if (wantz) materialize_q(n, a, lda, tau.data(), z, ldz);
  for (long j = 0; j < n; ++j)
    for (long i = 0; i < n; ++i) t[i + j * ldt] = a[i + j * lda];
  int info = schur_hess_complex<R>(n, t, ldt, wantz ? z : nullptr, ldz, maxiter, maxinner, false, w, stats);
  if (info != ORC_OK) return info;
  Mat<T> TT{t, ldt};
  for (long j = 1; j <= n; ++j)
    for (long i = j + 1; i <= n; ++i) TT(i, j) = T();  // triu(HH)
  if (scaled) {
    for (long j = 0; j < n; ++j) safescale(n, t + j * ldt, cscale, anrm);
    for (long j = 1; j <= n; ++j) w[j - 1] = TT(j, j);
  }
  return ORC_OK;
}

template <class R> int gschur(long n, R* a, long lda, R* t, long ldt, R* z, long ldz, cx<R>* w, bool wantz,
                              bool scale, long maxiter, long /*maxinner*/, long* stats = nullptr) {
  R cscale(1.0), anrm(1.0);
  bool scaled = false;
  if (scale) scaled = scale_matrix(n, a, lda, cscale, anrm);
  std::vector<R> tau(n > 1 ? n - 1 : 1);
  hessenberg(n, a, lda, tau.data());
  if (wantz) materialize_q(n, a, lda, tau.data(), z, ldz);
  for (long j = 0; j < n; ++j)
    for (long i = 0; i < n; ++i) t[i + j * ldt] = a[i + j * lda];
  int info = schur_hess_real<R>(n, t, ldt, wantz ? z : nullptr, ldz, maxiter, w, stats);
  if (info != ORC_OK) return info;
  if (scaled) {
    for (long j = 0; j < n; ++j) safescale(n, t + j * ldt, cscale, anrm);
    safescale(n, w, cscale, anrm);
  }
  return ORC_OK;
}

// ========================================================= eigenvectors

// --------------------------------------------- util.jl:128-293 (zlatrs)
template <class T> typename real_of<T>::type usolve(Mat<T> A, long n, T* x, typename real_of<T>::type* cnorm) {
  typedef typename real_of<T>::type R;
  auto cabs1half = [](const T& z) { return rt<R>::abs(real(z) / R(2.0)) + rt<R>::abs(imag(z) / R(2.0)); };
  const R half = R(1.0) / R(2.0);
  const R smallnum = safemin<R>() / rt<R>::eps();
  const R bignum = R(1.0) / smallnum;
  R xscale(1.0);
  R tmax = rt<R>::abs(cnorm[0]);
  for (long j = 2; j <= n; ++j) tmax = rmax(tmax, rt<R>::abs(cnorm[j - 1]));
  R tscale;
  if (tmax <= bignum * half) {
    tscale = R(1.0);
  } else {
    tscale = half / (smallnum * tmax);
    for (long j = 0; j < n; ++j) cnorm[j] = cnorm[j] * tscale;
  }
  R xmax = cabs1half(x[0]);
  for (long j = 2; j <= n; ++j) xmax = rmax(xmax, cabs1half(x[j - 1]));
  R xbound = xmax;
  R grow;
  if (tscale != R(1.0)) {
    grow = R(0.0);
  } else {
    grow = half / rmax(xbound, smallnum);
    xbound = grow;
    bool toosmall = false;
    for (long j = n; j >= 1; --j) {
      toosmall = (grow <= smallnum);
      if (toosmall) break;
      T tjjs = A(j, j);
      R tjj = abs1(tjjs);
      if (tjj >= smallnum) {
        xbound = rmin(xbound, rmin(R(1.0), tjj) * grow);
      } else {
        xbound = R(0.0);
      }
      if (tjj + cnorm[j - 1] >= smallnum) {
        grow = grow * (tjj / (tjj + cnorm[j - 1]));
      } else {
        grow = R(0.0);
      }
    }
    if (!toosmall) grow = xbound;
  }
  if (grow * tscale > smallnum) {
    for (long j = n; j >= 1; --j) {
      x[j - 1] = x[j - 1] / A(j, j);
      T xj = x[j - 1];
      for (long i = j - 1; i >= 1; --i) x[i - 1] = x[i - 1] - A(i, j) * xj;
    }
  } else {
    if (xmax > bignum * half) {
      xscale = (bignum * half) / xmax;
      for (long i = 0; i < n; ++i) x[i] = x[i] * xscale;
      xmax = bignum;
    } else {
      xmax = xmax * R(2.0);
    }
    for (long j = n; j >= 1; --j) {
      R xj = abs1(x[j - 1]);
      T tjjs = A(j, j) * tscale;
      R tjj = abs1(tjjs);
      if (tjj > smallnum) {
        if (tjj < R(1.0)) {
          if (xj > tjj * bignum) {
            R rec = R(1.0) / xj;
            for (long i = 0; i < n; ++i) x[i] = x[i] * rec;
            xscale = xscale * rec;
            xmax = xmax * rec;
          }
        }
        x[j - 1] = x[j - 1] / tjjs;
        xj = abs1(x[j - 1]);
      } else if (tjj > R(0.0)) {
        if (xj > tjj * bignum) {
          R rec = (tjj * bignum) / xj;
          if (cnorm[j - 1] > R(1.0)) rec = rec / cnorm[j - 1];
          for (long i = 0; i < n; ++i) x[i] = x[i] * rec;
          xscale = xscale * rec;
          xmax = xmax * rec;
        }
        x[j - 1] = x[j - 1] / tjjs;
        xj = abs1(x[j - 1]);
      } else {
        for (long i = 0; i < n; ++i) x[i] = T();
        x[j - 1] = T(R(1.0));
        xscale = R(0.0);
        xmax = R(0.0);
      }
      if (xj > R(1.0)) {
        R rec = R(1.0) / xj;
        if (cnorm[j - 1] > (bignum - xmax) * rec) {
          rec = rec * half;
          for (long i = 0; i < n; ++i) x[i] = x[i] * rec;
          xscale = xscale * rec;
        }
      } else if (xj * cnorm[j - 1] > bignum - xmax) {
        for (long i = 0; i < n; ++i) x[i] = x[i] * half;
        xscale = xscale * half;
      }
      if (j > 1) {
        T xjt = x[j - 1] * tscale;
        for (long i = 1; i <= j - 1; ++i) x[i - 1] = x[i - 1] - xjt * A(i, j);
        xmax = abs1(x[0]);
        for (long i = 2; i <= j - 1; ++i) xmax = rmax(xmax, abs1(x[i - 1]));
      }
    }
    xscale = xscale / tscale;
  }
  if (tscale != R(1.0)) {
    for (long i = 0; i < n; ++i) cnorm[i] = cnorm[i] * (R(1.0) / tscale);
  }
  return xscale;
}

// --------------------------------------------- util.jl:301-461 (zlatrs, ^H)
template <class T> typename real_of<T>::type cusolve(Mat<T> A, long n, T* x, typename real_of<T>::type* cnorm) {
  typedef typename real_of<T>::type R;
  auto cabs1half = [](const T& z) { return rt<R>::abs(real(z) / R(2.0)) + rt<R>::abs(imag(z) / R(2.0)); };
  const R half = R(1.0) / R(2.0);
  const R smallnum = safemin<R>() / rt<R>::eps();
  const R bignum = R(1.0) / smallnum;
  R xscale(1.0);
  R tmax = rt<R>::abs(cnorm[0]);
  for (long j = 2; j <= n; ++j) tmax = rmax(tmax, rt<R>::abs(cnorm[j - 1]));
  R tscale;
  if (tmax <= bignum * half) {
    tscale = R(1.0);
  } else {
    tscale = half / (smallnum * tmax);
    for (long j = 0; j < n; ++j) cnorm[j] = cnorm[j] * tscale;
  }
  R xmax = cabs1half(x[0]);
  for (long j = 2; j <= n; ++j) xmax = rmax(xmax, cabs1half(x[j - 1]));
  R xbound = xmax;
  R grow;
  if (tscale != R(1.0)) {
    grow = R(0.0);
  } else {
    grow = half / rmax(xbound, smallnum);
    xbound = grow;
    bool toosmall = false;
    for (long j = 1; j <= n; ++j) {
      toosmall = (grow <= smallnum);
      if (toosmall) break;
      R xj = R(1.0) + cnorm[j - 1];
      grow = rmin(grow, xbound / xj);
      T tjjs = A(j, j);
      R tjj = abs1(tjjs);
      if (tjj >= smallnum) {
        if (xj > tjj) xbound = xbound * (tjj / xj);
      } else {
        xbound = R(0.0);
      }
    }
    if (!toosmall) grow = rmin(grow, xbound);
  }
  if (grow * tscale > smallnum) {
    for (long j = 1; j <= n; ++j) {
      T zz = x[j - 1];
      for (long i = 1; i <= j - 1; ++i) zz = zz - conj(A(i, j)) * x[i - 1];
      x[j - 1] = zz / conj(A(j, j));
    }
  } else {
    if (xmax > bignum * half) {
      xscale = (bignum * half) / xmax;
      for (long i = 0; i < n; ++i) x[i] = x[i] * xscale;
      xmax = bignum;
    } else {
      xmax = xmax * R(2.0);
    }
    for (long j = 1; j <= n; ++j) {
      R xj = abs1(x[j - 1]);
      T uscale = T(tscale);
      R rec = R(1.0) / rmax(xmax, R(1.0));
      T tjjs = conj(A(j, j)) * tscale;
      if (cnorm[j - 1] > (bignum - xj) * rec) {
        rec = rec * half;
        tjjs = conj(A(j, j)) * tscale;
        R tjj = abs1(tjjs);
        if (tjj > R(1.0)) {
          rec = rmin(R(1.0), rec * tjj);
          uscale = uscale / tjjs;
        }
        if (rec < R(1.0)) {
          for (long i = 0; i < n; ++i) x[i] = x[i] * rec;
          xscale = xscale * rec;
          xmax = xmax * rec;
        }
      }
      T csumj = T();
      for (long i = 1; i <= j - 1; ++i) csumj = csumj + (conj(A(i, j)) * uscale) * x[i - 1];
      bool unscaled = (uscale == T(tscale)) || rt<R>::isnan(uscale.re) || rt<R>::isnan(imag(uscale));
      if (unscaled) {
        x[j - 1] = x[j - 1] - csumj;
        xj = abs1(x[j - 1]);
        tjjs = conj(A(j, j)) * tscale;
        R tjj = abs1(tjjs);
        if (tjj > smallnum) {
          if (tjj < R(1.0)) {
            if (xj > tjj * bignum) {
              R rec2 = R(1.0) / xj;
              for (long i = 0; i < n; ++i) x[i] = x[i] * rec2;
              xscale = xscale * rec2;
              xmax = xmax * rec2;
            }
          }
          x[j - 1] = x[j - 1] / tjjs;
        } else if (tjj > R(0.0)) {
          if (xj > tjj * bignum) {
            R rec2 = (tjj * bignum) / xj;
            for (long i = 0; i < n; ++i) x[i] = x[i] * rec2;
            xscale = xscale * rec2;
            xmax = xmax * rec2;
          }
          x[j - 1] = x[j - 1] / tjjs;
        } else {
          for (long i = 0; i < n; ++i) x[i] = T();
          x[j - 1] = T(R(1.0));
          xscale = R(0.0);
          xmax = R(0.0);
        }
      } else {
        x[j - 1] = x[j - 1] / tjjs - csumj;
      }
      xmax = rmax(xmax, abs1(x[j - 1]));
    }
    xscale = xscale / tscale;
  }
  if (tscale != R(1.0)) {
    for (long i = 0; i < n; ++i) cnorm[i] = cnorm[i] * (R(1.0) / tscale);
  }
  return xscale;
}

// ------------------------------------------------ util.jl:603-687
template <class R> R safe_lu_solve(long n, cx<R>* ap, cx<R>* b, bool& perturbed) {
  typedef cx<R> T;
  Mat<T> A{ap, n};
  perturbed = false;
  const R smallnum = rt<R>::fmin() / rt<R>::eps();
  R smin(0.0);
  for (long i = 0; i < n * n; ++i) smin = rmax(smin, absv(ap[i]));
  smin = rmax(rt<R>::eps() * smin, smallnum);
  long ipsv = 0, jpsv = 0;
  std::vector<long> jpiv(n, 0);
  for (long i = 1; i <= n - 1; ++i) {
    R xmax(0.0);
    for (long ip = i; ip <= n; ++ip)
      for (long jp = i; jp <= n; ++jp)
        if (absv(A(ip, jp)) >= xmax) {
          xmax = absv(A(ip, jp));
          ipsv = ip;
          jpsv = jp;
        }
    if (ipsv != i) {
      for (long j = 1; j <= n; ++j) std::swap(A(ipsv, j), A(i, j));
      std::swap(b[i - 1], b[ipsv - 1]);
    }
    if (jpsv != i) {
      for (long j = 1; j <= n; ++j) std::swap(A(j, jpsv), A(j, i));
    }
    jpiv[i - 1] = jpsv;
    if (absv(A(i, i)) < smin) {
      perturbed = true;
      A(i, i) = T(smin);
    }
    for (long j = i + 1; j <= n; ++j) {
      A(j, i) = A(j, i) / A(i, i);
      b[j - 1] = b[j - 1] - A(j, i) * b[i - 1];
      for (long k = i + 1; k <= n; ++k) A(j, k) = A(j, k) - A(j, i) * A(i, k);
    }
  }
  if (absv(A(n, n)) < smin) {
    perturbed = true;
    A(n, n) = T(smin);
  }
  jpiv[n - 1] = n;
  R scale(1.0);
  const R a = R(8.0) * smallnum;
  bool needscl = false;
  for (long i = 1; i <= n; ++i)
    if (a * absv(b[i - 1]) > absv(A(i, i))) {
      needscl = true;
      break;
    }
  if (needscl) {
    R bm(0.0);
    for (long i = 0; i < n; ++i) bm = rmax(bm, absv(b[i]));
    scale = (R(1.0) / R(8.0)) / bm;
    for (long i = 0; i < n; ++i) b[i] = b[i] * scale;
  }
  for (long i = 1; i <= n; ++i) {
    long k = n + 1 - i;
    T t = T(R(1.0)) / A(k, k);
    b[k - 1] = b[k - 1] * t;
    for (long j = k + 1; j <= n; ++j) b[k - 1] = b[k - 1] - (t * A(k, j)) * b[j - 1];
  }
  for (long i = 1; i <= n - 1; ++i) {
    long jj = jpiv[n - i - 1];
    if (jj != n - i) std::swap(b[n - i - 1], b[jj - 1]);
  }
  return scale;
}

// ------------------------------------------------ util.jl:698-751
// Solves (a*A - bw*D) X = s*B for na = 1 or 2, D = I.  A is given by its
// entries in column-major order (a11, a21, a12, a22).  Returns s; B is
// overwritten by X and xnorm receives the reference's xnorm.
template <class R> R xsolve(R a, long na, const R* Ae, cx<R> bw, cx<R>* B, R smin, R& xnorm) {
  typedef cx<R> T;
  R s(1.0);
  const R small = R(2.0) * safemin<R>();
  const R bignum = s / small;
  smin = rmax(smin, small);
  if (na == 1) {
    R csr = a * Ae[0] - bw.re;
    R csi = -bw.im;
    R cnorm = rt<R>::abs(csr) + rt<R>::abs(csi);
    if (cnorm < smin) {
      csr = smin;
      csi = R(0.0);
      cnorm = smin;
    }
    R bnorm = abs1(B[0]);
    if (cnorm < R(1.0) && bnorm > R(1.0) && bnorm > bignum * cnorm) s = R(1.0) / bnorm;
    T x = (B[0] * s) / T(csr, csi);
    xnorm = abs1(x);
    B[0] = x;
    return s;
  }
  R cmax(0.0);
  T M[4];
  for (long j = 0; j < 2; ++j)
    for (long i = 0; i < 2; ++i) {
      T z = T(a * Ae[i + 2 * j]);
      if (i == j) z = z - bw;
      M[i + 2 * j] = z;
      cmax = rmax(cmax, abs1(z));
    }
  if (cmax < smin) {
    R bnorm = rmax(abs1(B[0]), abs1(B[1]));
    if (smin < R(1.0) && bnorm > R(1.0) && bnorm > bignum * smin) s = R(1.0) / bnorm;
    R t1 = s / smin;
    B[0] = B[0] * t1;
    B[1] = B[1] * t1;
  } else {
    bool pert;
    s = safe_lu_solve<R>(2, M, B, pert);
  }
  xnorm = rt<R>::sqrt(abs2(B[0]) + abs2(B[1]));
  return s;
}

// ------------------------------------------------ util.jl:572-592
template <class R> void enormalize(long n, cx<R>* v, long ldv) {
  Mat<cx<R>> V{v, ldv};
  for (long j = 1; j <= n; ++j) {
    R nrm = norm2(n, &V(1, j), 1);
    R s = R(1.0) / nrm;
    R t = abs2(V(1, j));
    long i0 = 1;
    for (long i = 2; i <= n; ++i) {
      R u = abs2(V(i, j));
      if (u > t) {
        i0 = i;
        t = u;
      }
    }
    cx<R> tt = (s * conj(V(i0, j))) / rt<R>::sqrt(t);
    for (long i = 1; i <= n; ++i) V(i, j) = V(i, j) * tt;
    V(i0, j) = cx<R>(V(i0, j).re);
  }
}

// ------------------------------- vectors.jl:45-131 (ztrevc-style right)
template <class R> void eigvecs_right_complex(long n, cx<R>* tp, long ldt, const cx<R>* zp, long ldz, cx<R>* vp,
                                              long ldv) {
  typedef cx<R> T;
  Mat<T> TT{tp, ldt};
  Mat<const T> Z{zp, ldz};
  Mat<T> V{vp, ldv};
  const R ulp = rt<R>::eps();
  const R smallnum = safemin<R>() * R(double(n));
  std::vector<T> v(n), tdiag(n);
  std::vector<R> tnorms(n, R(0.0));
  for (long j = 1; j <= n; ++j) tdiag[j - 1] = TT(j, j);
  for (long j = 2; j <= n; ++j)
    for (long i = 1; i <= j - 1; ++i) tnorms[j - 1] = tnorms[j - 1] + absv(TT(i, j));
  for (long ki = n; ki >= 1; --ki) {
    R smin = rmax(ulp * abs1(TT(ki, ki)), smallnum);
    v[0] = T(R(1.0));
    for (long k = 1; k <= ki - 1; ++k) v[k - 1] = -TT(k, ki);
    for (long k = 1; k <= ki - 1; ++k) {
      TT(k, k) = TT(k, k) - TT(ki, ki);
      if (abs1(TT(k, k)) < smin) TT(k, k) = T(smin);
    }
    R vscale;
    if (ki > 1) {
      vscale = usolve<T>(TT, ki - 1, v.data(), tnorms.data());
      v[ki - 1] = T(vscale);
    } else {
      vscale = R(1.0);
    }
    if (zp) {
      for (long j = 1; j <= n; ++j) {
        V(j, ki) = vscale * Z(j, ki);
        for (long i = 1; i <= ki - 1; ++i) V(j, ki) = V(j, ki) + Z(j, i) * v[i - 1];
      }
    } else {
      for (long j = 1; j <= ki; ++j) V(j, ki) = v[j - 1];
      for (long j = ki + 1; j <= n; ++j) V(j, ki) = T();
    }
    R t0 = abs1(V(1, ki));
    for (long i = 2; i <= n; ++i) t0 = rmax(t0, abs1(V(i, ki)));
    R remax = R(1.0) / t0;
    for (long i = 1; i <= n; ++i) V(i, ki) = V(i, ki) * remax;
    for (long k = 1; k <= ki - 1; ++k) TT(k, k) = tdiag[k - 1];
  }
}

// ------------------------------ vectors.jl:372-460 (ztrevc-style left)
template <class R> void eigvecs_left_complex(long n, cx<R>* tp, long ldt, const cx<R>* zp, long ldz, cx<R>* vp,
                                             long ldv) {
  typedef cx<R> T;
  Mat<T> TT{tp, ldt};
  Mat<const T> Z{zp, ldz};
  Mat<T> V{vp, ldv};
  const R ulp = rt<R>::eps();
  const R smallnum = safemin<R>() * R(double(n));
  std::vector<T> v(n), tdiag(n);
  std::vector<R> tnorms(n, R(0.0));
  for (long j = 1; j <= n; ++j) tdiag[j - 1] = TT(j, j);
  for (long j = 2; j <= n; ++j)
    for (long i = 1; i <= j - 1; ++i) tnorms[j - 1] = tnorms[j - 1] + abs1(TT(i, j));
  for (long ki = 1; ki <= n; ++ki) {
    R smin = rmax(ulp * abs1(TT(ki, ki)), smallnum);
    for (long k = ki + 1; k <= n; ++k) v[k - 1] = -conj(TT(ki, k));
    for (long k = 1; k <= ki; ++k) v[k - 1] = T();
    for (long k = ki + 1; k <= n; ++k) {
      TT(k, k) = TT(k, k) - TT(ki, ki);
      if (abs1(TT(k, k)) < smin) TT(k, k) = T(smin);
    }
    R vscale;
    if (ki < n) {
      Mat<T> sub{&TT(ki + 1, ki + 1), ldt};
      // the reference passes the full tnorms here (vectors.jl:417-420)
      vscale = cusolve<T>(sub, n - ki, &v[ki], tnorms.data());
      v[ki - 1] = T(vscale);
    } else {
      v[n - 1] = T(R(1.0));
      vscale = R(1.0);
    }
    if (zp) {
      for (long j = 1; j <= n; ++j) {
        V(j, ki) = vscale * Z(j, ki);
        for (long i = ki + 1; i <= n; ++i) V(j, ki) = V(j, ki) + Z(j, i) * v[i - 1];
      }
    } else {
      for (long j = ki; j <= n; ++j) V(j, ki) = v[j - 1];
      for (long j = 1; j <= ki - 1; ++j) V(j, ki) = T();
    }
    R t0 = abs1(V(1, ki));
    for (long i = 2; i <= n; ++i) t0 = rmax(t0, abs1(V(i, ki)));
    R remax = R(1.0) / t0;
    for (long i = 1; i <= n; ++i) V(i, ki) = V(i, ki) * remax;
    for (long k = ki + 1; k <= n; ++k) TT(k, k) = tdiag[k - 1];
  }
}

// ------------------------------- vectors.jl:133-362 (dtrevc-style right)
template <class R> void eigvecs_right_real(long n, const R* tp, long ldt, const R* zp, long ldz, cx<R>* vp,
                                           long ldv) {
  typedef cx<R> C;
  Mat<const R> T{tp, ldt};
  Mat<const R> Z{zp, ldz};
  Mat<C> V{vp, ldv};
  const R ulp = rt<R>::eps();
  const R smallnum = safemin<R>() * (R(double(n)) / ulp);
  const R bignum = (R(1.0) - ulp) / smallnum;
  for (long j = 1; j <= n; ++j)
    for (long i = 1; i <= n; ++i) V(i, j) = C();
  const bool have_z = (zp != nullptr);
  std::vector<C> v(n);
  std::vector<R> tnorms(n, R(0.0));
  for (long j = 2; j <= n; ++j)
    for (long i = 1; i <= j - 1; ++i) tnorms[j - 1] = tnorms[j - 1] + rt<R>::abs(T(i, j));
  int ip = 0;
  long isel = n;
  const R one(1.0);
  for (long ki = n; ki >= 1; --ki) {
    if (ip == 1) {
      ip = 0;
      if (have_z) {
        for (long i = 1; i <= n; ++i) V(i, ki) = conj(V(i, ki + 1));
      } else {
        for (long i = 1; i <= n; ++i) V(i, isel + 1) = conj(V(i, isel + 2));
      }
      continue;
    }
    if (ki > 1 && T(ki, ki - 1) != R(0.0)) ip = -1;
    R wr = T(ki, ki);
    R wi = (ip == 0) ? R(0.0) : rt<R>::sqrt(rt<R>::abs(T(ki, ki - 1))) * rt<R>::sqrt(rt<R>::abs(T(ki - 1, ki)));
    R smin = rmax(ulp * (rt<R>::abs(wr) + rt<R>::abs(wi)), smallnum);
    if (ip == 0) {
      v[ki - 1] = C(one);
      for (long i = 1; i <= ki - 1; ++i) v[i - 1] = C(-T(i, ki));
      long jnxt = ki - 1;
      for (long j = ki - 1; j >= 1; --j) {
        if (j > jnxt) continue;
        long j1 = j, j2 = j;
        if (j > 1 && T(j, j - 1) != R(0.0)) {
          j1 = j - 1;
          jnxt = j - 2;
        }
        if (j1 == j2) {
          C vt1[1] = {v[j - 1]};
          R Ae[1] = {T(j, j)};
          R t1;
          R vscale = xsolve<R>(one, 1, Ae, C(wr), vt1, smin, t1);
          if (t1 > one && tnorms[j - 1] > bignum / t1) {
            vt1[0] = vt1[0] / t1;
            vscale = vscale / t1;
          }
          if (vscale != one)
            for (long i = 1; i <= ki; ++i) v[i - 1] = v[i - 1] * vscale;
          v[j - 1] = vt1[0];
          for (long i = 1; i <= j - 1; ++i) v[i - 1] = v[i - 1] - vt1[0] * T(i, j);
        } else {
          C vt2[2] = {v[j - 2], v[j - 1]};
          R Ae[4] = {T(j - 1, j - 1), T(j, j - 1), T(j - 1, j), T(j, j)};
          R t1;
          R vscale = xsolve<R>(one, 2, Ae, C(wr), vt2, smin, t1);
          if (t1 > one) {
            R beta = rmax(tnorms[j - 2], tnorms[j - 1]);
            if (beta > bignum / t1) {
              vt2[0] = vt2[0] / t1;
              vt2[1] = vt2[1] / t1;
              vscale = vscale / t1;
            }
          }
          if (vscale != one)
            for (long i = 1; i <= ki; ++i) v[i - 1] = v[i - 1] * vscale;
          v[j - 2] = vt2[0];
          v[j - 1] = vt2[1];
          for (long i = 1; i <= j - 2; ++i) v[i - 1] = (v[i - 1] - vt2[0] * T(i, j - 1) - vt2[1] * T(i, j));
        }
      }
      if (!have_z) {
        R m(0.0);
        for (long i = 1; i <= ki; ++i) m = rmax(m, absv(v[i - 1]));
        R vscale = one / m;
        for (long i = 1; i <= ki; ++i) V(i, isel) = vscale * v[i - 1];
      } else {
        if (ki == 1) {
          for (long i = 1; i <= n; ++i) V(i, ki) = C(Z(i, ki));
        } else {
          for (long i = 1; i <= n; ++i) {
            C s;
            for (long k = 1; k <= ki; ++k) s = s + Z(i, k) * v[k - 1];
            V(i, ki) = s;
          }
          R m(0.0);
          for (long i = 1; i <= n; ++i) m = rmax(m, absv(V(i, ki)));
          R vscale = one / m;
          for (long i = 1; i <= n; ++i) V(i, ki) = V(i, ki) * vscale;
        }
      }
    } else {
      if (rt<R>::abs(T(ki - 1, ki)) >= rt<R>::abs(T(ki, ki - 1))) {
        v[ki - 2] = C(one);
        v[ki - 1] = C(R(0.0), wi / T(ki - 1, ki));
      } else {
        v[ki - 2] = C(-wi / T(ki, ki - 1));
        v[ki - 1] = C(R(0.0), one);
      }
      for (long i = 1; i <= ki - 2; ++i)
        v[i - 1] = C(-v[ki - 2].re * T(i, ki - 1), -(v[ki - 1].im * T(i, ki)));
      long jnxt = ki - 2;
      for (long j = ki - 2; j >= 1; --j) {
        if (j > jnxt) continue;
        long j1 = j, j2 = j;
        jnxt = j - 1;
        if (j > 1 && T(j, j - 1) != R(0.0)) {
          j1 = j - 1;
          jnxt = j - 2;
        }
        if (j1 == j2) {
          C vt1[1] = {v[j - 1]};
          R Ae[1] = {T(j, j)};
          R t1;
          R vscale = xsolve<R>(one, 1, Ae, C(wr, wi), vt1, smin, t1);
          if (t1 > one && tnorms[j - 1] > bignum / t1) {
            vt1[0] = vt1[0] / t1;
            vscale = vscale / t1;
          }
          if (vscale != one)
            for (long i = 1; i <= ki; ++i) v[i - 1] = v[i - 1] * vscale;
          v[j - 1] = vt1[0];
          for (long i = 1; i <= j - 1; ++i) v[i - 1] = v[i - 1] - vt1[0] * T(i, j);
        } else {
          C vt2[2] = {v[j - 2], v[j - 1]};
          R Ae[4] = {T(j - 1, j - 1), T(j, j - 1), T(j - 1, j), T(j, j)};
          R t1;
          R vscale = xsolve<R>(one, 2, Ae, C(wr, wi), vt2, smin, t1);
          if (t1 > one) {
            R beta = rmax(tnorms[j - 2], tnorms[j - 1]);
            if (beta > bignum / t1) {
              vt2[0] = vt2[0] / t1;
              vt2[1] = vt2[1] / t1;
              vscale = vscale / t1;
            }
          }
          if (vscale != one)
            for (long i = 1; i <= ki; ++i) v[i - 1] = v[i - 1] * vscale;
          v[j - 2] = vt2[0];
          v[j - 1] = vt2[1];
          for (long i = 1; i <= j - 2; ++i) v[i - 1] = (v[i - 1] - vt2[0] * T(i, j - 1) - vt2[1] * T(i, j));
        }
      }
      if (!have_z) {
        for (long i = 1; i <= ki; ++i) V(i, isel) = conj(v[i - 1]);
        R t1(0.0);
        for (long k = 1; k <= ki; ++k) t1 = rmax(t1, abs1(V(k, isel)));
        t1 = one / t1;
        for (long i = 1; i <= ki; ++i) V(i, isel) = V(i, isel) * t1;
      } else {
        for (long i = 1; i <= n; ++i) {
          C s;
          for (long k = 1; k <= ki; ++k) s = s + Z(i, k) * v[k - 1];
          V(i, ki) = s;
        }
        R m(0.0);
        for (long i = 1; i <= n; ++i) m = rmax(m, absv(V(i, ki)));
        R vscale = one / m;
        for (long i = 1; i <= n; ++i) V(i, ki) = vscale * conj(V(i, ki));
      }
    }
    isel -= (ip == 0) ? 1 : 2;
    if (ip == -1) ip = 1;
  }
}

// -------------------------------- vectors.jl:462-698 (dtrevc-style left)
template <class R> void eigvecs_left_real(long n, const R* tp, long ldt, const R* zp, long ldz, cx<R>* vp,
                                          long ldv) {
  typedef cx<R> C;
  Mat<const R> T{tp, ldt};
  Mat<const R> Z{zp, ldz};
  Mat<C> V{vp, ldv};
  const R ulp = rt<R>::eps();
  const R smallnum = safemin<R>() * (R(double(n)) / ulp);
  const R bignum = (R(1.0) - ulp) / smallnum;
  for (long j = 1; j <= n; ++j)
    for (long i = 1; i <= n; ++i) V(i, j) = C();
  const bool have_z = (zp != nullptr);
  std::vector<C> v(n);
  std::vector<R> tnorms(n, R(0.0));
  for (long j = 2; j <= n; ++j)
    for (long i = 1; i <= j - 1; ++i) tnorms[j - 1] = tnorms[j - 1] + rt<R>::abs(T(i, j));
  int ip = 0;
  long isel = 1;
  const R one(1.0);
  for (long ki = 1; ki <= n; ++ki) {
    if (ip == -1) {
      ip = 0;
      if (have_z) {
        for (long i = 1; i <= n; ++i) V(i, ki) = conj(V(i, ki - 1));
      } else {
        for (long i = 1; i <= n; ++i) V(i, isel - 1) = conj(V(i, isel - 2));
      }
      continue;
    }
    if (ki < n && T(ki + 1, ki) != R(0.0)) ip = 1;
    R wr = T(ki, ki);
    R wi = (ip == 0) ? R(0.0) : rt<R>::sqrt(rt<R>::abs(T(ki, ki + 1))) * rt<R>::sqrt(rt<R>::abs(T(ki + 1, ki)));
    R smin = rmax(ulp * (rt<R>::abs(wr) + rt<R>::abs(wi)), smallnum);
    if (ip == 0) {
      v[ki - 1] = C(one);
      for (long i = ki + 1; i <= n; ++i) v[i - 1] = C(-T(ki, i));
      R vmax = one, vcrit = bignum;
      long jnxt = ki + 1;
      for (long j = ki + 1; j <= n; ++j) {
        if (j < jnxt) continue;
        long j1 = j, j2 = j;
        jnxt = j + 1;
        if (j < n && T(j + 1, j) != R(0.0)) {
          j2 = j + 1;
          jnxt = j + 2;
        }
        if (j1 == j2) {
          if (tnorms[j - 1] > vcrit) {
            R t1 = one / vmax;
            for (long i = ki; i <= n; ++i) v[i - 1] = v[i - 1] * t1;
            vmax = one;
            vcrit = bignum;
          }
          C d;
          for (long i = ki + 1; i <= j - 1; ++i) d = d + T(i, j) * v[i - 1];
          C vt1[1] = {v[j - 1] - d};
          R Ae[1] = {T(j, j)};
          R xn;
          R vscale = xsolve<R>(one, 1, Ae, C(wr), vt1, smin, xn);
          if (vscale != one)
            for (long i = ki; i <= n; ++i) v[i - 1] = v[i - 1] * vscale;
          v[j - 1] = vt1[0];
          vmax = rmax(absv(v[j - 1]), vmax);
          vcrit = bignum / vmax;
        } else {
          R beta = rmax(tnorms[j], tnorms[j - 1]);
          if (beta > vcrit) {
            R t1 = one / vmax;
            for (long i = ki; i <= n; ++i) v[i - 1] = v[i - 1] * t1;
            vmax = one;
            vcrit = bignum;
          }
          C d0, d1;
          for (long i = ki + 1; i <= j - 1; ++i) {
            d0 = d0 + T(i, j) * v[i - 1];
            d1 = d1 + T(i, j + 1) * v[i - 1];
          }
          C vt2[2] = {v[j - 1] - d0, v[j] - d1};
          // transpose of the 2x2 diagonal block
          R Ae[4] = {T(j, j), T(j, j + 1), T(j + 1, j), T(j + 1, j + 1)};
          R xn;
          R vscale = xsolve<R>(one, 2, Ae, C(wr), vt2, smin, xn);
          if (vscale != one)
            for (long i = ki; i <= n; ++i) v[i - 1] = v[i - 1] * vscale;
          v[j - 1] = vt2[0];
          v[j] = vt2[1];
          vmax = rmax3(absv(v[j - 1]), absv(v[j]), vmax);
          vcrit = bignum / vmax;
        }
      }
      if (!have_z) {
        R m(0.0);
        for (long i = ki; i <= n; ++i) m = rmax(m, absv(v[i - 1]));
        R vscale = one / m;
        for (long i = ki; i <= n; ++i) V(i, isel) = vscale * v[i - 1];
      } else {
        if (ki == n) {
          for (long i = 1; i <= n; ++i) V(i, ki) = C(Z(i, ki));
        } else {
          for (long i = 1; i <= n; ++i) {
            C s;
            for (long k = ki; k <= n; ++k) s = s + Z(i, k) * v[k - 1];
            V(i, ki) = s;
          }
          R m(0.0);
          for (long i = 1; i <= n; ++i) m = rmax(m, absv(V(i, ki)));
          R vscale = one / m;
          for (long i = 1; i <= n; ++i) V(i, ki) = V(i, ki) * vscale;
        }
      }
    } else {
      if (rt<R>::abs(T(ki, ki + 1)) >= rt<R>::abs(T(ki + 1, ki))) {
        v[ki - 1] = C(wi / T(ki, ki + 1));
        v[ki] = C(R(0.0), one);
      } else {
        v[ki - 1] = C(one);
        v[ki] = C(R(0.0), -wi / T(ki + 1, ki));
      }
      for (long i = ki + 2; i <= n; ++i) v[i - 1] = C(-v[ki - 1].re * T(ki, i), -(v[ki].im * T(ki + 1, i)));
      R vmax = one, vcrit = bignum;
      long jnxt = ki + 2;
      for (long j = ki + 1; j <= n; ++j) {
        if (j < jnxt) continue;
        long j1 = j, j2 = j;
        jnxt = j + 1;
        if (j < n && T(j + 1, j) != R(0.0)) {
          j1 = j + 1;  // vectors.jl:627 sets j1 (not j2); kept as in the reference
          jnxt = j + 2;
        }
        if (j1 == j2) {
          if (tnorms[j - 1] > vcrit) {
            R t1 = one / vmax;
            for (long i = ki; i <= n; ++i) v[i - 1] = v[i - 1] * t1;
            vmax = one;
            vcrit = bignum;
          }
          C d;
          for (long i = ki + 2; i <= j - 1; ++i) d = d + T(i, j) * v[i - 1];
          C vt1[1] = {v[j - 1] - d};
          R Ae[1] = {T(j, j)};
          R xn;
          R vscale = xsolve<R>(one, 1, Ae, C(wr, -wi), vt1, smin, xn);
          if (vscale != one)
            for (long i = ki; i <= n; ++i) v[i - 1] = v[i - 1] * vscale;
          v[j - 1] = vt1[0];
          vmax = rmax(abs1(v[j - 1]), vmax);
          vcrit = bignum / vmax;
        } else {
          // with j1 = j+1 the overflow guard reads tnorms[j-1] (vectors.jl:653)
          R beta = rmax(tnorms[j - 2], tnorms[j - 1]);
          if (beta > vcrit) {
            R t1 = one / vmax;
            for (long i = ki; i <= n; ++i) v[i - 1] = v[i - 1] * t1;
            vmax = one;
            vcrit = bignum;
          }
          C d0, d1;
          for (long i = ki + 2; i <= j - 1; ++i) {
            d0 = d0 + T(i, j) * v[i - 1];
            d1 = d1 + T(i, j + 1) * v[i - 1];
          }
          C vt2[2] = {v[j - 1] - d0, v[j] - d1};
          R Ae[4] = {T(j, j), T(j, j + 1), T(j + 1, j), T(j + 1, j + 1)};
          R xn;
          R vscale = xsolve<R>(one, 2, Ae, C(wr, -wi), vt2, smin, xn);
          if (vscale != one)
            for (long i = ki; i <= n; ++i) v[i - 1] = v[i - 1] * vscale;
          v[j - 1] = vt2[0];
          v[j] = vt2[1];
          vmax = rmax3(abs1(vt2[0]), abs1(vt2[1]), vmax);
          vcrit = bignum / vmax;
        }
      }
      if (!have_z) {
        for (long i = ki; i <= n; ++i) V(i, isel) = v[i - 1];
        R t1(0.0);
        for (long k = ki; k <= n; ++k) t1 = rmax(t1, abs1(V(k, isel)));
        t1 = one / t1;
        for (long i = ki; i <= n; ++i) V(i, isel) = V(i, isel) * t1;
      } else {
        for (long i = 1; i <= n; ++i) {
          C s;
          for (long k = ki; k <= n; ++k) s = s + Z(i, k) * v[k - 1];
          V(i, ki) = s;
        }
        R m(0.0);
        for (long i = 1; i <= n; ++i) m = rmax(m, absv(V(i, ki)));
        R vscale = one / m;
        for (long i = 1; i <= n; ++i) V(i, ki) = V(i, ki) * vscale;
      }
    }
    isel += (ip == 0) ? 1 : 2;
    if (ip == 1) ip = -1;
  }
}

// ---------------------------------------------------- triang.jl:9-43
// Real standard-form Schur (T, Z) -> complex triangular (Tc, Zc).
template <class R> void triangularize(long n, const R* tr, long ldt, const R* zr, long ldz, cx<R>* tc, cx<R>* zc) {
  typedef cx<R> C;
  Mat<const R> Tr{tr, ldt};
  Mat<C> T{tc, n};
  Mat<C> Z{zc, n};
  for (long j = 1; j <= n; ++j)
    for (long i = 1; i <= n; ++i) {
      T(i, j) = C(Tr(i, j));
      Z(i, j) = C(zr[(i - 1) + (j - 1) * ldz]);
    }
  for (long j = n; j >= 2; --j) {
    if (Tr(j, j - 1) != R(0.0)) {
      R s = rt<R>::sqrt(rt<R>::abs(Tr(j, j - 1)));
      R c = rt<R>::sqrt(rt<R>::abs(Tr(j - 1, j)));
      R r = rt<R>::hypot(s, c);
      C gc = C(c / r);            // Givens c (complex-typed, real valued)
      C gs = C(R(0.0), -(s / r));  // Givens s = -im*(s/r)
      // lmul!(G, T): rows j-1, j
      for (long k = 1; k <= n; ++k) {
        C a1 = T(j - 1, k), a2 = T(j, k);
        T(j - 1, k) = gc * a1 + gs * a2;
        T(j, k) = -conj(gs) * a1 + gc * a2;
      }
      // rmul!(T, G') and rmul!(Z, G'): columns j-1, j times G^H
      for (long k = 1; k <= n; ++k) {
        C a1 = T(k, j - 1), a2 = T(k, j);
        T(k, j - 1) = a1 * conj(gc) + a2 * conj(gs);
        T(k, j) = -a1 * gs + a2 * conj(gc);
        C b1 = Z(k, j - 1), b2 = Z(k, j);
        Z(k, j - 1) = b1 * conj(gc) + b2 * conj(gs);
        Z(k, j) = -b1 * gs + b2 * conj(gc);
      }
    }
  }
  for (long j = 1; j <= n; ++j)
    for (long i = j + 1; i <= n; ++i) T(i, j) = C();
}

// ------------------------------------------------- balance.jl:33-199
// balance!(A; scale, permute) with p = 1, algo = :pr (xGEBAL's scheme).
// On exit a holds the balanced matrix, ilo/ihi (1-based) the bounds of the
// unreduced block, sp[0..n-1] the permutation record (sp[ms-1] = js), d the
// diagonal scale factors.  Returns 0, or -3 when a NaN was met while
// scaling (the reference's error("NaN encountered while balancing")).
template <class R> R bal_abs2max(long m, const R* x, long inc) {
  R ax0 = rt<R>::abs(x[0]);  // norm1est.jl:76-87
  for (long i = 1; i < m; ++i) {
    R ax = rt<R>::abs(x[i * inc]);
    if (ax > ax0) ax0 = ax;
  }
  return ax0;
}
template <class R> R bal_abs2max(long m, const cx<R>* x, long inc) {
  R ax0 = abs2(x[0]);  // norm1est.jl:89-100
  for (long i = 1; i < m; ++i) {
    R ax = abs2(x[i * inc]);
    if (ax > ax0) ax0 = ax;
  }
  return rt<R>::sqrt(ax0);
}
// Julia's max / min propagate a NaN operand
template <class R> R jmax(const R& a, const R& b) { return (rt<R>::isnan(a) || rt<R>::isnan(b)) ? a + b : rmax(a, b); }
template <class R> R jmin(const R& a, const R& b) { return (rt<R>::isnan(a) || rt<R>::isnan(b)) ? a + b : rmin(a, b); }
template <class T> int balance(long n, T* a, long lda, bool scale, bool permute, long& ilo, long& ihi, long* sp,
                               typename real_of<T>::type* d, bool& trivial) {
  typedef typename real_of<T>::type R;
  Mat<T> A{a, lda};
  const R beta = R(2.0);
  const R factor = R(0.95);
  const R sfmin1 = rt<R>::fmin() / rt<R>::eps();
  const R sfmin2 = sfmin1 * beta;
  const R sfmax2 = R(1.0) / sfmin2;
  for (long i = 0; i < n; ++i) {
    d[i] = R(1.0);
    sp[i] = 0;
  }
  ilo = 1;
  ihi = n;
  trivial = true;
  auto swap_cols_rows = [&](long js, long ms, long rlast, long cfirst) {
    for (long i = 1; i <= rlast; ++i) std::swap(A(i, js), A(i, ms));
    for (long i = cfirst; i <= n; ++i) std::swap(A(js, i), A(ms, i));
  };
  if (permute) {
    ihi = n + 1;
    long js = 0;
    while (ihi > 1) {
      ihi -= 1;
      bool exch = false;
      long ms = 0;
      for (long j = ihi; j >= 1; --j) {
        exch = true;
        js = j;
        for (long i = 1; i <= ihi; ++i) {
          if (i == j) continue;
          if (!(A(j, i) == T(R(0.0)))) exch = false;
        }
        if (exch) {
          ms = ihi;
          break;
        }
      }
      if (exch) {
        sp[ms - 1] = js;
        if (js != ms) {
          trivial = false;
          swap_cols_rows(js, ms, ihi, ilo);
        }
      } else {
        break;
      }
    }
    if (ihi > 1) {
      ilo = 0;
      while (ilo < n) {
        ilo += 1;
        bool exch = false;
        long ms = 0;
        js = 0;
        for (long j = ilo; j <= ihi; ++j) {
          js = j;
          exch = true;
          for (long i = ilo; i <= ihi; ++i) {
            if (i == j) continue;
            if (!(A(i, j) == T(R(0.0)))) exch = false;
          }
          if (exch) {
            ms = ilo;
            break;
          }
        }
        if (exch) {
          sp[ms - 1] = js;
          if (ms != js) {
            trivial = false;
            swap_cols_rows(js, ms, ihi, ilo);
          }
        } else {
          break;
        }
      }
    }
  }
  if (scale) {
    bool converged = false;
    while (!converged) {
      converged = true;
      for (long i = ilo; i <= ihi; ++i) {
        const long m = ihi - ilo + 1;
        R c = norm2(m, &A(ilo, i), 1);
        R r = norm2(m, &A(i, ilo), lda);
        R ca = bal_abs2max(m, &A(ilo, i), 1);
        R ra = bal_abs2max(m, &A(i, ilo), lda);
        // a NaN norm would make the reference's sweep loop run forever; it
        // is reported as the reference's NaN error instead
        if (rt<R>::isnan(c) || rt<R>::isnan(r)) return -3;
        if (c == R(0.0) || r == R(0.0)) continue;
        R g = r / beta;
        R sv = c + r;
        R f = R(1.0);
        while (c < r / beta) {
          if (c >= g || jmax(f, jmax(c, ca)) >= sfmax2 || jmin(r, jmin(g, ra)) <= sfmin2)
            break;
          if (rt<R>::isnan(c + f + ca + r + g + ra)) return -3;
          f *= beta;
          c *= beta;
          ca *= beta;
          r /= beta;
          g /= beta;
          ra /= beta;
        }
        g = c / beta;
        while (r <= c / beta) {
          if (g < r || jmax(r, ra) >= sfmax2 || jmin(jmin(f, c), jmin(g, ca)) <= sfmin2) break;
          f /= beta;
          c /= beta;
          g /= beta;
          ca /= beta;
          r *= beta;
          ra *= beta;
        }
        if (f != R(1.0)) trivial = false;
        if (c + r >= factor * sv) continue;
        converged = false;
        d[i - 1] *= f;
        const R fi = R(1.0) / f;
        for (long k = ilo; k <= n; ++k) A(i, k) = A(i, k) * fi;
        for (long k = 1; k <= ihi; ++k) A(k, i) = A(k, i) * f;
      }
    }
  }
  return 0;
}

// ----------------------------------------------- balance.jl:201-261
// lmul!(B, V) (right eigenvectors, inverse = false) or ldiv!(B, V) (left
// eigenvectors, inverse = true) on the n x m columns of V.
template <class R> void balance_apply(long n, long m, long ilo, long ihi, const long* sp, const R* d, bool trivial,
                                      cx<R>* v, long ldv, bool inverse) {
  if (trivial) return;
  Mat<cx<R>> V{v, ldv};
  if (ilo != ihi)
    for (long j = 1; j <= m; ++j)
      for (long i = 1; i <= n; ++i) V(i, j) = V(i, j) * (inverse ? R(1.0) / d[i - 1] : d[i - 1]);
  for (long j = ilo - 1; j >= 1; --j) {
    const long mm = sp[j - 1];
    if (mm == j) continue;
    for (long i = 1; i <= m; ++i) std::swap(V(j, i), V(mm, i));
  }
  for (long j = ihi + 1; j <= n; ++j) {
    const long mm = sp[j - 1];
    if (mm == j) continue;
    for (long i = 1; i <= m; ++i) std::swap(V(j, i), V(mm, i));
  }
}

// ======================================== reordering and conditioning
// ordschur.jl:23-73 (gordschur!, _trexchange!), sylvester.jl:24-139
// (trsylvester!, adjtrsylvester!), norm1est.jl:21-100 (norm1est!,
// _findamax), ordschur.jl:75-131 (eigvalscond, subspacesep); complex
// Schur forms.

// Julia's floatmin2: 2^trunc(log2(floatmin / eps) / 2) (2^-485 for Float64)
template <class R> R floatmin2() {
  R x(1.0);
  const R lim = rt<R>::fmin() / rt<R>::eps();
  // the exponent of floatmin / eps, halved toward zero
  int e = 0;
  R y(1.0);
  while (y > lim) {
    y = y * R(0.5);
    ++e;
  }
  for (int i = 0; i < e / 2; ++i) x = x * R(0.5);
  return x;
}

// LinearAlgebra.givensAlgorithm(f, g) for complex arguments (the xLARTG
// scheme of LAPACK 3.x with power-of-two scaling; its abs1 is
// max(|re|, |im|)): cs real, sn complex, r = cs f + sn g.
template <class R> void givens_cx(cx<R> f, cx<R> g, R& cs, cx<R>& sn, cx<R>& r) {
  typedef cx<R> C;
  auto am = [](const C& z) { return rmax(rt<R>::abs(z.re), rt<R>::abs(z.im)); };
  const R one(1.0);
  const R safmin = rt<R>::fmin();
  const R safmn2 = floatmin2<R>();
  const R safmx2 = one / safmn2;
  R scalepar = rmax(am(f), am(g));
  C fs = f, gs = g;
  int count = 0;
  if (scalepar >= safmx2) {
    while (true) {
      count += 1;
      fs = fs * safmn2;
      gs = gs * safmn2;
      scalepar = scalepar * safmn2;
      if (scalepar < safmx2 || count >= 20) break;
    }
  } else if (scalepar <= safmn2) {
    if (g == C()) {
      cs = one;
      sn = C();
      r = f;
      return;
    }
    while (true) {
      count -= 1;
      fs = fs * safmx2;
      gs = gs * safmx2;
      scalepar = scalepar * safmx2;
      if (scalepar > safmn2) break;
    }
  }
  const R f2 = abs2(fs), g2 = abs2(gs);
  if (f2 <= rmax(g2, one) * safmin) {
    // F is very small
    if (f == C()) {
      cs = R(0.0);
      r = C(rt<R>::hypot(g.re, g.im));
      const R d = rt<R>::hypot(gs.re, gs.im);
      sn = C(gs.re / d, -gs.im / d);
      return;
    }
    const R f2s = rt<R>::hypot(fs.re, fs.im);
    const R g2s = rt<R>::sqrt(g2);
    cs = f2s / g2s;
    C ff;
    if (am(f) > one) {
      const R d = rt<R>::hypot(f.re, f.im);
      ff = C(f.re / d, f.im / d);
    } else {
      const R dr = safmx2 * f.re, di = safmx2 * f.im;
      const R d = rt<R>::hypot(dr, di);
      ff = C(dr / d, di / d);
    }
    sn = ff * C(gs.re / g2s, -gs.im / g2s);
    r = cs * f + sn * g;
  } else {
    // the common case
    const R f2s = rt<R>::sqrt(one + g2 / f2);
    r = C(f2s * fs.re, f2s * fs.im);
    cs = one / f2s;
    const R d = f2 + g2;
    sn = C(r.re / d, r.im / d);
    sn = sn * conj(gs);
    if (count > 0)
      for (int i = 0; i < count; ++i) r = r * safmx2;
    else
      for (int i = 0; i < -count; ++i) r = r * safmn2;
  }
}

// _trexchange!(T, Z, iold, inew) (ordschur.jl:52-73; Golub & Van Loan
// Algorithm 7.6.1): adjacent swaps by Givens rotations, lmul!(G, T),
// rmul!(T, G'), rmul!(Z, G').
template <class R> void trexchange(long n, cx<R>* tp, long ldt, cx<R>* zp, long ldz, bool wantz, long iold,
                                   long inew) {
  typedef cx<R> C;
  Mat<C> T{tp, ldt};
  Mat<C> Z{zp, ldz};
  if (n < 1 || iold == inew) return;
  const long step = iold < inew ? 1 : -1;
  const long k0 = iold < inew ? iold : iold - 1;
  const long k1 = iold < inew ? inew - 1 : inew;
  for (long k = k0;; k += step) {
    const C Tkk = T(k, k), Tpp = T(k + 1, k + 1);
    R c;
    C sn, r;
    givens_cx<R>(T(k, k + 1), Tpp - Tkk, c, sn, r);
    for (long j = 1; j <= n; ++j) {
      const C a1 = T(k, j), a2 = T(k + 1, j);
      T(k, j) = c * a1 + sn * a2;
      T(k + 1, j) = -conj(sn) * a1 + c * a2;
    }
    // G' = Givens(c, -sn): column k gets a1 c - a2 conj(-sn), column k+1
    // a1 (-sn) + a2 c
    const C ms = -sn;
    for (long i = 1; i <= n; ++i) {
      const C a1 = T(i, k), a2 = T(i, k + 1);
      T(i, k) = a1 * c - a2 * conj(ms);
      T(i, k + 1) = a1 * ms + a2 * c;
    }
    if (wantz)
      for (long i = 1; i <= n; ++i) {
        const C a1 = Z(i, k), a2 = Z(i, k + 1);
        Z(i, k) = a1 * c - a2 * conj(ms);
        Z(i, k + 1) = a1 * ms + a2 * c;
      }
    if (k == k1) break;
  }
}

// gordschur!(T, Z, select) (ordschur.jl:23-43): the selected eigenvalues to
// the front, in order; T made exactly upper triangular; w = diag(T).
template <class R> void gordschur(long n, cx<R>* tp, long ldt, cx<R>* zp, long ldz, bool wantz, const int* select,
                                  cx<R>* w) {
  typedef cx<R> C;
  Mat<C> T{tp, ldt};
  long ks = 0;
  for (long k = 1; k <= n; ++k)
    if (select[k - 1]) {
      ks += 1;
      if (k != ks) trexchange<R>(n, tp, ldt, zp, ldz, wantz, k, ks);
    }
  for (long j = 1; j <= n; ++j)
    for (long i = j + 1; i <= n; ++i) T(i, j) = C();
  for (long j = 1; j <= n; ++j) w[j - 1] = T(j, j);
}

// the entrywise max |a| of Julia's norm(A, Inf) for an m x m block
template <class R> R maxabs(long m, const cx<R>* ap, long lda) {
  R mx(0.0);
  for (long j = 0; j < m; ++j)
    for (long i = 0; i < m; ++i) mx = rmax(mx, absv(ap[i + j * lda]));
  return mx;
}

// trsylvester!(A, B, C) (sylvester.jl:24-75): A X - X B = scale C for
// upper triangular A (m x m) and B (n x n); C (m x n, ldc) overwritten by
// X.  Returns scale; perturbed set when a diagonal sum was replaced by smin.
template <class R> R trsylvester(long m, const cx<R>* ap, long lda, long n, const cx<R>* bp, long ldb, cx<R>* cp,
                                 long ldc, bool& perturbed) {
  typedef cx<R> C;
  Mat<const C> A{ap, lda};
  Mat<const C> B{bp, ldb};
  Mat<C> X{cp, ldc};
  R scale(1.0);
  const R tiny = rt<R>::eps();
  const R small = safemin<R>() * R((double)m) * R((double)n) / tiny;
  const R bignum = R(1.0) / small;
  const R smin = rmax(small, tiny * maxabs<R>(m, ap, lda) * maxabs<R>(n, bp, ldb));
  perturbed = false;
  for (long l = 1; l <= n; ++l)
    for (long k = m; k >= 1; --k) {
      C suml, sumr;
      for (long j = k + 1; j <= m; ++j) suml += A(k, j) * X(j, l);
      for (long j = 1; j <= l - 1; ++j) sumr += X(k, j) * B(j, l);
      const C v = X(k, l) - (suml - sumr);  // isgn = -1
      R scaloc(1.0);
      C a11 = A(k, k) - B(l, l);
      R da11 = abs1(a11);
      if (da11 <= smin) {
        a11 = C(smin);
        da11 = smin;
        perturbed = true;
      }
      const R db = abs1(v);
      if (da11 < R(1.0) && db > R(1.0) && db > bignum * da11) scaloc = R(1.0) / db;
      const C x11 = (v * scaloc) / a11;
      if (scaloc != R(1.0)) {
        for (long jj = 1; jj <= n; ++jj)
          for (long ii = 1; ii <= m; ++ii) X(ii, jj) = scaloc * X(ii, jj);
        scale = scale * scaloc;
      }
      X(k, l) = x11;
    }
  return scale;
}

// adjtrsylvester!(A, B, C) (sylvester.jl:86-139): A^H X - X B^H = scale C.
template <class R> R adjtrsylvester(long m, const cx<R>* ap, long lda, long n, const cx<R>* bp, long ldb, cx<R>* cp,
                                    long ldc, bool& perturbed) {
  typedef cx<R> C;
  Mat<const C> A{ap, lda};
  Mat<const C> B{bp, ldb};
  Mat<C> X{cp, ldc};
  R scale(1.0);
  const R tiny = rt<R>::eps();
  const R small = safemin<R>() * R((double)m) * R((double)n) / tiny;
  const R bignum = R(1.0) / small;
  const R smin = rmax(small, tiny * maxabs<R>(m, ap, lda) * maxabs<R>(n, bp, ldb));
  perturbed = false;
  for (long l = n; l >= 1; --l)
    for (long k = 1; k <= m; ++k) {
      C suml, sumr;
      for (long j = 1; j <= k - 1; ++j) suml += conj(A(j, k)) * X(j, l);
      for (long j = l + 1; j <= n; ++j) sumr += conj(X(k, j)) * B(l, j);
      const C v = X(k, l) - (suml - conj(sumr));  // isgn = -1
      R scaloc(1.0);
      C a11 = conj(A(k, k) - B(l, l));
      R da11 = abs1(a11);
      if (da11 < smin) {
        a11 = C(smin);
        da11 = smin;
        perturbed = true;
      }
      const R db = abs1(v);
      if (da11 < R(1.0) && db > R(1.0) && db > bignum * da11) scaloc = R(1.0) / db;
      const C x11 = (v * scaloc) / a11;
      if (scaloc != R(1.0)) {
        for (long jj = 1; jj <= n; ++jj)
          for (long ii = 1; ii <= m; ++ii) X(ii, jj) = scaloc * X(ii, jj);
        scale = scale * scaloc;
      }
      X(k, l) = x11;
    }
  return scale;
}

// _findamax (norm1est.jl:86-100), complex: sqrt of the first largest |x|^2
template <class R> R findamax(long n, const cx<R>* x, long& i0) {
  R ax0 = abs2(x[0]);
  i0 = 0;
  for (long i = 1; i < n; ++i) {
    const R ax = abs2(x[i]);
    if (ax > ax0) {
      ax0 = ax;
      i0 = i;
    }
  }
  return rt<R>::sqrt(ax0);
}

template <class R> R norm1(long n, const cx<R>* x) {
  R s(0.0);
  for (long i = 0; i < n; ++i) s = s + absv(x[i]);
  return s;
}

// norm1est!(applyA!, applyAH!, y) (norm1est.jl:21-75), Hager's estimator
// with Higham's modifications; the operators act in place on x.
template <class R, class FA, class FH> R norm1est(long n, FA applyA, FH applyAH, int maxiter = 5) {
  typedef cx<R> C;
  std::vector<C> x((size_t)n, C(R(1.0) / R((double)n))), y((size_t)n);
  applyA(x.data());
  if (n == 1) return absv(x[0]);
  R gamma = norm1<R>(n, x.data());
  const R tiny = safemin<R>();
  auto sgn = [&]() {
    for (long i = 0; i < n; ++i) {
      const R a = absv(x[i]);
      x[i] = a > tiny ? x[i] / a : C(R(1.0));
    }
  };
  sgn();
  applyAH(x.data());
  long j0;
  R ax0 = findamax<R>(n, x.data(), j0);
  for (int iter = 1; iter <= maxiter; ++iter) {
    for (long i = 0; i < n; ++i) x[i] = C();
    x[j0] = C(R(1.0));
    applyA(x.data());
    y = x;
    const R oldg = gamma;
    gamma = norm1<R>(n, y.data());
    if (gamma <= oldg) break;
    sgn();
    applyAH(x.data());
    const long jlast = j0;
    ax0 = findamax<R>(n, x.data(), j0);
    if (absv(x[jlast]) == ax0) break;
  }
  // alternative estimate (Higham 1990)
  R isign(1.0);
  for (long i = 0; i < n; ++i) {
    x[i] = C(isign * (R(1.0) + R((double)i) / R((double)(n - 1))));
    isign = -isign;
  }
  applyA(x.data());
  const R t = R(2.0) * norm1<R>(n, x.data()) / (R(3.0) * R((double)n));
  return rmax(t, gamma);
}

// eigvalscond(S, nsub) (ordschur.jl:85-98) and subspacesep(S, nsub)
// (ordschur.jl:111-131) of the upper triangular T (n x n); 0 < nsub < n.
template <class R> R eigvalscond(long n, const cx<R>* tp, long ldt, long nsub) {
  typedef cx<R> C;
  const long m = nsub, n2 = n - nsub;
  std::vector<C> X((size_t)(m * n2));
  for (long j = 0; j < n2; ++j)
    for (long i = 0; i < m; ++i) X[i + j * m] = tp[i + (m + j) * ldt];
  bool pert;
  const R scale = trsylvester<R>(m, tp, ldt, n2, tp + m + m * ldt, ldt, X.data(), m, pert);
  R rn(0.0);
  for (const C& v : X) rn = rn + abs2(v);
  rn = rt<R>::sqrt(rn);
  return rn > R(0.0) ? scale / (rt<R>::sqrt(scale * scale / rn + rn) * rt<R>::sqrt(rn)) : R(1.0);
}

template <class R> R subspacesep(long n, const cx<R>* tp, long ldt, long nsub) {
  const long m = nsub, n2 = n - nsub;
  R scale(1.0);
  bool pert;
  auto fA = [&](cx<R>* x) { scale = trsylvester<R>(m, tp, ldt, n2, tp + m + m * ldt, ldt, x, m, pert); };
  auto fH = [&](cx<R>* x) { scale = adjtrsylvester<R>(m, tp, ldt, n2, tp + m + m * ldt, ldt, x, m, pert); };
  const R est = norm1est<R>(m * n2, fA, fH);
  return scale / est;
}


// ===================================================================
// Hermitian / real symmetric path: _hehessenberg! (hessenberg.jl:19-95)
// with rank2update! (:97-130), _materializeQ (:133-232), and the
// symmetric tridiagonal QL/QR iteration (symtridiag.jl:14-427).
// ===================================================================

// LinearAlgebra.givensAlgorithm(f, g) for real arguments (LAPACK 3.1 xLARTG:
// safmn2 = floatmin2 scaling, and the sign fix when |f| > |g| and cs < 0).
template <class R> void givens_re(R f, R g, R& cs, R& sn, R& r) {
  const R safmn2 = floatmin2<R>();
  const R safmx2 = R(1.0) / safmn2;
  if (g == R(0.0)) {
    cs = R(1.0);
    sn = R(0.0);
    r = f;
    return;
  }
  if (f == R(0.0)) {
    cs = R(0.0);
    sn = R(1.0);
    r = g;
    return;
  }
  R f1 = f, g1 = g;
  R sc = rmax(rt<R>::abs(f1), rt<R>::abs(g1));
  if (sc >= safmx2) {
    int count = 0;
    while (true) {
      count += 1;
      f1 = f1 * safmn2;
      g1 = g1 * safmn2;
      sc = rmax(rt<R>::abs(f1), rt<R>::abs(g1));
      if (sc < safmx2 || count >= 20) break;
    }
    r = rt<R>::sqrt(f1 * f1 + g1 * g1);
    cs = f1 / r;
    sn = g1 / r;
    for (int i = 0; i < count; ++i) r = r * safmx2;
  } else if (sc <= safmn2) {
    int count = 0;
    while (true) {
      count += 1;
      f1 = f1 * safmx2;
      g1 = g1 * safmx2;
      sc = rmax(rt<R>::abs(f1), rt<R>::abs(g1));
      if (sc > safmn2) break;
    }
    r = rt<R>::sqrt(f1 * f1 + g1 * g1);
    cs = f1 / r;
    sn = g1 / r;
    for (int i = 0; i < count; ++i) r = r * safmn2;
  } else {
    r = rt<R>::sqrt(f1 * f1 + g1 * g1);
    cs = f1 / r;
    sn = g1 / r;
  }
  if (rt<R>::abs(f) > rt<R>::abs(g) && cs < R(0.0)) {
    cs = -cs;
    sn = -sn;
    r = -r;
  }
}

// rank2update!(A, -1, x, y, Val(:L)) (hessenberg.jl:115-130) on the m x m
// block at a: A <- A - x y^H - y x^H over the lower triangle, diagonal kept
// real.
template <class T> void rank2update_l(long m, T* a, long lda, const T* x, const T* y) {
  Mat<T> A{a, lda};
  for (long j = 1; j <= m; ++j) {
    if (x[j - 1] == T() && y[j - 1] == T()) {
      A(j, j) = T(real(A(j, j)));
      continue;
    }
    const T t1 = -conj(y[j - 1]);
    const T t2 = conj(-x[j - 1]);
    A(j, j) = T(real(A(j, j)) + real(x[j - 1] * t1 + y[j - 1] * t2));
    for (long i = j + 1; i <= m; ++i) A(i, j) = A(i, j) + (x[i - 1] * t1 + y[i - 1] * t2);
  }
}
// rank2update!(A, -1, x, y, Val(:U)) (hessenberg.jl:97-113)
template <class T> void rank2update_u(long m, T* a, long lda, const T* x, const T* y) {
  Mat<T> A{a, lda};
  for (long j = 1; j <= m; ++j) {
    if (x[j - 1] == T() && y[j - 1] == T()) {
      A(j, j) = T(real(A(j, j)));
      continue;
    }
    const T t1 = -conj(y[j - 1]);
    const T t2 = conj(-x[j - 1]);
    for (long i = 1; i <= j - 1; ++i) A(i, j) = A(i, j) + (x[i - 1] * t1 + y[i - 1] * t2);
    A(j, j) = T(real(A(j, j)) + real(x[j - 1] * t1 + y[j - 1] * t2));
  }
}

// w = tau * A x for the Hermitian m x m block at a given by its lower
// (upper = false) or upper triangle, in the order of the reference BLAS
// xHEMV / xSYMV (the reference calls mul!(w, Hermitian(A, uplo), x, tau,
// false), which BLAS evaluates in its own order).
template <class T> void hemv(long m, const T* a, long lda, bool upper, T tau, const T* x, T* w) {
  Mat<const T> A{a, lda};
  for (long i = 0; i < m; ++i) w[i] = T();
  for (long j = 1; j <= m; ++j) {
    const T t1 = tau * x[j - 1];
    T t2 = T();
    if (upper) {
      for (long i = 1; i <= j - 1; ++i) {
        w[i - 1] = w[i - 1] + t1 * A(i, j);
        t2 = t2 + conj(A(i, j)) * x[i - 1];
      }
      w[j - 1] = w[j - 1] + t1 * real(A(j, j)) + tau * t2;
    } else {
      w[j - 1] = w[j - 1] + t1 * real(A(j, j));
      for (long i = j + 1; i <= m; ++i) {
        w[i - 1] = w[i - 1] + t1 * A(i, j);
        t2 = t2 + conj(A(i, j)) * x[i - 1];
      }
      w[j - 1] = w[j - 1] + tau * t2;
    }
  }
}

template <class T> T dotc(long m, const T* x, const T* y) {
  T s = T();
  for (long i = 0; i < m; ++i) s = s + conj(x[i]) * y[i];
  return s;
}

// _hehessenberg!(Hermitian(A, uplo)) (hessenberg.jl:27-95; LAPACK xHETD2):
// A (n x n, lda) holds the Hermitian matrix in its lower (upper = false) or
// upper triangle.  On exit d (n) and e (n-1) hold the real symmetric
// tridiagonal matrix, tau (n-1) the reflector scalars, and A the reflectors
// (L: reflector i in A[i+2:n, i]; U: in A[1:i-1, i+1]).
template <class T> void hetrd(long n, T* a, long lda, bool upper, T* tau, typename real_of<T>::type* d,
                              typename real_of<T>::type* e) {
  typedef typename real_of<T>::type R;
  Mat<T> A{a, lda};
  if (n == 0) return;
  std::vector<T> xi((size_t)n), w((size_t)n);
  if (!upper) {
    A(1, 1) = T(real(A(1, 1)));
    for (long i = 1; i <= n - 1; ++i) {
      const long m = n - i;
      T* x = &A(i + 1, i);
      const T ti = reflector(m, x, 1);
      e[i - 1] = real(x[0]);
      if (!(ti == T())) {
        x[0] = T(1.0);
        hemv<T>(m, &A(i + 1, i + 1), lda, false, ti, x, w.data());
        const T alpha = -(ti / R(2.0)) * dotc<T>(m, w.data(), x);
        for (long k = 0; k < m; ++k) w[k] = w[k] + alpha * x[k];
        rank2update_l<T>(m, &A(i + 1, i + 1), lda, x, w.data());
      } else {
        A(i + 1, i + 1) = T(real(A(i + 1, i + 1)));
      }
      A(i + 1, i) = T(e[i - 1]);
      d[i - 1] = real(A(i, i));
      tau[i - 1] = ti;
    }
    d[n - 1] = real(A(n, n));
  } else {
    A(n, n) = T(real(A(n, n)));
    for (long i = n - 1; i >= 1; --i) {
      xi[0] = A(i, i + 1);
      for (long k = 1; k <= i - 1; ++k) xi[k] = A(k, i + 1);
      const T ti = reflector(i, xi.data(), 1);
      e[i - 1] = real(xi[0]);
      for (long k = 1; k <= i - 1; ++k) A(k, i + 1) = xi[k];
      if (!(ti == T())) {
        A(i, i + 1) = T(1.0);
        const T* x = &A(1, i + 1);
        hemv<T>(i, &A(1, 1), lda, true, ti, x, w.data());
        const T alpha = -(ti / R(2.0)) * dotc<T>(i, w.data(), x);
        for (long k = 0; k < i; ++k) w[k] = w[k] + alpha * x[k];
        rank2update_u<T>(i, &A(1, 1), lda, x, w.data());
      } else {
        A(i, i) = T(real(A(i, i)));
      }
      A(i, i + 1) = T(e[i - 1]);
      d[i] = real(A(i + 1, i + 1));
      tau[i - 1] = ti;
    }
    d[0] = real(A(1, 1));
  }
}

// _materializeQ(H, Val(:U)) (hessenberg.jl:198-232): Q with A = Q T Q^H
// from the reflectors of the upper-triangle reduction.
template <class T> void materialize_q_u(long n, const T* a, long lda, const T* tau, T* q, long ldq) {
  Mat<const T> F{a, lda};
  Mat<T> A{q, ldq};
  for (long j = 1; j <= n; ++j)
    for (long i = 1; i <= n; ++i) A(i, j) = F(i, j);
  for (long j = 1; j <= n - 1; ++j) {
    for (long i = 1; i <= j - 1; ++i) A(i, j) = A(i, j + 1);
    for (long i = j; i <= n; ++i) A(i, j) = T();
  }
  for (long i = 1; i <= n - 1; ++i) A(i, n) = T();
  A(n, n) = T(1.0);
  std::vector<T> w((size_t)std::max<long>(n - 1, 1));
  for (long i = 1; i <= n - 1; ++i) {
    A(i, i) = T(1.0);
    const T ti = tau[i - 1];
    for (long j = 1; j <= i - 1; ++j) w[j - 1] = dotc<T>(i, &A(1, j), &A(1, i));
    for (long j = 1; j <= i - 1; ++j) {
      const T t = ti * conj(w[j - 1]);
      for (long k = 1; k <= i; ++k) A(k, j) = A(k, j) - A(k, i) * t;
    }
    for (long k = 1; k <= i - 1; ++k) A(k, i) = A(k, i) * (-ti);
    A(i, i) = T(1.0) - ti;
    for (long k = i + 1; k <= n; ++k) A(k, i) = T();
  }
}

// _dgz_sym2x2(a, b, c, wantvec) (symtridiag.jl:375-427; LAPACK xLAEV2):
// eigenvalues rt1 (larger modulus), rt2 and the rotation (cs, sn) of
// [a b; b c].
template <class R> void sym2x2(R a, R b, R c, bool wantvec, R& rt1, R& rt2, R& cs1, R& sn1) {
  const R sm = a + c, df = a - c, adf = rt<R>::abs(df), tb = R(2.0) * b, ab = rt<R>::abs(tb);
  R acmx, acmn;
  if (rt<R>::abs(a) > rt<R>::abs(c)) {
    acmx = a;
    acmn = c;
  } else {
    acmx = c;
    acmn = a;
  }
  const R rt_ = jhypot<R>(ab, adf);
  int sgn1;
  if (sm < R(0.0)) {
    rt1 = (sm - rt_) / R(2.0);
    sgn1 = -1;
    rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
  } else if (sm > R(0.0)) {
    rt1 = (sm + rt_) / R(2.0);
    sgn1 = 1;
    rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
  } else {
    rt1 = rt_ / R(2.0);
    rt2 = -rt_ / R(2.0);
    sgn1 = 1;
  }
  cs1 = R(0.0);
  sn1 = R(0.0);
  if (!wantvec) return;
  R cs;
  int sgn2;
  if (df > R(0.0)) {
    cs = df + rt_;
    sgn2 = 1;
  } else {
    cs = df - rt_;
    sgn2 = -1;
  }
  const R acs = rt<R>::abs(cs);
  if (acs > ab) {
    const R ct = -tb / cs;
    sn1 = R(1.0) / rt<R>::sqrt(R(1.0) + ct * ct);
    cs1 = ct * sn1;
  } else {
    if (ab == R(0.0)) {
      cs1 = R(1.0);
      sn1 = R(0.0);
    } else {
      const R tn = -cs / tb;
      cs1 = R(1.0) / rt<R>::sqrt(R(1.0) + tn * tn);
      sn1 = tn * cs1;
    }
  }
  if (sgn1 == sgn2) {
    const R tn = cs1;
    cs1 = -sn1;
    sn1 = tn;
  }
}

// rmul!(Z, Givens(i, i+1, c, s)') on columns i, i+1 (1-based) of the
// zrows x n matrix Z: Z[:, i] = c a1 + s a2, Z[:, i+1] = -s a1 + c a2.
template <class R, class TZ> void rot_cols(TZ* z, long ldz, long zrows, long i, R c, R s) {
  Mat<TZ> Z{z, ldz};
  for (long r = 1; r <= zrows; ++r) {
    const TZ a1 = Z(r, i), a2 = Z(r, i + 1);
    Z(r, i) = a1 * c + a2 * s;
    Z(r, i + 1) = (-a1) * s + a2 * c;
  }
}

// _st_findsplit_fwd! (symtridiag.jl:354-372)
template <class R> long st_findsplit(R* d, R* e, long l1, long n) {
  long msplit = n;
  for (long m = l1; m <= n - 1; ++m) {
    const R tst = rt<R>::abs(e[m - 1]);
    if (tst == R(0.0)) {
      msplit = m;
      break;
    }
    if (tst <= rt<R>::sqrt(rt<R>::abs(d[m - 1])) * rt<R>::sqrt(rt<R>::abs(d[m])) * rt<R>::eps()) {
      e[m - 1] = R(0.0);
      msplit = m;
      break;
    }
  }
  return msplit;
}

// _gschur!(A::SymTridiagonal, QRIteration(), Z) (symtridiag.jl:77-172;
// LAPACK xSTEQR) with _st_QLIter! (:175-262) and _st_QRIter! (:265-352):
// d (n) is overwritten with the eigenvalues (unsorted), e (n-1) destroyed,
// and the zrows x n matrix Z (real or complex; null: none) multiplied by
// the rotations.  The reference's iteration counter never advances
// (symtridiag.jl:141), so it has no limit; this restatement stops after
// maxsweeps QL/QR sweeps in all (30 n, LAPACK's limit, by default) and
// returns the number of nonzero off-diagonal entries then left (LAPACK's
// info), 0 on convergence.  *sweeps receives the sweep count.
template <class R, class TZ>
long steqr(long n, R* d, R* e, TZ* z, long ldz, long zrows, long maxsweeps, long* sweeps) {
  const bool wantz = z != nullptr && zrows > 0;
  const R epsT = rt<R>::eps(), eps2 = epsT * epsT;
  const R safmin = rt<R>::fmin(), safmax = R(1.0) / safmin;
  const R ssfmax = rt<R>::sqrt(safmax) / R(3.0);
  const R ssfmin = rt<R>::sqrt(safmin) / eps2;
  long niter = 0;
  if (maxsweeps < 0) maxsweeps = 30 * n;
  if (sweeps) *sweeps = 0;
  std::vector<R> gc((size_t)std::max<long>(n, 1)), gs((size_t)std::max<long>(n, 1));
  // 1-based views
  auto D = [&](long i) -> R& { return d[i - 1]; };
  auto E = [&](long i) -> R& { return e[i - 1]; };
  long l1 = 1;
  while (l1 <= n) {
    if (l1 > 1) E(l1 - 1) = R(0.0);
    const long msplit0 = st_findsplit<R>(d, e, l1, n);
    long l = l1, lsave = l, lend = msplit0, lendsave = lend;
    l1 = msplit0 + 1;
    if (lend == l) continue;
    // opnorm(SymTridiagonal(d[l:lend], e[l:lend-1]), Inf): largest row sum
    R anorm(0.0);
    for (long i = l; i <= lend; ++i) {
      R s(0.0);
      if (i > l) s = s + rt<R>::abs(E(i - 1));
      s = s + rt<R>::abs(D(i));
      if (i < lend) s = s + rt<R>::abs(E(i));
      anorm = rmax(anorm, s);
    }
    int iscale = 0;
    if (anorm == R(0.0)) continue;
    if (anorm > ssfmax) {
      iscale = 1;
      safescale<R>(lend - l + 1, &D(l), anorm, ssfmax);
      safescale<R>(lend - l, &E(l), anorm, ssfmax);
    } else if (anorm < ssfmin) {
      iscale = 2;
      safescale<R>(lend - l + 1, &D(l), anorm, ssfmin);
      safescale<R>(lend - l, &E(l), anorm, ssfmin);
    }
    if (rt<R>::abs(D(lend)) < rt<R>::abs(D(l))) {
      lend = lsave;
      l = lendsave;
    }
    bool stop = false;
    if (lend > l) {
      // QL iteration
      while (l <= lend) {
        long msplit = lend;
        for (long m = l; m <= lend - 1; ++m) {
          if (E(m) * E(m) <= (eps2 * rt<R>::abs(D(m))) * rt<R>::abs(D(m + 1)) + safmin) {
            msplit = m;
            break;
          }
        }
        if (msplit < lend) E(msplit) = R(0.0);
        R p = D(l);
        if (msplit != l) {
          if (msplit == l + 1) {
            R rt1, rt2, cs, sn;
            sym2x2<R>(D(l), E(l), D(l + 1), wantz, rt1, rt2, cs, sn);
            if (wantz) rot_cols<R, TZ>(z, ldz, zrows, l, cs, sn);
            D(l) = rt1;
            D(l + 1) = rt2;
            E(l) = R(0.0);
            l += 2;
            continue;
          }
          if (niter >= maxsweeps) {
            stop = true;
            break;
          }
          niter += 1;
          R g = (D(l + 1) - p) / (R(2.0) * E(l));
          R r = jhypot<R>(g, R(1.0));
          g = D(msplit) - p + (E(l) / (g + rt<R>::copysign(r, g)));
          R s(1.0), c(1.0);
          p = R(0.0);
          for (long i = msplit - 1; i >= l; --i) {
            const R f = s * E(i);
            const R b = c * E(i);
            givens_re<R>(g, f, c, s, r);
            if (i < msplit - 1) E(i + 1) = r;
            g = D(i + 1) - p;
            r = (D(i) - g) * s + R(2.0) * c * b;
            p = s * r;
            D(i + 1) = g + p;
            g = c * r - b;
            gc[i - 1] = c;
            gs[i - 1] = -s;
          }
          if (wantz)
            for (long i = msplit - 1; i >= l; --i) rot_cols<R, TZ>(z, ldz, zrows, i, gc[i - 1], gs[i - 1]);
          D(l) = D(l) - p;
          E(l) = g;
        } else {
          D(l) = p;
          l += 1;
        }
      }
    } else {
      // QR iteration
      while (l >= lend) {
        long msplit = lend;
        for (long m = l; m >= lend + 1; --m) {
          if (E(m - 1) * E(m - 1) <= (eps2 * rt<R>::abs(D(m))) * rt<R>::abs(D(m - 1)) + safmin) {
            msplit = m;
            break;
          }
        }
        if (msplit > lend) E(msplit - 1) = R(0.0);
        R p = D(l);
        if (msplit != l) {
          if (msplit == l - 1) {
            R rt1, rt2, cs, sn;
            sym2x2<R>(D(l - 1), E(l - 1), D(l), wantz, rt1, rt2, cs, sn);
            if (wantz) rot_cols<R, TZ>(z, ldz, zrows, l - 1, cs, sn);
            D(l - 1) = rt1;
            D(l) = rt2;
            E(l - 1) = R(0.0);
            l -= 2;
            continue;
          }
          if (niter >= maxsweeps) {
            stop = true;
            break;
          }
          niter += 1;
          R g = (D(l - 1) - p) / (R(2.0) * E(l - 1));
          R r = jhypot<R>(g, R(1.0));
          g = D(msplit) - p + (E(l - 1) / (g + rt<R>::copysign(r, g)));
          R s(1.0), c(1.0);
          p = R(0.0);
          for (long i = msplit; i <= l - 1; ++i) {
            const R f = s * E(i);
            const R b = c * E(i);
            givens_re<R>(g, f, c, s, r);
            if (i != msplit) E(i - 1) = r;
            g = D(i) - p;
            r = (D(i + 1) - g) * s + R(2.0) * c * b;
            p = s * r;
            D(i) = g + p;
            g = c * r - b;
            gc[i - 1] = c;
            gs[i - 1] = s;
          }
          if (wantz)
            for (long i = msplit; i <= l - 1; ++i) rot_cols<R, TZ>(z, ldz, zrows, i, gc[i - 1], gs[i - 1]);
          D(l) = D(l) - p;
          E(l - 1) = g;
        } else {
          D(l) = p;
          l -= 1;
        }
      }
    }
    if (iscale == 1) {
      safescale<R>(lendsave - lsave + 1, &D(lsave), ssfmax, anorm);
      safescale<R>(lendsave - lsave, &E(lsave), ssfmax, anorm);
    } else if (iscale == 2) {
      safescale<R>(lendsave - lsave + 1, &D(lsave), ssfmin, anorm);
      safescale<R>(lendsave - lsave, &E(lsave), ssfmin, anorm);
    }
    if (stop) {
      if (sweeps) *sweeps = niter;
      long nbad = 0;
      for (long i = 1; i <= n - 1; ++i) nbad += (E(i) != R(0.0)) ? 1 : 0;
      return nbad > 0 ? nbad : 1;
    }
  }
  if (sweeps) *sweeps = niter;
  return 0;
}

// Julia's isless on floats: a total order with -0.0 < 0.0 and NaN last.
template <class R> bool jl_isless(R a, R b) {
  const bool an = rt<R>::isnan(a), bn = rt<R>::isnan(b);
  if (an || bn) return !an && bn;
  if (a == b) return std::signbit(rt<R>::to_double(a)) && !std::signbit(rt<R>::to_double(b));
  return a < b;
}

// sorteig!(λ, V, eigsortby) for real eigenvalues (LinearAlgebra; called by
// geigen!, GenericSchur.jl:133): a stable ascending sort under isless,
// the columns of V (vrows x n) permuted alike; no change when sorted.
template <class R, class TV> void sorteig(long n, R* w, TV* v, long ldv, long vrows) {
  std::vector<long> p((size_t)n);
  for (long i = 0; i < n; ++i) p[i] = i;
  std::stable_sort(p.begin(), p.end(), [&](long a, long b) { return jl_isless<R>(w[a], w[b]); });
  std::vector<R> w2((size_t)n);
  for (long i = 0; i < n; ++i) w2[i] = w[p[i]];
  for (long i = 0; i < n; ++i) w[i] = w2[i];
  if (!v) return;
  std::vector<TV> tmp((size_t)(vrows * n));
  for (long j = 0; j < n; ++j)
    for (long r = 0; r < vrows; ++r) tmp[r + j * vrows] = v[r + p[j] * ldv];
  for (long j = 0; j < n; ++j)
    for (long r = 0; r < vrows; ++r) v[r + j * ldv] = tmp[r + j * vrows];
}

// geigen! / geigvals! for a Hermitian (or real symmetric) matrix with
// QRIteration (GenericSchur.jl:124-146): _hehessenberg!, _materializeQ,
// gschur!(H, V) (symtridiag.jl:14-41) and sorteig! (when sort).  A (n x n)
// is overwritten by the reduction; w (n) receives the eigenvalues, v (n x n,
// null: eigenvalues only) the eigenvectors.  Returns steqr's info.
template <class T> long heev(long n, T* a, long lda, bool upper, typename real_of<T>::type* w, T* v, long ldv,
                             bool sort, long* sweeps) {
  typedef typename real_of<T>::type R;
  if (sweeps) *sweeps = 0;
  if (n == 0) return 0;
  std::vector<T> tau((size_t)std::max<long>(n - 1, 1));
  std::vector<R> e((size_t)std::max<long>(n - 1, 1));
  hetrd<T>(n, a, lda, upper, tau.data(), w, e.data());
  if (v) {
    if (upper) materialize_q_u<T>(n, a, lda, tau.data(), v, ldv);
    else materialize_q<T>(n, a, lda, tau.data(), v, ldv);
  }
  const long info = steqr<R, T>(n, w, e.data(), v, ldv, v ? n : 0, -1, sweeps);
  if (sort) {
    bool sorted = true;
    for (long i = 1; i < n; ++i) sorted = sorted && !jl_isless<R>(w[i], w[i - 1]);
    if (!sorted) sorteig<R, T>(n, w, v, ldv, v ? n : 0);
  }
  return info;
}


// ===================================================================
// Reordering of a real Schur form: gordschur! (rordschur.jl:3-45),
// _trexchange! (:54-165, LAPACK xTREXC), _swap1or2 (:168-297, xLAEXC) and
// _syl1or2 (:300-466, xLASY2).  BLAS calls of the reference on length-2
// and length-3 vectors are evaluated in the order of the reference BLAS
// (xDOT, xAXPY, xGEMV, xGER); LinearAlgebra's reflector! (stdlib) and its
// short-vector norm (generic_norm2) are restated.
// ===================================================================

// jmax (defined with the balancing code) is Julia's NaN-propagating max.

// LinearAlgebra.generic_norm2 (the norm of vectors shorter than 32):
// unscaled sum of squares unless n maxabs^2 overflows or maxabs^2 underflows
template <class R> R jl_norm2_short(long n, const R* x) {
  if (n == 0) return R(0.0);
  R maxabs = rt<R>::abs(x[0]);
  for (long i = 1; i < n; ++i) maxabs = jmax(maxabs, rt<R>::abs(x[i]));
  const bool inf = !rt<R>::isnan(maxabs) && !(maxabs <= rt<R>::fmax());
  if (maxabs == R(0.0) || inf) return maxabs;
  const R t = (R(double(n)) * maxabs) * maxabs;
  const bool finite = !rt<R>::isnan(t) && t <= rt<R>::fmax();
  if (finite && !(maxabs * maxabs == R(0.0))) {
    R sum = x[0] * x[0];
    for (long i = 1; i < n; ++i) sum = sum + x[i] * x[i];
    return rt<R>::sqrt(sum);
  }
  R y = rt<R>::abs(x[0]) / maxabs;
  R sum = y * y;
  for (long i = 1; i < n; ++i) {
    y = rt<R>::abs(x[i]) / maxabs;
    sum = sum + y * y;
  }
  return maxabs * rt<R>::sqrt(sum);
}

// LinearAlgebra.reflector!(x) (stdlib): x[0] <- -nu, x[1:] <- x[1:] / (x0 + nu),
// returns tau = (x0 + nu) / nu with nu = copysign(norm(x), x0).
template <class R> R jl_reflector(long n, R* x) {
  if (n == 0) return R(0.0);
  R xi1 = x[0];
  const R normu = jl_norm2_short<R>(n, x);
  if (normu == R(0.0)) return R(0.0);
  const R nu = rt<R>::copysign(normu, xi1);
  xi1 = xi1 + nu;
  x[0] = -nu;
  for (long i = 1; i < n; ++i) x[i] = x[i] / xi1;
  return xi1 / nu;
}

// _syl1or2(Tl, Tr, B) (rordschur.jl:300-466; xLASY2 with isgn = -1): X with
// Tl X - X Tr = scale B, orders n1, n2 in {1, 2}; matrices 2 x 2 column-major
// (element (i, j) at [i + 2 j]).  Returns scale.
template <class R> R syl1or2(int n1, int n2, const R* tl, const R* tr, const R* b, R* x) {
  const R epst = rt<R>::eps();
  const R smallnum = rt<R>::fmin() / epst;
  auto TL = [&](int i, int j) { return tl[(i - 1) + 2 * (j - 1)]; };
  auto TR = [&](int i, int j) { return tr[(i - 1) + 2 * (j - 1)]; };
  auto B = [&](int i, int j) { return b[(i - 1) + 2 * (j - 1)]; };
  R scale(1.0);
  if (n1 == 1 && n2 == 1) {
    R tau1 = TL(1, 1) - TR(1, 1);
    R beta = rt<R>::abs(tau1);
    if (beta <= smallnum) {
      tau1 = smallnum;
      beta = smallnum;
    }
    const R gamma = rt<R>::abs(B(1, 1));
    if (smallnum * gamma > beta) scale = R(1.0) / gamma;
    x[0] = (B(1, 1) * scale) / tau1;
    return scale;
  }
  if (n1 == 2 && n2 == 2) {
    R smin = jmax(jmax(jmax(rt<R>::abs(TR(1, 1)), rt<R>::abs(TR(2, 1))), jmax(rt<R>::abs(TR(1, 2)), rt<R>::abs(TR(2, 2)))),
                  jmax(jmax(rt<R>::abs(TL(1, 1)), rt<R>::abs(TL(2, 1))), jmax(rt<R>::abs(TL(1, 2)), rt<R>::abs(TL(2, 2)))));
    smin = jmax(epst * smin, smallnum);
    R TT[4][4] = {};
    TT[0][0] = TL(1, 1) - TR(1, 1);
    TT[1][1] = TL(2, 2) - TR(1, 1);
    TT[2][2] = TL(1, 1) - TR(2, 2);
    TT[3][3] = TL(2, 2) - TR(2, 2);
    TT[0][1] = TL(1, 2);
    TT[1][0] = TL(2, 1);
    TT[2][3] = TL(1, 2);
    TT[3][2] = TL(2, 1);
    TT[0][2] = -TR(2, 1);
    TT[1][3] = -TR(2, 1);
    TT[2][0] = -TR(1, 2);
    TT[3][1] = -TR(1, 2);
    R bv[4] = {B(1, 1), B(2, 1), B(1, 2), B(2, 2)};
    int ipsv = 0, jpsv = 0;
    int jpiv[4] = {0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
      R xmax(0.0);
      for (int ip = i; ip < 4; ++ip)
        for (int jp = i; jp < 4; ++jp)
          if (rt<R>::abs(TT[ip][jp]) > xmax) {
            xmax = rt<R>::abs(TT[ip][jp]);
            ipsv = ip;
            jpsv = jp;
          }
      if (ipsv != i) {
        for (int k = 0; k < 4; ++k) std::swap(TT[ipsv][k], TT[i][k]);
        std::swap(bv[i], bv[ipsv]);
      }
      if (jpsv != i)
        for (int k = 0; k < 4; ++k) std::swap(TT[k][jpsv], TT[k][i]);
      jpiv[i] = jpsv;
      if (rt<R>::abs(TT[i][i]) < smin) TT[i][i] = smin;
      for (int j = i + 1; j < 4; ++j) {
        TT[j][i] = TT[j][i] / TT[i][i];
        bv[j] = bv[j] - TT[j][i] * bv[i];
        for (int k = i + 1; k < 4; ++k) TT[j][k] = TT[j][k] - TT[j][i] * TT[i][k];
      }
    }
    if (rt<R>::abs(TT[3][3]) < smin) TT[3][3] = smin;
    const R a = R(8.0) * smallnum;
    if ((a * rt<R>::abs(bv[0]) > TT[0][0]) || (a * rt<R>::abs(bv[1]) > TT[1][1]) || (a * rt<R>::abs(bv[2]) > TT[2][2]) ||
        (a * rt<R>::abs(bv[3]) > TT[3][3])) {
      const R bmax = jmax(jmax(rt<R>::abs(bv[0]), rt<R>::abs(bv[1])), jmax(rt<R>::abs(bv[2]), rt<R>::abs(bv[3])));
      scale = (R(1.0) / R(8.0)) / bmax;
      for (int k = 0; k < 4; ++k) bv[k] = bv[k] * scale;
    }
    R xt[4] = {};
    for (int i = 1; i <= 4; ++i) {
      const int k = 4 - i;
      const R t = R(1.0) / TT[k][k];
      xt[k] = bv[k] * t;
      for (int j = k + 1; j < 4; ++j) xt[k] = xt[k] - (t * TT[k][j]) * xt[j];
    }
    for (int i = 1; i <= 3; ++i) {
      const int jj = jpiv[3 - i];
      if (jj != 3 - i) std::swap(xt[3 - i], xt[jj]);
    }
    x[0] = xt[0];
    x[1] = xt[1];
    x[2] = xt[2];
    x[3] = xt[3];
    return scale;
  }
  R smin, tmp[4], btmp[2];
  if (n1 == 1) {
    smin = jmax(epst * jmax(rt<R>::abs(TL(1, 1)),
                            jmax(jmax(rt<R>::abs(TR(1, 1)), rt<R>::abs(TR(2, 1))), jmax(rt<R>::abs(TR(1, 2)), rt<R>::abs(TR(2, 2))))),
                smallnum);
    tmp[0] = TL(1, 1) - TR(1, 1);
    tmp[1] = -TR(1, 2);
    tmp[2] = -TR(2, 1);
    tmp[3] = TL(1, 1) - TR(2, 2);
    btmp[0] = B(1, 1);
    btmp[1] = B(1, 2);
  } else {
    smin = jmax(epst * jmax(rt<R>::abs(TR(1, 1)),
                            jmax(jmax(rt<R>::abs(TL(1, 1)), rt<R>::abs(TL(2, 1))), jmax(rt<R>::abs(TL(1, 2)), rt<R>::abs(TL(2, 2))))),
                smallnum);
    tmp[0] = TL(1, 1) - TR(1, 1);
    tmp[1] = TL(2, 1);
    tmp[2] = TL(1, 2);
    tmp[3] = TL(2, 2) - TR(1, 1);
    btmp[0] = B(1, 1);
    btmp[1] = B(2, 1);
  }
  // findmax(abs.(tmp)): the first index of the largest value
  int ipiv = 0;
  R amax = rt<R>::abs(tmp[0]);
  for (int k = 1; k < 4; ++k)
    if (rt<R>::abs(tmp[k]) > amax) {
      amax = rt<R>::abs(tmp[k]);
      ipiv = k;
    }
  R u11 = tmp[ipiv];
  if (amax <= smin) u11 = smin;
  static const int inds[4][3] = {{2, 1, 3}, {3, 0, 2}, {0, 3, 1}, {1, 2, 0}};
  static const bool xsw[4] = {false, false, true, true}, bsw[4] = {false, true, false, true};
  const R u12 = tmp[inds[ipiv][0]];
  const R l21 = tmp[inds[ipiv][1]] / u11;
  R u22 = tmp[inds[ipiv][2]] - u12 * l21;
  if (rt<R>::abs(u22) <= smin) u22 = smin;
  R b1, b2;
  if (bsw[ipiv]) {
    b1 = btmp[1];
    b2 = btmp[0] - l21 * btmp[1];
  } else {
    b1 = btmp[0];
    b2 = btmp[1] - l21 * btmp[0];
  }
  if (((R(2.0) * smallnum) * rt<R>::abs(b2) > rt<R>::abs(u22)) || ((R(2.0) * smallnum) * rt<R>::abs(b1) > rt<R>::abs(u11))) {
    scale = (R(1.0) / R(2.0)) / jmax(rt<R>::abs(b1), rt<R>::abs(b2));
    b1 = scale * b1;
    b2 = scale * b2;
  }
  const R x22 = b2 / u22;
  R x1 = b1 / u11 - (u12 / u11) * x22, x2 = x22;
  if (xsw[ipiv]) std::swap(x1, x2);
  if (n1 == 1) {  // X is 1 x 2
    x[0] = x1;
    x[2] = x2;
  } else {  // X is 2 x 1
    x[0] = x1;
    x[1] = x2;
  }
  return scale;
}

// Householder H = I - tau [1; v1; v2][1; v1; v2]^T (householder.jl:105-172):
// lmul!(H', A[r0:r0+2, c0:c1]) and rmul!(A[r0:r1, c0:c0+2], H); 1-based.
template <class R> void hh3_lmul(Mat<R> A, long r0, long c0, long c1, R v1, R v2, R tau) {
  for (long j = c0; j <= c1; ++j) {
    R va = A(r0, j);
    va = va + (v1 * A(r0 + 1, j) + v2 * A(r0 + 2, j));
    va = tau * va;
    A(r0, j) = A(r0, j) - va;
    A(r0 + 1, j) = A(r0 + 1, j) + (-va) * v1;
    A(r0 + 2, j) = A(r0 + 2, j) + (-va) * v2;
  }
}
template <class R> void hh3_rmul(Mat<R> A, long r0, long r1, long c0, R v1, R v2, R tau) {
  for (long i = r0; i <= r1; ++i) {
    R x = A(i, c0 + 1) * v1 + A(i, c0 + 2) * v2;
    x = x + A(i, c0);
    A(i, c0) = A(i, c0) + (-tau) * x;
    A(i, c0 + 1) = A(i, c0 + 1) + x * ((-tau) * v1);
    A(i, c0 + 2) = A(i, c0 + 2) + x * ((-tau) * v2);
  }
}
// the shifted reflector [v1; v2; 1] of the n1 = 1, n2 = 2 swap
// (_reflectorYappl!, rordschur.jl:511-540: A <- H' A on rows r0..r0+2;
// _reflectorYlppa!, :485-509: A <- A H on columns c0..c0+2)
template <class R> void hy3_lmul(Mat<R> A, long r0, long c0, long c1, R v1, R v2, R tau) {
  for (long j = c0; j <= c1; ++j) {
    R va = A(r0 + 2, j);
    va = va + v1 * A(r0, j);
    va = va + v2 * A(r0 + 1, j);
    va = tau * va;
    A(r0 + 2, j) = A(r0 + 2, j) - va;
    A(r0, j) = A(r0, j) - v1 * va;
    A(r0 + 1, j) = A(r0 + 1, j) - v2 * va;
  }
}
template <class R> void hy3_rmul(Mat<R> A, long r0, long r1, long c0, R v1, R v2, R tau) {
  for (long i = r0; i <= r1; ++i) {
    R x = A(i, c0 + 2);
    x = x + A(i, c0) * v1;
    x = x + A(i, c0 + 1) * v2;
    x = tau * x;
    A(i, c0 + 2) = A(i, c0 + 2) - x;
    A(i, c0) = A(i, c0) - x * v1;
    A(i, c0 + 1) = A(i, c0 + 1) - x * v2;
  }
}
// rows (i, i+1) <- G (i, i+1) for columns c0..c1 and columns (i, i+1) <- (i, i+1) G'
// for rows r0..r1, G = Givens(i, i+1, c, s)
template <class R> void g2_lmul(Mat<R> A, long i, long c0, long c1, R c, R s) {
  for (long j = c0; j <= c1; ++j) {
    const R a1 = A(i, j), a2 = A(i + 1, j);
    A(i, j) = c * a1 + s * a2;
    A(i + 1, j) = -s * a1 + c * a2;
  }
}
template <class R> void g2_rmul(Mat<R> A, long r0, long r1, long i, R c, R s) {
  for (long r = r0; r <= r1; ++r) {
    const R a1 = A(r, i), a2 = A(r, i + 1);
    A(r, i) = a1 * c + a2 * s;
    A(r, i + 1) = -a1 * s + a2 * c;
  }
}

// _swap1or2(T, Z, j1, n1, n2) (rordschur.jl:168-297; LAPACK xLAEXC)
template <class R> bool swap1or2(long n, R* t, long ldt, R* z, long ldz, long j1, int n1, int n2) {
  Mat<R> T{t, ldt};
  Mat<R> Z{z, ldz};
  if (j1 + n1 > n) return true;
  const long j2 = j1 + 1, j3 = j1 + 2, j4 = j1 + 3;
  if (n1 == 1 && n2 == 1) {
    const R T11 = T(j1, j1), T22 = T(j2, j2);
    R c, s, r;
    givens_re<R>(T(j1, j2), T22 - T11, c, s, r);
    if (j3 <= n) g2_lmul<R>(T, j1, j3, n, c, s);
    g2_rmul<R>(T, 1, j1 - 1, j1, c, s);
    T(j1, j1) = T22;
    T(j2, j2) = T11;
    if (z) g2_rmul<R>(Z, 1, n, j1, c, s);
    return true;
  }
  const int nd = n1 + n2;
  R Dm[16];
  Mat<R> D{Dm, 4};
  R dnorm(0.0);
  for (int j = 1; j <= nd; ++j)
    for (int i = 1; i <= nd; ++i) {
      D(i, j) = T(j1 + i - 1, j1 + j - 1);
      dnorm = (i == 1 && j == 1) ? rt<R>::abs(D(i, j)) : jmax(dnorm, rt<R>::abs(D(i, j)));
    }
  const R smallnum = rt<R>::fmin() / rt<R>::eps();
  const R thresh = jmax(R(10.0) * rt<R>::eps() * dnorm, smallnum);
  R tl[4] = {}, tr[4] = {}, bb[4] = {}, X[4] = {};
  for (int j = 0; j < n1; ++j)
    for (int i = 0; i < n1; ++i) tl[i + 2 * j] = D(i + 1, j + 1);
  for (int j = 0; j < n2; ++j)
    for (int i = 0; i < n2; ++i) tr[i + 2 * j] = D(n1 + i + 1, n1 + j + 1);
  for (int j = 0; j < n2; ++j)
    for (int i = 0; i < n1; ++i) bb[i + 2 * j] = D(i + 1, n1 + j + 1);
  const R scale = syl1or2<R>(n1, n2, tl, tr, bb, X);
  auto x = [&](int i, int j) { return X[(i - 1) + 2 * (j - 1)]; };
  if (n1 == 1 && n2 == 2) {
    R u[3] = {x(1, 2), scale, x(1, 1)};
    const R tau = jl_reflector<R>(3, u);
    const R T11 = T(j1, j1);
    hy3_lmul<R>(D, 1, 1, 3, u[1], u[2], tau);
    hy3_rmul<R>(D, 1, 3, 1, u[1], u[2], tau);
    if (jmax(jmax(rt<R>::abs(D(3, 1)), rt<R>::abs(D(3, 2))), rt<R>::abs(D(3, 3) - T11)) > thresh) return false;
    hy3_lmul<R>(T, j1, j1, n, u[1], u[2], tau);
    hy3_rmul<R>(T, 1, j2, j1, u[1], u[2], tau);
    T(j3, j1) = R(0.0);
    T(j3, j2) = R(0.0);
    T(j3, j3) = T11;
    if (z) hy3_rmul<R>(Z, 1, n, j1, u[1], u[2], tau);
  } else if (n1 == 2 && n2 == 1) {
    R u[3] = {-x(1, 1), -x(2, 1), scale};
    const R tau = jl_reflector<R>(3, u);
    u[0] = R(1.0);
    const R T33 = T(j3, j3);
    hh3_lmul<R>(D, 1, 1, 3, u[1], u[2], tau);
    hh3_rmul<R>(D, 1, 3, 1, u[1], u[2], tau);
    if (jmax(jmax(rt<R>::abs(D(1, 1) - T33), rt<R>::abs(D(2, 1))), rt<R>::abs(D(3, 1))) > thresh) return false;
    hh3_rmul<R>(T, 1, j3, j1, u[1], u[2], tau);
    hh3_lmul<R>(T, j1, j2, n, u[1], u[2], tau);
    T(j1, j1) = T33;
    T(j2, j1) = R(0.0);
    T(j3, j1) = R(0.0);
    if (z) hh3_rmul<R>(Z, 1, n, j1, u[1], u[2], tau);
  } else {
    R u1[3] = {-x(1, 1), -x(2, 1), scale};
    const R tau1 = jl_reflector<R>(3, u1);
    const R tt = -tau1 * (x(1, 2) + u1[1] * x(2, 2));
    R u2[3] = {-tt * u1[1] - x(2, 2), -tt * u1[2], scale};
    const R tau2 = jl_reflector<R>(3, u2);
    hh3_lmul<R>(D, 1, 1, 4, u1[1], u1[2], tau1);
    hh3_rmul<R>(D, 1, 4, 1, u1[1], u1[2], tau1);
    hh3_lmul<R>(D, 2, 1, 4, u2[1], u2[2], tau2);
    hh3_rmul<R>(D, 1, 4, 2, u2[1], u2[2], tau2);
    if (jmax(jmax(rt<R>::abs(D(3, 1)), rt<R>::abs(D(3, 2))), jmax(rt<R>::abs(D(4, 1)), rt<R>::abs(D(4, 2)))) > thresh)
      return false;
    hh3_lmul<R>(T, j1, j1, n, u1[1], u1[2], tau1);
    hh3_rmul<R>(T, 1, j4, j1, u1[1], u1[2], tau1);
    hh3_lmul<R>(T, j2, j1, n, u2[1], u2[2], tau2);
    hh3_rmul<R>(T, 1, j4, j2, u2[1], u2[2], tau2);
    T(j3, j1) = R(0.0);
    T(j3, j2) = R(0.0);
    T(j4, j1) = R(0.0);
    T(j4, j2) = R(0.0);
    if (z) {
      hh3_rmul<R>(Z, 1, n, j1, u1[1], u1[2], tau1);
      hh3_rmul<R>(Z, 1, n, j2, u2[1], u2[2], tau2);
    }
  }
  auto standardize = [&](long k) {
    R a = T(k, k), b = T(k, k + 1), c = T(k + 1, k), d = T(k + 1, k + 1), cs, sn;
    cx<R> w1, w2;
    gs2x2<R>(a, b, c, d, cs, sn, w1, w2);
    g2_lmul<R>(T, k, k + 2, n, cs, sn);
    g2_rmul<R>(T, 1, k - 1, k, cs, sn);
    T(k, k) = a;
    T(k, k + 1) = b;
    T(k + 1, k) = c;
    T(k + 1, k + 1) = d;
    if (z) g2_rmul<R>(Z, 1, n, k, cs, sn);
  };
  if (n2 == 2) standardize(j1);
  if (n1 == 2) standardize(j1 + n2);
  return true;
}

// _trexchange!(T, Z, iold, inew) (rordschur.jl:54-165; LAPACK xTREXC):
// returns false when a swap was refused; inew receives the final position.
template <class R> bool trexchange_re(long n, R* t, long ldt, R* z, long ldz, long iold, long& inew) {
  Mat<R> T{t, ldt};
  if (n <= 1 || iold == inew) return true;
  if (iold > 1 && T(iold, iold - 1) != R(0.0)) iold -= 1;
  auto nbsize = [&](long j) { return (j < n && T(j + 1, j) != R(0.0)) ? 2 : 1; };
  int nbold = nbsize(iold);
  if (inew > 1 && T(inew, inew - 1) != R(0.0)) inew -= 1;
  const int nbnew = nbsize(inew);
  if (iold == inew) return true;
  long ihere = iold;
  bool ok = true;
  if (iold < inew) {
    if (nbold == 2 && nbnew == 1) inew -= 1;
    else if (nbold == 1 && nbnew == 2) inew += 1;
    while (ihere < inew) {
      if (nbold == 1 || nbold == 2) {
        const int nbnext = nbsize(ihere + nbold);
        ok = swap1or2<R>(n, t, ldt, z, ldz, ihere, nbold, nbnext);
        if (!ok) { inew = ihere; return false; }
        ihere += nbnext;
        if (nbold == 2 && T(ihere + 1, ihere) == R(0.0)) nbold = 3;
      } else {
        int nbnext = nbsize(ihere + 1);
        ok = swap1or2<R>(n, t, ldt, z, ldz, ihere + 1, 1, nbnext);
        if (!ok) { inew = ihere; return false; }
        if (nbnext == 1) {
          ok = swap1or2<R>(n, t, ldt, z, ldz, ihere, 1, nbnext);
          ihere += 1;
        } else {
          if (T(ihere + 2, ihere + 1) == R(0.0)) nbnext = 1;
          if (nbnext == 2) {
            ok = swap1or2<R>(n, t, ldt, z, ldz, ihere, 1, nbnext);
            if (!ok) { inew = ihere; return false; }
            ihere += 2;
          } else {
            ok = swap1or2<R>(n, t, ldt, z, ldz, ihere, 1, 1);
            ok = swap1or2<R>(n, t, ldt, z, ldz, ihere + 1, 1, 1);
            ihere += 2;
          }
        }
      }
    }
  } else {
    while (ihere > inew) {
      if (nbold == 1 || nbold == 2) {
        const int nbnext = (ihere >= 3 && T(ihere - 1, ihere - 2) != R(0.0)) ? 2 : 1;
        ok = swap1or2<R>(n, t, ldt, z, ldz, ihere - nbnext, nbnext, nbold);
        if (!ok) { inew = ihere; return false; }
        ihere -= nbnext;
        if (nbold == 2 && T(ihere + 1, ihere) == R(0.0)) nbold = 3;
      } else {
        int nbnext = (ihere >= 3 && T(ihere - 1, ihere - 2) != R(0.0)) ? 2 : 1;
        ok = swap1or2<R>(n, t, ldt, z, ldz, ihere - nbnext, nbnext, 1);
        if (!ok) { inew = ihere; return false; }
        if (nbnext == 1) {
          ok = swap1or2<R>(n, t, ldt, z, ldz, ihere, nbnext, 1);
          ihere -= 1;
        } else {
          if (T(ihere, ihere - 1) == R(0.0)) nbnext = 1;
          if (nbnext == 2) {
            ok = swap1or2<R>(n, t, ldt, z, ldz, ihere - 1, 2, 1);
            if (!ok) { inew = ihere; return false; }
            ihere -= 2;
          } else {
            ok = swap1or2<R>(n, t, ldt, z, ldz, ihere, 1, 1);
            ok = swap1or2<R>(n, t, ldt, z, ldz, ihere - 1, 1, 1);
            ihere -= 2;
          }
        }
      }
    }
  }
  inew = ihere;
  return ok;
}

// gordschur!(T, Z, select) for a real Schur form (rordschur.jl:3-45):
// returns 0, or k (1-based) when the swap moving block k was refused
// (IllConditionException(k)); w (n complex) receives the eigenvalues.
template <class R> long gordschur_re(long n, R* t, long ldt, R* z, long ldz, const int* select, cx<R>* w) {
  Mat<R> T{t, ldt};
  long ks = 0;
  bool pair = false;
  for (long k = 1; k <= n; ++k) {
    if (pair) {
      pair = false;
      continue;
    }
    bool swap = select[k - 1] != 0;
    if (k < n && T(k + 1, k) != R(0.0)) {
      pair = true;
      swap = swap || select[k] != 0;
    }
    if (swap) {
      ks += 1;
      if (k != ks) {
        long inew = ks;
        const bool ok = trexchange_re<R>(n, t, ldt, z, ldz, k, inew);
        ks = inew;
        if (!ok) return k;
      }
      if (pair) ks += 1;
    }
  }
  for (long k = 1; k <= n; ++k) w[k - 1] = cx<R>(T(k, k));
  for (long k = 1; k <= n - 1; ++k)
    if (T(k + 1, k) != R(0.0)) {
      const R y = rt<R>::sqrt(rt<R>::abs(T(k, k + 1))) * rt<R>::sqrt(rt<R>::abs(T(k + 1, k)));
      w[k - 1] = w[k - 1] + cx<R>(R(0.0), y);
      w[k] = w[k] - cx<R>(R(0.0), y);
    }
  return 0;
}
}  // namespace orc
