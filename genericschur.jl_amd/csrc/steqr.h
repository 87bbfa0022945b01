// Symmetric tridiagonal QL/QR iteration with accumulation of the rotations,
// following GenericSchur.jl's _gschur!(A::SymTridiagonal, QRIteration(), Z)
// (symtridiag.jl:77-172, a translation of LAPACK xSTEQR) with _st_QLIter!
// (:175-262), _st_QRIter! (:265-352), _st_findsplit_fwd! (:354-372) and
// _dgz_sym2x2 (:375-427).
//
// Work split on the GPU: one lane (the producer) runs the scalar recurrence
// on d and e in LDS and appends every rotation it makes to a log in LDS;
// the other wavefronts of the workgroup (the appliers) apply the log to the
// rows of Z, each lane one row at a time, with the rotations of a sweep
// chained through a register.  Two log buffers alternate, so the appliers
// work on round k's rotations while the producer makes round k+1's; one
// workgroup barrier separates rounds.  Each element of Z receives exactly
// the reference's operations in the reference's order
// (rmul!(Z, Givens(i, i+1, c, s)'): Z[:, i] = c a1 + s a2,
// Z[:, i+1] = -s a1 + c a2), and the recurrence is evaluated without
// contraction, so d, e and Z match the CPU restatement bit for bit.
//
// The reference's sweep counter never advances (symtridiag.jl:141: the
// statement `niter + - nit` discards its value), so it has no iteration
// limit; here the producer stops after `maxsweeps` QL/QR sweeps in all
// (30 n, LAPACK's limit) and reports the number of nonzero off-diagonal
// entries left, so that every launch ends.
#pragma once
#include "dev_common.h"

namespace gs {

// LinearAlgebra.givensAlgorithm(f, g) for real f, g (LAPACK 3.1 xLARTG):
// floatmin2 = 2^-485 scaling and the sign fix when |f| > |g| and cs < 0.
GS_DEV void givens_re(double f, double g, double& cs, double& sn, double& r) {
#pragma clang fp contract(off)
  const double safmn2 = 0x1p-485, safmx2 = 0x1p485;
  if (g == 0.0) {
    cs = 1.0;
    sn = 0.0;
    r = f;
    return;
  }
  if (f == 0.0) {
    cs = 0.0;
    sn = 1.0;
    r = g;
    return;
  }
  double f1 = f, g1 = g;
  double sc = fmax(fabs(f1), fabs(g1));
  if (sc >= safmx2) {
    int count = 0;
    while (true) {
      count += 1;
      f1 = f1 * safmn2;
      g1 = g1 * safmn2;
      sc = fmax(fabs(f1), fabs(g1));
      if (sc < safmx2 || count >= 20) break;
    }
    r = sqrt(f1 * f1 + g1 * g1);
    cs = f1 / r;
    sn = g1 / r;
    for (int i = 0; i < count; ++i) r = r * safmx2;
  } else if (sc <= safmn2) {
    int count = 0;
    // the scaled value exceeds safmn2 after at most 3 steps for finite
    // nonzero input; the cap keeps a NaN from looping
    while (true) {
      count += 1;
      f1 = f1 * safmx2;
      g1 = g1 * safmx2;
      sc = fmax(fabs(f1), fabs(g1));
      if (sc > safmn2 || count >= 20) break;
    }
    r = sqrt(f1 * f1 + g1 * g1);
    cs = f1 / r;
    sn = g1 / r;
    for (int i = 0; i < count; ++i) r = r * safmn2;
  } else {
    r = sqrt(f1 * f1 + g1 * g1);
    cs = f1 / r;
    sn = g1 / r;
  }
  if (fabs(f) > fabs(g) && cs < 0.0) {
    cs = -cs;
    sn = -sn;
    r = -r;
  }
}

// _dgz_sym2x2(a, b, c, wantvec) (symtridiag.jl:375-427; LAPACK xLAEV2)
GS_DEV void sym2x2(double a, double b, double c, bool wantvec, double& rt1, double& rt2, double& cs1, double& sn1) {
#pragma clang fp contract(off)
  const double sm = a + c, df = a - c, adf = fabs(df), tb = 2.0 * b, ab = fabs(tb);
  double acmx, acmn;
  if (fabs(a) > fabs(c)) {
    acmx = a;
    acmn = c;
  } else {
    acmx = c;
    acmn = a;
  }
  const double rt = jl_hypot(ab, adf);
  int sgn1;
  if (sm < 0.0) {
    rt1 = (sm - rt) / 2.0;
    sgn1 = -1;
    rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
  } else if (sm > 0.0) {
    rt1 = (sm + rt) / 2.0;
    sgn1 = 1;
    rt2 = (acmx / rt1) * acmn - (b / rt1) * b;
  } else {
    rt1 = rt / 2.0;
    rt2 = -rt / 2.0;
    sgn1 = 1;
  }
  cs1 = 0.0;
  sn1 = 0.0;
  if (!wantvec) return;
  double cs;
  int sgn2;
  if (df > 0.0) {
    cs = df + rt;
    sgn2 = 1;
  } else {
    cs = df - rt;
    sgn2 = -1;
  }
  if (fabs(cs) > ab) {
    const double ct = -tb / cs;
    sn1 = 1.0 / sqrt(1.0 + ct * ct);
    cs1 = ct * sn1;
  } else if (ab == 0.0) {
    cs1 = 1.0;
    sn1 = 0.0;
  } else {
    const double tn = -cs / tb;
    cs1 = 1.0 / sqrt(1.0 + tn * tn);
    sn1 = tn * cs1;
  }
  if (sgn1 == sgn2) {
    const double tn = cs1;
    cs1 = -sn1;
    sn1 = tn;
  }
}

// d or e of the recurrence: entry i at p[i * stride] (stride 1 in LDS for
// one matrix, the lane count where each lane of a wavefront owns a matrix
// and the arrays are interleaved)
struct StVec {
  double* p;
  int stride;
  GS_DEV double& operator[](int i) const { return p[i * stride]; }
};

// safescale!(x, cfrom, cto) (util.jl:41-80) on x[i0 .. i0 + len - 1]
GS_DEV void st_safescale(StVec x, int i0, int len, double cfrom, double cto) {
  const double smlnum = GS_FMIN, bignum = 1.0 / smlnum;
  double cfromc = cfrom, ctoc = cto;
  bool done = false;
  for (int guard = 0; !done && guard < 64; ++guard) {
    double mul;
    const double cfrom1 = cfromc * smlnum;
    if (cfrom1 == cfromc) {
      mul = ctoc / cfromc;
      done = true;
    } else {
      const double cto1 = ctoc / bignum;
      if (cto1 == ctoc) {
        mul = cto;
        done = true;
        cfromc = 1.0;
      } else if (fabs(cfrom1) > fabs(ctoc) && ctoc != 0.0) {
        mul = smlnum;
        cfromc = 1.0;
      } else if (fabs(cto1) > fabs(cfromc)) {
        mul = bignum;
        cfromc = cfrom1;
      } else {
        mul = ctoc / cfromc;
        done = true;
      }
    }
    for (int i = 0; i < len; ++i) x[i0 + i] = x[i0 + i] * mul;
  }
}

// Rotation logs.  A segment is cnt rotations on column pairs (i, i+1)
// (1-based) with i = i0, i0 + dir, ...; rotation (c, s) maps
// Z[:, i] = c a1 + s a2, Z[:, i+1] = -s a1 + c a2.  The producer calls
// begin() before a segment's rotations, put() for each, end() after.
//
// StLog: in LDS, rotations in c[] and s[], segments in seg[] (i0, cnt, dir).
struct StLog {
  double* c;   // cap
  double* s;   // cap
  int* seg;    // 3 * cap: i0, cnt, dir
  int nrot, nseg, cap;
  GS_DEV bool room(int bound) const { return nrot + bound <= cap; }
  GS_DEV int begin() const { return nrot; }
  GS_DEV void put(int base, int k, double cc, double ss) {
    c[base + k] = cc;
    s[base + k] = ss;
  }
  GS_DEV void end(int base, int i0, int cnt, int dir) {
    (void)base;
    seg[3 * nseg] = i0;
    seg[3 * nseg + 1] = cnt;
    seg[3 * nseg + 2] = dir;
    nrot += cnt;
    nseg += 1;
  }
};

// StLogG: one matrix's record stream in global memory, 16-byte records: a
// segment header (i0, dir * cnt, as doubles) followed by its rotations
// (c, s).  pos counts the records written.
struct StLogG {
  double2* rec;
  int pos, cap;
  GS_DEV bool room(int bound) const { return pos + 1 + bound <= cap; }
  GS_DEV int begin() const { return pos + 1; }
  GS_DEV void put(int base, int k, double cc, double ss) { rec[base + k] = make_double2(cc, ss); }
  GS_DEV void end(int base, int i0, int cnt, int dir) {
    rec[base - 1] = make_double2((double)i0, (double)(dir * cnt));
    pos = base + cnt;
  }
};

// Producer state (symtridiag.jl:77-172 unrolled into resumable steps).
struct StState {
  int n, l1, l, lend, lsave, lendsave, iscale, sweeps, maxsweeps, info;
  bool inblk, ql, stop, done;
  double anorm;
};

GS_DEV void st_init(StState& S, int n, int maxsweeps) {
  S.n = n;
  S.l1 = 1;
  S.l = S.lend = S.lsave = S.lendsave = 0;
  S.iscale = 0;
  S.sweeps = 0;
  S.maxsweeps = maxsweeps;
  S.info = 0;
  S.inblk = S.ql = S.stop = false;
  S.done = n <= 0;
  S.anorm = 0.0;
}

// Upper bound on the rotations the next st_step can log.
GS_DEV int st_next_bound(const StState& S) { return S.inblk ? abs(S.lend - S.l) : S.n - 1; }

// One step of the producer: the set-up of the next unreduced block, one QL
// or QR iteration, or the end of a block.  d, e: 0-based; D(i) = d[i-1] is
// the reference's 1-based d[i].  Rotations go to lg when non-null.
template <class Log> GS_DEV void st_step(StState& S, StVec d, StVec e, Log* lg) {
#pragma clang fp contract(off)
  const int n = S.n;
  const double eps = GS_EPS, eps2 = GS_EPS * GS_EPS, safmin = GS_FMIN;
  const double ssfmax = sqrt(1.0 / GS_FMIN) / 3.0;
  const double ssfmin = sqrt(GS_FMIN) / eps2;
#define D(i) d[(i) - 1]
#define E(i) e[(i) - 1]
  if (!S.inblk) {
    if (S.l1 > n) {
      S.done = true;
      return;
    }
    if (S.l1 > 1) E(S.l1 - 1) = 0.0;
    // _st_findsplit_fwd!
    int msplit = n;
    for (int m = S.l1; m <= n - 1; ++m) {
      const double tst = fabs(E(m));
      if (tst == 0.0) {
        msplit = m;
        break;
      }
      if (tst <= sqrt(fabs(D(m))) * sqrt(fabs(D(m + 1))) * eps) {
        E(m) = 0.0;
        msplit = m;
        break;
      }
    }
    S.l = S.l1;
    S.lsave = S.l;
    S.lend = msplit;
    S.lendsave = msplit;
    S.l1 = msplit + 1;
    if (S.lend == S.l) return;
    // opnorm(SymTridiagonal(d[l:lend], e[l:lend-1]), Inf): largest row sum
    double anorm = 0.0;
    for (int i = S.l; i <= S.lend; ++i) {
      double s = 0.0;
      if (i > S.l) s = s + fabs(E(i - 1));
      s = s + fabs(D(i));
      if (i < S.lend) s = s + fabs(E(i));
      anorm = fmax(anorm, s);
    }
    S.anorm = anorm;
    S.iscale = 0;
    if (anorm == 0.0) return;
    if (anorm > ssfmax) {
      S.iscale = 1;
      st_safescale(d, S.l - 1, S.lend - S.l + 1, anorm, ssfmax);
      st_safescale(e, S.l - 1, S.lend - S.l, anorm, ssfmax);
    } else if (anorm < ssfmin) {
      S.iscale = 2;
      st_safescale(d, S.l - 1, S.lend - S.l + 1, anorm, ssfmin);
      st_safescale(e, S.l - 1, S.lend - S.l, anorm, ssfmin);
    }
    if (fabs(D(S.lend)) < fabs(D(S.l))) {
      S.lend = S.lsave;
      S.l = S.lendsave;
    }
    S.ql = S.lend > S.l;
    S.inblk = true;
    return;
  }
  const bool more = !S.stop && (S.ql ? (S.l <= S.lend) : (S.l >= S.lend));
  if (!more) {
    // end of the block: undo the scaling
    const int ls = S.lsave, le = S.lendsave;
    if (S.iscale == 1) {
      st_safescale(d, ls - 1, le - ls + 1, ssfmax, S.anorm);
      st_safescale(e, ls - 1, le - ls, ssfmax, S.anorm);
    } else if (S.iscale == 2) {
      st_safescale(d, ls - 1, le - ls + 1, ssfmin, S.anorm);
      st_safescale(e, ls - 1, le - ls, ssfmin, S.anorm);
    }
    S.inblk = false;
    if (S.stop) {
      int nbad = 0;
      for (int i = 1; i <= n - 1; ++i) nbad += (E(i) != 0.0) ? 1 : 0;
      S.info = nbad > 0 ? nbad : 1;
      S.done = true;
    }
    return;
  }
  // _st_QLIter! (dir = +1: l < lend, the block swept towards l from its far
  // end) and _st_QRIter! (dir = -1) loop bodies as one: the two are mirror
  // images, with the same operations on mirrored indices, so lanes of a
  // wavefront in either iteration share the path.
  const int l = S.l, lend = S.lend;
  const int dir = S.ql ? 1 : -1;
  const bool wantz = lg != nullptr;
  // search for a small off-diagonal entry: pairs (m, m + dir), e index
  // m (QL) or m - 1 (QR)
  // (four tests at a time with their loads issued together; the first
  // pair that passes, as one test after another)
  auto small = [&](int m) {
    const int em = dir > 0 ? m : m - 1;
    return E(em) * E(em) <= (eps2 * fabs(D(m))) * fabs(D(m + dir)) + safmin;
  };
  int msplit = lend;
  {
    const int len = dir > 0 ? lend - l : l - lend;
    int t = 0;
    for (; t + 4 <= len; t += 4) {
      const int m0 = l + dir * t;
      const bool h0 = small(m0), h1 = small(m0 + dir), h2 = small(m0 + 2 * dir), h3 = small(m0 + 3 * dir);
      if (h0 || h1 || h2 || h3) {
        msplit = h0 ? m0 : (h1 ? m0 + dir : (h2 ? m0 + 2 * dir : m0 + 3 * dir));
        t = len + 1;
        break;
      }
    }
    for (; t < len; ++t)
      if (small(l + dir * t)) {
        msplit = l + dir * t;
        break;
      }
  }
  if (msplit != lend) E(dir > 0 ? msplit : msplit - 1) = 0.0;
  double p = D(l);
  if (msplit != l) {
    if (msplit == l + dir) {
      // the remaining 2 x 2 block (lo, lo + 1)
      const int lo = dir > 0 ? l : l - 1;
      double rt1, rt2, cs, sn;
      sym2x2(D(lo), E(lo), D(lo + 1), wantz, rt1, rt2, cs, sn);
      if (wantz) {
        const int base = lg->begin();
        lg->put(base, 0, cs, sn);
        lg->end(base, lo, 1, 1);
      }
      D(lo) = rt1;
      D(lo + 1) = rt2;
      E(lo) = 0.0;
      S.l = l + 2 * dir;
      return;
    }
    if (S.sweeps >= S.maxsweeps) {
      S.stop = true;
      return;
    }
    S.sweeps += 1;
    const int el = dir > 0 ? l : l - 1;
    double g = (D(l + dir) - p) / (2.0 * E(el));
    double r = jl_hypot(g, 1.0);
    g = D(msplit) - p + (E(el) / (g + copysign(r, g)));
    double s = 1.0, c = 1.0;
    p = 0.0;
    // rotations j = j0, j0 - dir, ..., j1 on pairs (j, j + 1): QL updates
    // d[j + 1] against d[j], QR d[j] against d[j + 1]
    const int j0 = dir > 0 ? msplit - 1 : msplit;
    const int cnt = dir > 0 ? msplit - l : l - msplit;
    const int r0 = wantz ? lg->begin() : 0;
    // the entries of rotation k + 1 (e[jj - dir], d[up - dir], d[far - dir])
    // are loaded during rotation k, which writes only e[jj + dir], d[up]
    double ej = 0.0, dup = 0.0, dfar = 0.0;
    if (cnt > 0) {
      ej = E(j0);
      dup = D(dir > 0 ? j0 + 1 : j0);
      dfar = D(dir > 0 ? j0 : j0 + 1);
    }
    for (int k = 0; k < cnt; ++k) {
      const int jj = j0 - dir * k;
      const int up = dir > 0 ? jj + 1 : jj, far = dir > 0 ? jj : jj + 1;
      const double ek = ej, dupk = dup, dfark = dfar;
      if (k + 1 < cnt) {
        ej = E(jj - dir);
        dup = D(up - dir);
        dfar = D(far - dir);
      }
      const double f = s * ek;
      const double b = c * ek;
      givens_re(g, f, c, s, r);
      if (k > 0) E(jj + dir) = r;
      g = dupk - p;
      r = (dfark - g) * s + 2.0 * c * b;
      p = s * r;
      D(up) = g + p;
      g = c * r - b;
      if (wantz) lg->put(r0, k, c, dir > 0 ? -s : s);
    }
    if (wantz) lg->end(r0, j0, cnt, -dir);
    D(l) = D(l) - p;
    E(el) = g;
  } else {
    D(l) = p;
    S.l = l + dir;
  }
#undef D
#undef E
}

// Runs the producer until the log cannot take another step's rotations or
// the recurrence is done.
template <class Log> GS_DEV void st_produce(StState& S, StVec d, StVec e, Log* lg) {
  while (!S.done) {
    if (lg && !lg->room(st_next_bound(S))) break;
    st_step(S, d, e, lg);
  }
}

// Real / complex Z element helpers for the appliers
GS_DEV double st_rot1(double a1, double a2, double c, double s) {
#pragma clang fp contract(off)
  return a1 * c + a2 * s;
}
GS_DEV double st_rot2(double a1, double a2, double c, double s) {
#pragma clang fp contract(off)
  return (-a1) * s + a2 * c;
}
GS_DEV zd st_rot1(zd a1, zd a2, double c, double s) { return zmk(st_rot1(a1.x, a2.x, c, s), st_rot1(a1.y, a2.y, c, s)); }
GS_DEV zd st_rot2(zd a1, zd a2, double c, double s) { return zmk(st_rot2(a1.x, a2.x, c, s), st_rot2(a1.y, a2.y, c, s)); }

// Applies the log to one row of Z: z(k) is the element of the row in the
// reference's (1-based) column k.
template <class TZ, class ZA> GS_DEV void st_apply_row(const StLog& lg, ZA z) {
  int off = 0;
  for (int q = 0; q < lg.nseg; ++q) {
    const int i0 = lg.seg[3 * q], cnt = lg.seg[3 * q + 1], dir = lg.seg[3 * q + 2];
    if (dir < 0) {
      TZ cur = z(i0 + 1);
      for (int k = 0; k < cnt; ++k) {
        const int i = i0 - k;
        const double c = lg.c[off + k], s = lg.s[off + k];
        const TZ a1 = z(i);
        z(i + 1) = st_rot2(a1, cur, c, s);
        cur = st_rot1(a1, cur, c, s);
      }
      z(i0 - cnt + 1) = cur;
    } else {
      TZ cur = z(i0);
      for (int k = 0; k < cnt; ++k) {
        const int i = i0 + k;
        const double c = lg.c[off + k], s = lg.s[off + k];
        const TZ a2 = z(i + 1);
        z(i) = st_rot1(cur, a2, c, s);
        cur = st_rot2(cur, a2, c, s);
      }
      z(i0 + cnt) = cur;
    }
    off += cnt;
  }
}

// A zero the compiler cannot see through: an index that makes a load of a
// uniform address a vector load, counted apart from the LDS accesses (a
// scalar load would share their counter, so every LDS wait would wait for
// it too).
GS_DEV int st_opaque_zero() {
  int z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// Replays a record stream (StLogG) of npos records onto one row of Z.  The
// rotations of a segment are taken four at a time with the four columns
// they read loaded first: within a segment each column is read by one
// rotation before the next rotation writes it, so the element operations
// and their order are those of one rotation after another.  The (c, s)
// records are vector loads kept four groups ahead in four register
// buffers, so that their latency overlaps the rotations of the groups
// before (indices clamped into the segment; a clamped load is not used).
template <class TZ, class ZA> GS_DEV void st_replay_row(const double2* __restrict__ rec, int npos, ZA z) {
  const int vz = st_opaque_zero();
  int q = 0;
  while (q < npos) {
    const double2 h = rec[q];
    const int i0 = (int)h.x, dc = (int)h.y;
    const int cnt = dc < 0 ? -dc : dc;
    const double2* r = rec + q + 1 + vz;
    const bool back = dc < 0;
    TZ cur = back ? z(i0 + 1) : z(i0);
    // four rotations k .. k + 3 with records c[0..3]
    auto grp = [&](int k, const double2* c) {
      if (back) {
        const int i = i0 - k;
        const TZ a0 = z(i), a1 = z(i - 1), a2 = z(i - 2), a3 = z(i - 3);
        z(i + 1) = st_rot2(a0, cur, c[0].x, c[0].y);
        cur = st_rot1(a0, cur, c[0].x, c[0].y);
        z(i) = st_rot2(a1, cur, c[1].x, c[1].y);
        cur = st_rot1(a1, cur, c[1].x, c[1].y);
        z(i - 1) = st_rot2(a2, cur, c[2].x, c[2].y);
        cur = st_rot1(a2, cur, c[2].x, c[2].y);
        z(i - 2) = st_rot2(a3, cur, c[3].x, c[3].y);
        cur = st_rot1(a3, cur, c[3].x, c[3].y);
      } else {
        const int i = i0 + k;
        const TZ a0 = z(i + 1), a1 = z(i + 2), a2 = z(i + 3), a3 = z(i + 4);
        z(i) = st_rot1(cur, a0, c[0].x, c[0].y);
        cur = st_rot2(cur, a0, c[0].x, c[0].y);
        z(i + 1) = st_rot1(cur, a1, c[1].x, c[1].y);
        cur = st_rot2(cur, a1, c[1].x, c[1].y);
        z(i + 2) = st_rot1(cur, a2, c[2].x, c[2].y);
        cur = st_rot2(cur, a2, c[2].x, c[2].y);
        z(i + 3) = st_rot1(cur, a3, c[3].x, c[3].y);
        cur = st_rot2(cur, a3, c[3].x, c[3].y);
      }
    };
    auto ld4 = [&](double2* b, int kk) {
      const int last = cnt - 1;
      b[0] = r[min(kk, last)];
      b[1] = r[min(kk + 1, last)];
      b[2] = r[min(kk + 2, last)];
      b[3] = r[min(kk + 3, last)];
    };
    int k = 0;
    if (cnt >= 4) {
      double2 b0[4], b1[4], b2[4], b3[4];
      ld4(b0, 0);
      ld4(b1, 4);
      ld4(b2, 8);
      ld4(b3, 12);
      for (; k + 16 <= cnt; k += 16) {
        grp(k, b0);
        ld4(b0, k + 16);
        grp(k + 4, b1);
        ld4(b1, k + 20);
        grp(k + 8, b2);
        ld4(b2, k + 24);
        grp(k + 12, b3);
        ld4(b3, k + 28);
      }
      // at most three more full groups, already in b0, b1, b2
      if (k + 4 <= cnt) {
        grp(k, b0);
        k += 4;
      }
      if (k + 4 <= cnt) {
        grp(k, b1);
        k += 4;
      }
      if (k + 4 <= cnt) {
        grp(k, b2);
        k += 4;
      }
    }
    for (; k < cnt; ++k) {
      const double2 cs = r[k];
      if (back) {
        const int i = i0 - k;
        const TZ a1 = z(i);
        z(i + 1) = st_rot2(a1, cur, cs.x, cs.y);
        cur = st_rot1(a1, cur, cs.x, cs.y);
      } else {
        const int i = i0 + k;
        const TZ a2 = z(i + 1);
        z(i) = st_rot1(cur, a2, cs.x, cs.y);
        cur = st_rot2(cur, a2, cs.x, cs.y);
      }
    }
    if (back) z(i0 - cnt + 1) = cur;
    else z(i0 + cnt) = cur;
    q += 1 + cnt;
  }
}

}  // namespace gs
