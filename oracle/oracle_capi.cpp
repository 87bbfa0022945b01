// TEST INFRASTRUCTURE ONLY — C entry points of the CPU oracle (see
// gschur_oracle.hpp).  Loaded by tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py through ctypes; never by the product library.
//
// Complex arrays are interleaved (re, im) doubles, column-major, like Julia's
// ComplexF64 and C `double _Complex`.
#include <omp.h>

#include <cstring>
#include <vector>

#include "gschur_oracle.hpp"

using namespace orc;
typedef cx<double> zc;

extern "C" {

int orc_abi_version() { return 4; }

// gschur!(A::StridedMatrix{ComplexF64}) — GenericSchur.jl:350-372
int orc_schur_z(long n, double* a, long lda, double* t, long ldt, double* z, long ldz, double* w, int wantz,
                int scale, long maxiter, long maxinner, long* stats) {
  return gschur<double>(n, (zc*)a, lda, (zc*)t, ldt, (zc*)z, ldz, (zc*)w, wantz != 0, scale != 0, maxiter,
                        maxinner, stats);
}

// gschur!(A::StridedMatrix{Float64}) — GenericSchur.jl:805-835
int orc_schur_d(long n, double* a, long lda, double* t, long ldt, double* z, long ldz, double* w, int wantz,
                int scale, long maxiter, long* stats) {
  return gschur<double>(n, a, lda, t, ldt, z, ldz, (zc*)w, wantz != 0, scale != 0, maxiter, 0, stats);
}

// gschur!(H::Hessenberg{ComplexF64}, Z) — GenericSchur.jl:194-335.  On exit h
// holds triu(HH).  z may be NULL (no Schur vectors).
int orc_schur_hess_z(long n, double* h, long ldh, double* z, long ldz, long maxiter, long maxinner, int checksd,
                     double* w) {
  int info = schur_hess_complex<double>(n, (zc*)h, ldh, (zc*)z, ldz, maxiter, maxinner, checksd != 0, (zc*)w);
  if (info == ORC_OK) {
    zc* H = (zc*)h;
    for (long j = 0; j < n; ++j)
      for (long i = j + 1; i < n; ++i) H[i + j * ldh] = zc();
  }
  return info;
}

// gschur!(H::Hessenberg{Float64}, Z) — GenericSchur.jl:513-699
int orc_schur_hess_d(long n, double* h, long ldh, double* z, long ldz, long maxiter, double* w) {
  return schur_hess_real<double>(n, h, ldh, z, ldz, maxiter, (zc*)w);
}

// _hessenberg! + _materializeQ — hessenberg.jl:3-17,150-166.  q may be NULL.
void orc_hessenberg_z(long n, double* a, long lda, double* tau, double* q, long ldq) {
  hessenberg<zc>(n, (zc*)a, lda, (zc*)tau);
  if (q) materialize_q<zc>(n, (zc*)a, lda, (zc*)tau, (zc*)q, ldq);
}
void orc_hessenberg_d(long n, double* a, long lda, double* tau, double* q, long ldq) {
  hessenberg<double>(n, a, lda, tau);
  if (q) materialize_q<double>(n, a, lda, tau, q, ldq);
}

// _reflector! — householder.jl:12-102.  Returns tau through `tau`.
void orc_reflector_z(long m, double* x, double* tau) {
  zc t = reflector<double>(m, (zc*)x, 1);
  tau[0] = t.re;
  tau[1] = t.im;
}
void orc_reflector_d(long m, double* x, double* tau) { tau[0] = reflector<double>(m, x, 1); }

// _gs2x2! — GenericSchur.jl:716-803.  abcd in/out; cssn = (cs, sn);
// w = (re w1, im w1, re w2, im w2).
void orc_gs2x2(double* abcd, double* cssn, double* w) {
  zc w1, w2;
  gs2x2<double>(abcd[0], abcd[1], abcd[2], abcd[3], cssn[0], cssn[1], w1, w2);
  w[0] = w1.re;
  w[1] = w1.im;
  w[2] = w2.re;
  w[3] = w2.im;
}

// geigvecs(S; left) — vectors.jl:12-20 (+ _enormalize! when normalize != 0).
// t is modified temporarily and restored (complex case).  z may be NULL.
void orc_eigvecs_z(long n, double* t, long ldt, const double* z, long ldz, double* v, long ldv, int left,
                   int normalize) {
  if (left)
    eigvecs_left_complex<double>(n, (zc*)t, ldt, (const zc*)z, ldz, (zc*)v, ldv);
  else
    eigvecs_right_complex<double>(n, (zc*)t, ldt, (const zc*)z, ldz, (zc*)v, ldv);
  if (normalize) enormalize<double>(n, (zc*)v, ldv);
}
void orc_eigvecs_d(long n, const double* t, long ldt, const double* z, long ldz, double* v, long ldv, int left,
                   int normalize) {
  if (left)
    eigvecs_left_real<double>(n, t, ldt, z, ldz, (zc*)v, ldv);
  else
    eigvecs_right_real<double>(n, t, ldt, z, ldz, (zc*)v, ldv);
  if (normalize) enormalize<double>(n, (zc*)v, ldv);
}

void orc_enormalize(long n, double* v, long ldv) { enormalize<double>(n, (zc*)v, ldv); }

// triangularize — triang.jl:9-43 (tc, zc are n x n, ld = n)
void orc_triangularize(long n, const double* t, long ldt, const double* z, long ldz, double* tc, double* zcm) {
  triangularize<double>(n, t, ldt, z, ldz, (zc*)tc, (zc*)zcm);
}

// Batched drivers over packed column-major matrices at element offsets
// (`off`), eigenvalues at the exclusive prefix sum of n.  OpenMP over
// matrices with `nthreads` threads (<= 0: library default).
int orc_schur_batched_z(long count, const int* n, const long* off, const double* A, double* T, double* Z, double* W,
                        int* info, int nthreads, int wantz) {
  std::vector<long> woff(count + 1, 0);
  for (long i = 0; i < count; ++i) woff[i + 1] = woff[i] + n[i];
  if (nthreads > 0) omp_set_num_threads(nthreads);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
  for (long i = 0; i < count; ++i) {
    const long m = n[i];
    std::vector<zc> a((size_t)m * m);
    std::memcpy((void*)a.data(), A + 2 * off[i], sizeof(zc) * m * m);
    info[i] = gschur<double>(m, a.data(), m, (zc*)(T + 2 * off[i]), m, (zc*)(Z + 2 * off[i]), m,
                             (zc*)(W + 2 * woff[i]), wantz != 0, true, 100 * m, 30 * m);
    bad += (info[i] != 0);
  }
  return bad;
}

int orc_schur_batched_d(long count, const int* n, const long* off, const double* A, double* T, double* Z, double* W,
                        int* info, int nthreads, int wantz) {
  std::vector<long> woff(count + 1, 0);
  for (long i = 0; i < count; ++i) woff[i + 1] = woff[i] + n[i];
  if (nthreads > 0) omp_set_num_threads(nthreads);
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
  for (long i = 0; i < count; ++i) {
    const long m = n[i];
    std::vector<double> a((size_t)m * m);
    std::memcpy((void*)a.data(), A + off[i], sizeof(double) * m * m);
    info[i] = gschur<double>(m, a.data(), m, T + off[i], m, Z + off[i], m, (zc*)(W + 2 * woff[i]), wantz != 0, true,
                             100 * m, 0);
    bad += (info[i] != 0);
  }
  return bad;
}

}  // extern "C"

// ------------------------------------------------ extended precision
// Eigenvalues of the (exactly representable) double matrix A computed by
// the same restated algorithm in quad-double (prec = 2: about 212 bits) or
// __float128 (prec = 1: 113 bits).  w receives (re_hi, re_lo, im_hi, im_lo)
// per eigenvalue.  Returns the info code.
template <class R> static void put_hilo(const R& x, double* out) {
  double hi = rt<R>::to_double(x);
  R rem = x - R(rt<R>::from(hi));
  out[0] = hi;
  out[1] = rt<R>::to_double(rem);
}

template <class R> static int eig_hp_z(long n, const double* a, double* w) {
  std::vector<cx<R>> A((size_t)n * n), T((size_t)n * n), W(n);
  for (long i = 0; i < n * n; ++i) A[i] = cx<R>(rt<R>::from(a[2 * i]), rt<R>::from(a[2 * i + 1]));
  int info = gschur<R>(n, A.data(), n, T.data(), n, nullptr, n, W.data(), false, true, 100 * n, 30 * n);
  for (long i = 0; i < n; ++i) {
    put_hilo(W[i].re, w + 4 * i);
    put_hilo(W[i].im, w + 4 * i + 2);
  }
  return info;
}
template <class R> static int eig_hp_d(long n, const double* a, double* w) {
  std::vector<R> A((size_t)n * n), T((size_t)n * n);
  std::vector<cx<R>> W(n);
  for (long i = 0; i < n * n; ++i) A[i] = rt<R>::from(a[i]);
  int info = gschur<R>(n, A.data(), n, T.data(), n, nullptr, n, W.data(), false, true, 100 * n, 0);
  for (long i = 0; i < n; ++i) {
    put_hilo(W[i].re, w + 4 * i);
    put_hilo(W[i].im, w + 4 * i + 2);
  }
  return info;
}

extern "C" {

int orc_eig_hp_z(long n, const double* a, double* w, int prec) {
  return prec >= 2 ? eig_hp_z<qd_real>(n, a, w) : eig_hp_z<__float128>(n, a, w);
}
int orc_eig_hp_d(long n, const double* a, double* w, int prec) {
  return prec >= 2 ? eig_hp_d<qd_real>(n, a, w) : eig_hp_d<__float128>(n, a, w);
}

// quad-double primitives, exported so tests can pin the arithmetic itself
void orc_qd_sqrt(const double* a, double* r) {
  qd_real x(a[0], a[1], a[2], a[3]);
  qd_real s = qd_sqrt(x);
  for (int i = 0; i < 4; ++i) r[i] = s.x[i];
}
void orc_qd_div(const double* a, const double* b, double* r) {
  qd_real x(a[0], a[1], a[2], a[3]), y(b[0], b[1], b[2], b[3]);
  qd_real s = x / y;
  for (int i = 0; i < 4; ++i) r[i] = s.x[i];
}
void orc_qd_mul(const double* a, const double* b, double* r) {
  qd_real x(a[0], a[1], a[2], a[3]), y(b[0], b[1], b[2], b[3]);
  qd_real s = x * y;
  for (int i = 0; i < 4; ++i) r[i] = s.x[i];
}
void orc_qd_add(const double* a, const double* b, double* r) {
  qd_real x(a[0], a[1], a[2], a[3]), y(b[0], b[1], b[2], b[3]);
  qd_real s = x + y;
  for (int i = 0; i < 4; ++i) r[i] = s.x[i];
}


// balance!(A; scale, permute) — balance.jl:33-199.  ilo_ihi_triv receives
// {ilo, ihi, trivial}; sp (n) the permutation record; d (n) the scale
// factors.  Returns 0, or -3 for a NaN met while scaling.
int orc_balance_z(long n, double* a, long lda, int scale, int permute, long* ilo_ihi_triv, long* sp, double* d) {
  long ilo = 1, ihi = n;
  bool triv = true;
  int r = balance<zc>(n, (zc*)a, lda, scale != 0, permute != 0, ilo, ihi, sp, d, triv);
  ilo_ihi_triv[0] = ilo;
  ilo_ihi_triv[1] = ihi;
  ilo_ihi_triv[2] = triv ? 1 : 0;
  return r;
}
int orc_balance_d(long n, double* a, long lda, int scale, int permute, long* ilo_ihi_triv, long* sp, double* d) {
  long ilo = 1, ihi = n;
  bool triv = true;
  int r = balance<double>(n, a, lda, scale != 0, permute != 0, ilo, ihi, sp, d, triv);
  ilo_ihi_triv[0] = ilo;
  ilo_ihi_triv[1] = ihi;
  ilo_ihi_triv[2] = triv ? 1 : 0;
  return r;
}
// lmul!(B, V) / ldiv!(B, V) — balance.jl:201-261, V complex n x m
void orc_balance_apply(long n, long m, const long* ilo_ihi_triv, const long* sp, const double* d, double* v, long ldv,
                       int inverse) {
  balance_apply<double>(n, m, ilo_ihi_triv[0], ilo_ihi_triv[1], sp, d, ilo_ihi_triv[2] != 0, (zc*)v, ldv,
                        inverse != 0);
}
// gordschur! (complex; select: n ints), eigvalscond, subspacesep
void orc_ordschur_z(long n, double* t, long ldt, double* z, long ldz, int wantz, const int* select, double* w) {
  gordschur<double>(n, (zc*)t, ldt, (zc*)z, ldz, wantz != 0, select, (zc*)w);
}
double orc_eigvalscond_z(long n, const double* t, long ldt, long nsub) {
  return eigvalscond<double>(n, (const zc*)t, ldt, nsub);
}
double orc_subspacesep_z(long n, const double* t, long ldt, long nsub) {
  return subspacesep<double>(n, (const zc*)t, ldt, nsub);
}
// trsylvester! (adj = 0) / adjtrsylvester! (adj = 1) on complex A, B, C
double orc_trsylvester_z(long m, const double* a, long lda, long n, const double* b, long ldb, double* c, long ldc,
                         int adj, int* perturbed) {
  bool p = false;
  const double s = adj ? adjtrsylvester<double>(m, (const zc*)a, lda, n, (const zc*)b, ldb, (zc*)c, ldc, p)
                       : trsylvester<double>(m, (const zc*)a, lda, n, (const zc*)b, ldb, (zc*)c, ldc, p);
  *perturbed = p ? 1 : 0;
  return s;
}
void orc_givens_z(const double* f, const double* g, double* out5) {
  double cs;
  zc sn, r;
  givens_cx<double>(zc(f[0], f[1]), zc(g[0], g[1]), cs, sn, r);
  out5[0] = cs;
  out5[1] = sn.re;
  out5[2] = sn.im;
  out5[3] = r.re;
  out5[4] = r.im;
}

// ---- Hermitian / real symmetric path ----
// _hehessenberg! (hessenberg.jl:27-95) + _materializeQ (:133-232): on exit a
// holds the reflectors, d/e the tridiagonal, tau the scalars, q (n x n, may
// be NULL) the transformation.  upper selects the triangle of a that is read.
void orc_hetrd_z(long n, double* a, long lda, int upper, double* d, double* e, double* tau, double* q, long ldq) {
  hetrd<zc>(n, (zc*)a, lda, upper != 0, (zc*)tau, d, e);
  if (q) {
    if (upper) materialize_q_u<zc>(n, (zc*)a, lda, (zc*)tau, (zc*)q, ldq);
    else materialize_q<zc>(n, (zc*)a, lda, (zc*)tau, (zc*)q, ldq);
  }
}
void orc_hetrd_d(long n, double* a, long lda, int upper, double* d, double* e, double* tau, double* q, long ldq) {
  hetrd<double>(n, a, lda, upper != 0, tau, d, e);
  if (q) {
    if (upper) materialize_q_u<double>(n, a, lda, tau, q, ldq);
    else materialize_q<double>(n, a, lda, tau, q, ldq);
  }
}
// _gschur!(SymTridiagonal(d, e), QRIteration(), Z) (symtridiag.jl:77-172);
// z (zrows x n, complex when zcplx) may be NULL; maxsweeps < 0: 30 n.
long orc_steqr(long n, double* d, double* e, double* z, long ldz, long zrows, int zcplx, long maxsweeps,
               long* sweeps) {
  if (zcplx) return steqr<double, zc>(n, d, e, (zc*)z, ldz, z ? zrows : 0, maxsweeps, sweeps);
  return steqr<double, double>(n, d, e, z, ldz, z ? zrows : 0, maxsweeps, sweeps);
}
// geigen! / geigvals! (GenericSchur.jl:124-146) on a Hermitian / real
// symmetric matrix; v may be NULL.
long orc_heev_z(long n, double* a, long lda, int upper, double* w, double* v, long ldv, int sort, long* sweeps) {
  return heev<zc>(n, (zc*)a, lda, upper != 0, w, (zc*)v, ldv, sort != 0, sweeps);
}
long orc_heev_d(long n, double* a, long lda, int upper, double* w, double* v, long ldv, int sort, long* sweeps) {
  return heev<double>(n, a, lda, upper != 0, w, v, ldv, sort != 0, sweeps);
}
// heev over a batch of uniform order n (column-major, stride n*n), OpenMP
// over matrices; the CPU baseline of the batched Hermitian eigensolver.
void orc_heev_batched(long count, long n, int cplx, double* a, double* w, double* v, int sort, int nthreads) {
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const long es = cplx ? 2 : 1;
#pragma omp parallel for schedule(dynamic)
  for (long k = 0; k < count; ++k) {
    double* ak = a + k * n * n * es;
    double* vk = v ? v + k * n * n * es : nullptr;
    if (cplx) heev<zc>(n, (zc*)ak, n, false, w + k * n, (zc*)vk, n, sort != 0, nullptr);
    else heev<double>(n, ak, n, false, w + k * n, vk, n, sort != 0, nullptr);
  }
}
void orc_givens_d(double f, double g, double* out3) {
  givens_re<double>(f, g, out3[0], out3[1], out3[2]);
}
void orc_sym2x2(double a, double b, double c, double* out4) {
  sym2x2<double>(a, b, c, true, out4[0], out4[1], out4[2], out4[3]);
}
double orc_jhypot(double x, double y) { return jhypot<double>(x, y); }

// gordschur!(T, Z, select) for a real Schur form (rordschur.jl:3-45): T, Z
// (may be NULL) overwritten, w (n complex) the eigenvalues; returns 0 or the
// 1-based index k of IllConditionException(k).
long orc_ordschur_d(long n, double* t, long ldt, double* z, long ldz, const int* select, double* w) {
  return gordschur_re<double>(n, t, ldt, z, ldz, select, (zc*)w);
}
// _syl1or2(Tl, Tr, B) (rordschur.jl:300-466): x (2 x 2 column-major) and scale
double orc_syl1or2(int n1, int n2, const double* tl, const double* tr, const double* b, double* x) {
  return syl1or2<double>(n1, n2, tl, tr, b, x);
}
}  // extern "C"
