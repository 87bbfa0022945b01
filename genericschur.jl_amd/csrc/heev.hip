// Hermitian / real symmetric eigenproblem on gfx950, following
// GenericSchur.jl's QRIteration path:
//   geigen! / geigvals!    GenericSchur.jl:124-146
//   tridiagonalization     hessenberg.jl:19-95 (_hehessenberg!, LAPACK xHETD2)
//                          with rank2update! (hessenberg.jl:97-130)
//   Q formation            hessenberg.jl:133-232 (_materializeQ)
//   Schur / eigensystem    symtridiag.jl:14-41 (gschur!(H::Hessenberg{SymTridiagonal}))
//   QL/QR iteration        symtridiag.jl:77-427 (steqr.h)
//   ordering               sorteig!(λ, V, eigsortby)
//
// heev_kernel: one workgroup per matrix, the matrix resident in LDS.  The
// upper-triangle form is reduced as the lower-triangle form of J A J (J the
// reversal permutation): the reference's U reduction applies, at each step,
// the same reflector as the L reduction of J A J with its tail in reverse
// order, so d, e and Q = J Q_L J follow by reversing indices (the tail's
// order only changes the rounding of its norm, which a parallel sum changes
// anyway).  Then the QL/QR iteration with one producer lane and the other
// wavefronts applying the rotation log to Q's rows, the stable sort and the
// write-out.
//
// steqr_kernel: the QL/QR iteration alone on (d, e) with Z in global memory
// (gschur!(H, Z) / _gschur!(A::SymTridiagonal, QRIteration(), Z)).
#include "dev_common.h"
#include "formq.h"
#include "internal.h"
#include "steqr.h"

namespace gs {

namespace {

GS_DEV double tconj(double x) { return x; }
GS_DEV zd tconj(zd x) { return zconj(x); }
GS_DEV double tre(double x) { return x; }
GS_DEV double tre(zd x) { return x.x; }
template <class T> GS_DEV T tfrom(double x);
template <> GS_DEV double tfrom<double>(double x) { return x; }
template <> GS_DEV zd tfrom<zd>(double x) { return zr(x); }
GS_DEV bool tzero(double x) { return x == 0.0; }
GS_DEV bool tzero(zd x) { return x.x == 0.0 && x.y == 0.0; }

// _reflector!(x) (householder.jl:12-102) on x[0 .. m-1] (LDS) by one
// wavefront: returns tau (the same in every lane); x[0] receives beta and
// x[1 ..] the scaled reflector tail.
GS_DEV double he_reflector(double* x, int m, int lane) {
  if (m <= 1) return 0.0;
  const double sfmin = 2.0 * GS_FMIN / GS_EPS;
  auto tailnorm = [&]() {
    double amax = 0.0;
    for (int r = 1 + lane; r < m; r += 64) amax = fmax(amax, fabs(x[r]));
    amax = wave_max(amax);
    if (amax == 0.0) return 0.0;
    const double ra = 1.0 / amax;
    double s = 0.0;
    for (int r = 1 + lane; r < m; r += 64) {
      const double a = x[r] * ra;
      s += a * a;
    }
    return amax * sqrt(wave_sum(s));
  };
  double alpha = x[0];
  double xnorm = tailnorm();
  if (xnorm == 0.0) return 0.0;
  double beta = -copysign(jl_hypot(alpha, xnorm), alpha);
  int kount = 0;
  if (fabs(beta) < sfmin) {
    const double rsfmin = 1.0 / sfmin;
    bool smallb = true;
    while (smallb) {
      kount += 1;
      for (int r = 1 + lane; r < m; r += 64) x[r] = x[r] * rsfmin;
      beta *= rsfmin;
      alpha *= rsfmin;
      smallb = (fabs(beta) < sfmin) && (kount < 20);
    }
    xnorm = tailnorm();
    beta = -copysign(jl_hypot(alpha, xnorm), alpha);
  }
  const double tau = (beta - alpha) / beta;
  const double t = 1.0 / (alpha - beta);
  for (int r = 1 + lane; r < m; r += 64) x[r] = x[r] * t;
  for (int j = 0; j < kount; ++j) beta *= sfmin;
  if (lane == 0) x[0] = beta;
  return tau;
}

GS_DEV zd he_reflector(zd* x, int m, int lane) {
  if (m < 1) return zr(0.0);
  const double sfmin = GS_FMIN / GS_EPS;
  auto tailnorm = [&]() {
    double amax = 0.0;
    for (int r = 1 + lane; r < m; r += 64) amax = fmax(amax, fmax(fabs(x[r].x), fabs(x[r].y)));
    amax = wave_max(amax);
    if (amax == 0.0) return 0.0;
    const double ra = 1.0 / amax;
    double s = 0.0;
    for (int r = 1 + lane; r < m; r += 64) {
      const double a = x[r].x * ra, b = x[r].y * ra;
      s += a * a + b * b;
    }
    return amax * sqrt(wave_sum(s));
  };
  zd alpha = x[0];
  double ar = alpha.x, ai = alpha.y;
  double xnorm = tailnorm();
  if (xnorm == 0.0 && ai == 0.0) return zr(0.0);
  double beta = -copysign(hypot3(ar, ai, xnorm), ar);
  int kount = 0;
  if (fabs(beta) < sfmin) {
    const double rsfmin = 1.0 / sfmin;
    bool smallb = true;
    while (smallb) {
      kount += 1;
      for (int r = 1 + lane; r < m; r += 64) x[r] = x[r] * rsfmin;
      beta *= rsfmin;
      ar *= rsfmin;
      ai *= rsfmin;
      smallb = (fabs(beta) < sfmin) && (kount < 20);
    }
    xnorm = tailnorm();
    alpha = zmk(ar, ai);
    beta = -copysign(hypot3(ar, ai, xnorm), ar);
  }
  const zd tau = zmk((beta - ar) / beta, -ai / beta);
  const zd t = zdiv(zr(1.0), alpha - zr(beta));
  for (int r = 1 + lane; r < m; r += 64) x[r] = x[r] * t;
  for (int j = 0; j < kount; ++j) beta *= sfmin;
  if (lane == 0) x[0] = zr(beta);
  return tau;
}

template <int NT, class F, class E> GS_DEV void tdots4e(int i0, int i1, int k0, int k1, F f, E epi, double) {
  ddots4e<NT>(i0, i1, k0, k1, f, epi);
}
template <int NT, class F, class E> GS_DEV void tdots4e(int i0, int i1, int k0, int k1, F f, E epi, zd) {
  zdots4e<NT>(i0, i1, k0, k1, f, epi);
}

template <int NT> GS_DEV double tblock_sum(double v, double* red) { return block_sum<NT>(v, red); }
template <int NT> GS_DEV zd tblock_sum(zd v, double* red) {
  const double x = block_sum<NT>(v.x, red);
  const double y = block_sum<NT>(v.y, red);
  return zmk(x, y);
}

template <int NT, int RM> GS_DEV void form_q(double* M, int ld, int n, const double* tau, double* wk) {
  d_form_q_inplace<NT, RM>(M, ld, n, tau, wk);
}
template <int NT, int RM> GS_DEV void form_q(zd* M, int ld, int n, const zd* tau, zd* wk) {
  z_form_q_inplace<NT, RM>(M, ld, n, tau, wk);
}

// _hehessenberg!(Hermitian(M, :L)) (hessenberg.jl:27-60) on the n x n matrix
// M in LDS (leading dimension ld; lower triangle read).  d (n), e (n-1)
// receive the tridiagonal matrix, tau (n-1) the scalars; reflector p stays
// in M[p+2 :, p].  v, w: LDS vectors of n elements.
template <int NT, class T>
GS_DEV void he_tridiag(T* M, int ld, int n, T* tau, T* v, T* w, double* d, double* e, double* red) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = NT / 64;
  if (tid == 0 && n > 0) M[0] = tfrom<T>(tre(M[0]));
  __syncthreads();
  for (int p = 0; p <= n - 2; ++p) {
    const int m = n - 1 - p;
    T* x = M + (p + 1) + (int64_t)p * ld;
    T* B = M + (p + 1) + (int64_t)(p + 1) * ld;  // trailing block, m x m
    if (wave == 0) {
      const T tp = he_reflector(x, m, lane);
      // each lane reads back the tail entries it scaled (x[1 + lane + 64 j])
      for (int k = 1 + lane; k < m; k += 64) v[k] = x[k];
      if (lane == 0) {
        v[0] = tfrom<T>(1.0);
        tau[p] = tp;
      }
    }
    __syncthreads();
    const T tp = tau[p];
    if (!tzero(tp)) {
      // w = tau * Hermitian(B, :L) v
      tdots4e<NT>(
          0, m, 0, m,
          [&](int k, int i) {
            T h;
            if (k < i) h = B[i + (int64_t)k * ld];
            else if (k == i) h = tfrom<T>(tre(B[i + (int64_t)i * ld]));
            else h = tconj(B[k + (int64_t)i * ld]);
            return h * v[k];
          },
          [&](int i, T acc) { w[i] = tp * acc; }, T());
      __syncthreads();
      // alpha = -(tau / 2) (w^H v);  w += alpha v
      T part = tfrom<T>(0.0);
      for (int k = tid; k < m; k += NT) part = part + tconj(w[k]) * v[k];
      const T dotwv = tblock_sum<NT>(part, red);
      const T alpha = -(tp * 0.5) * dotwv;
      for (int k = tid; k < m; k += NT) w[k] = w[k] + alpha * v[k];
      __syncthreads();
      // rank2update!(B, -1, v, w, Val(:L))
      for (int c = wave; c < m; c += NW) {
        const T vc = v[c], wc = w[c];
        T* col = B + (int64_t)c * ld;
        if (tzero(vc) && tzero(wc)) {
          if (lane == 0) col[c] = tfrom<T>(tre(col[c]));
          continue;
        }
        const T t1 = -tconj(wc), t2 = tconj(-vc);
        for (int r = c + lane; r < m; r += 64) {
          const T upd = v[r] * t1 + w[r] * t2;
          col[r] = (r == c) ? tfrom<T>(tre(col[r]) + tre(upd)) : col[r] + upd;
        }
      }
    } else if (tid == 0) {
      B[0] = tfrom<T>(tre(B[0]));
    }
    if (tid == 0) {
      e[p] = tre(x[0]);
      d[p] = tre(M[p + (int64_t)p * ld]);
    }
    __syncthreads();
  }
  if (tid == 0 && n > 0) d[n - 1] = tre(M[(n - 1) + (int64_t)(n - 1) * ld]);
  __syncthreads();
}

// Julia's isless on Float64: -0.0 before 0.0, NaN last.
GS_DEV bool jl_isless(double a, double b) {
  const bool an = a != a, bn = b != b;
  if (an || bn) return !an && bn;
  if (a == b) return __builtin_signbit(a) && !__builtin_signbit(b);
  return a < b;
}

// Log buffers and round control in LDS: buffer i holds c (cap doubles),
// s (cap doubles) and the segments (3 cap ints); the pointers are computed
// from the buffer index (an indexed pointer array would live in scratch).
struct StMem {
  double* base;
  int cap;
  int* ctrl;  // nrot[2], nseg[2], done[2]
  GS_DEV double* c(int i) const { return base + 2 * i * cap; }
  GS_DEV double* s(int i) const { return base + (2 * i + 1) * cap; }
  GS_DEV int* seg(int i) const { return reinterpret_cast<int*>(base + 4 * cap) + 3 * i * cap; }
};
GS_DEV StMem st_mem(void* base, int cap, int* ctrl) {
  StMem m;
  m.base = static_cast<double*>(base);
  m.cap = cap;
  m.ctrl = ctrl;
  return m;
}
constexpr size_t st_log_bytes(int cap) { return (size_t)cap * 56; }

// The QL/QR iteration on d, e (LDS, stride 1) from the producer state S
// (thread 0's copy counts) with the rotations applied to nrows rows of Z by
// waves 1 .. NT/64 - 1; zat(r, k) is the element of row r in the
// reference's column k (1-based).  On return thread 0's S is done.
template <int NT, class TZ, class ZAt>
GS_DEV void st_solve(StState& S, double* d, double* e, bool wantz, int nrows, ZAt zat, const StMem& mem) {
  const int tid = threadIdx.x;
  const StVec dv{d, 1}, ev{e, 1};
  if (!wantz) {
    if (tid == 0) st_produce<StLog>(S, dv, ev, nullptr);
    __syncthreads();
    return;
  }
  auto apply = [&](int bj) {
    StLog L;
    L.c = mem.c(bj);
    L.s = mem.s(bj);
    L.seg = mem.seg(bj);
    L.nrot = mem.ctrl[bj];
    L.nseg = mem.ctrl[2 + bj];
    L.cap = mem.cap;
    for (int r = tid - 64; r < nrows; r += NT - 64) st_apply_row<TZ>(L, [&](int k) -> TZ& { return zat(r, k); });
  };
  for (int k = 0;; ++k) {
    const int bi = k & 1;
    if (tid == 0) {
      StLog L;
      L.c = mem.c(bi);
      L.s = mem.s(bi);
      L.seg = mem.seg(bi);
      L.nrot = 0;
      L.nseg = 0;
      L.cap = mem.cap;
      st_produce(S, dv, ev, &L);
      mem.ctrl[bi] = L.nrot;
      mem.ctrl[2 + bi] = L.nseg;
      mem.ctrl[4 + bi] = S.done ? 1 : 0;
    } else if (tid >= 64 && k > 0) {
      apply(bi ^ 1);
    }
    __syncthreads();
    if (mem.ctrl[4 + bi]) {
      if (tid >= 64 && mem.ctrl[bi] > 0) apply(bi);
      __syncthreads();
      break;
    }
  }
}

}  // namespace

// ------------------------------------------------------------------------
// Phase 1: _hehessenberg! + _materializeQ of `count` matrices of order n,
// one workgroup per matrix, the matrix in LDS.  Writes d (W, stride sW), e
// (E, stride sE) in the reference's order and, when Z is non-null, Q (Z,
// ld ldz, stride sZ) in the reference's coordinates.
template <class T, int NT, int RM> __global__ void __launch_bounds__(NT) heev_tq_kernel(HeevArgs a) {
  extern __shared__ double smem[];
  const int b = blockIdx.x;
  const int n = a.n, ld = n | 1, tid = threadIdx.x;
  const bool upper = a.upper != 0;
  T* M = reinterpret_cast<T*>(smem);
  char* ub = reinterpret_cast<char*>(M + (size_t)ld * n);
  T* tau = reinterpret_cast<T*>(ub);
  T* v = tau + n;
  T* w = v + n;
  double* d = reinterpret_cast<double*>(ub + heev_union_bytes(n, sizeof(T)));
  double* e = d + n;
  double* red = e + n;
  // lower triangle of the working matrix (J A J for the upper form)
  const T* Ag = static_cast<const T*>(a.A) + (int64_t)b * a.sA;
  for (int idx = tid; idx < n * n; idx += NT) {
    const int i = idx % n, j = idx / n;
    if (i >= j) M[i + (int64_t)j * ld] = upper ? Ag[(n - 1 - i) + (int64_t)(n - 1 - j) * a.lda] : Ag[i + (int64_t)j * a.lda];
  }
  __syncthreads();
  he_tridiag<NT, T>(M, ld, n, tau, v, w, d, e, red);
  const bool wantq = a.Z != nullptr;
  if (wantq) form_q<NT, RM>(M, ld, n, tau, w);
  // d, e in the reference's order (reversed for the upper form)
  for (int i = tid; i < n; i += NT) a.W[(int64_t)b * a.sW + i] = upper ? d[n - 1 - i] : d[i];
  for (int i = tid; i < n - 1; i += NT) a.E[(int64_t)b * a.sE + i] = upper ? e[n - 2 - i] : e[i];
  if (wantq) {
    T* Zg = static_cast<T*>(a.Z) + (int64_t)b * a.sZ;
    for (int idx = tid; idx < n * n; idx += NT) {
      const int r = idx % n, j = idx / n;
      Zg[r + (int64_t)j * a.ldz] = upper ? M[(n - 1 - r) + (int64_t)(n - 1 - j) * ld] : M[r + (int64_t)j * ld];
    }
  }
}

// Phase 2: the QL/QR recurrence of 64 matrices per workgroup, one per lane
// (symtridiag.jl:77-427 via st_step), d and e interleaved in LDS (entry i of
// lane l at [i * 64 + l]), each lane's rotations to its record stream in
// HBM (LOG, cap records per matrix).  A lane whose stream would overflow
// stops at a step boundary and leaves its state for phase 3.
__global__ void __launch_bounds__(64) steqr_lanes_kernel(HeevArgs a) {
  extern __shared__ double smem[];
  const int lane = threadIdx.x, n = a.n;
  const int b = blockIdx.x * 64 + lane;
  if (b >= a.count) return;
  const StVec d{smem + lane, 64}, e{smem + (size_t)64 * n + lane, 64};
  for (int i = 0; i < n; ++i) d[i] = a.W[(int64_t)b * a.sW + i];
  for (int i = 0; i < n - 1; ++i) e[i] = a.E[(int64_t)b * a.sE + i];
  StState S;
  st_init(S, n, a.maxsweeps);
  StState* states = static_cast<StState*>(a.state);
  if (a.log) {
    StLogG L;
    L.rec = static_cast<double2*>(a.log) + (int64_t)b * a.logcap;
    L.pos = 0;
    L.cap = a.logcap;
    st_produce(S, d, e, &L);
    a.logpos[b] = L.pos;
  } else {
    st_produce<StLogG>(S, d, e, nullptr);
  }
  for (int i = 0; i < n; ++i) a.W[(int64_t)b * a.sW + i] = d[i];
  for (int i = 0; i < n - 1; ++i) a.E[(int64_t)b * a.sE + i] = e[i];
  states[b] = S;
}

// Phase 3: per matrix, the record stream replayed onto Q's rows (Q in LDS,
// one row per thread), the recurrence finished in place for a matrix whose
// stream overflowed (producer lane and appliers as for one matrix), the
// stable sort (sorteig!) and the write-out of W, V and info.
static_assert(heev_max_n(false) <= 128 && heev_max_n(true) <= 128, "phase 3 replays one row per thread");
template <class T, int NT> __global__ void __launch_bounds__(NT) heev_apply_kernel(HeevArgs a) {
  extern __shared__ double smem[];
  const int b = blockIdx.x;
  const int n = a.n, ld = n | 1, tid = threadIdx.x;
  const bool wantz = a.Z != nullptr;
  T* M = reinterpret_cast<T*>(smem);
  char* ub = reinterpret_cast<char*>(M + (wantz ? (size_t)ld * n : 0));
  double* d = reinterpret_cast<double*>(ub);
  double* e = d + n;
  int* ctrl = reinterpret_cast<int*>(e + n);
  int* perm = ctrl + 8;
  void* logb = static_cast<void*>(perm + ((n + 3) / 4) * 4);
  T* Zg = static_cast<T*>(a.Z) + (wantz ? (int64_t)b * a.sZ : 0);
  for (int i = tid; i < n; i += NT) d[i] = a.W[(int64_t)b * a.sW + i];
  for (int i = tid; i < n - 1; i += NT) e[i] = a.E[(int64_t)b * a.sE + i];
  if (wantz)
    for (int idx = tid; idx < n * n; idx += NT) {
      const int r = idx % n, j = idx / n;
      M[r + (int64_t)j * ld] = Zg[r + (int64_t)j * a.ldz];
    }
  __syncthreads();
  StState* states = static_cast<StState*>(a.state);
  StState S = states[b];
  if (wantz) {
    const double2* rec = static_cast<const double2*>(a.log) + (int64_t)b * a.logcap;
    const int npos = a.logpos[b];
    // one row per thread (n <= NT)
    const int r = tid;
    if (r < n) st_replay_row<T>(rec, npos, [&](int k) -> T& { return M[r + (int64_t)(k - 1) * ld]; });
    __syncthreads();
  }
  if (!S.done) {
    // the stream overflowed: finish the recurrence here
    const StMem mem = st_mem(logb, n, ctrl);
    st_solve<NT, T>(S, d, e, wantz, n, [&](int r, int k) -> T& { return M[r + (int64_t)(k - 1) * ld]; }, mem);
    if (tid == 0) states[b] = S;
    __syncthreads();
    S = states[b];
  }
  // sorteig!: stable ascending order under isless; perm[j] = source column
  for (int i = tid; i < n; i += NT) {
    int rank = i;
    if (a.sort) {
      const double di = d[i];
      rank = 0;
      for (int j = 0; j < n; ++j) {
        const double dj = d[j];
        rank += (jl_isless(dj, di) || (j < i && !jl_isless(di, dj))) ? 1 : 0;
      }
    }
    perm[rank] = i;
  }
  __syncthreads();
  for (int j = tid; j < n; j += NT) a.W[(int64_t)b * a.sW + j] = d[perm[j]];
  if (wantz)
    for (int idx = tid; idx < n * n; idx += NT) {
      const int r = idx % n, j = idx / n;
      Zg[r + (int64_t)j * a.ldz] = M[r + (int64_t)perm[j] * ld];
    }
  if (tid == 0) {
    if (a.info) a.info[b] = S.info;
    if (a.sweeps) a.sweeps[b] = S.sweeps;
  }
}

// _gschur!(SymTridiagonal(d, e), QRIteration(), Z): d, e in global memory
// (n, n - 1), Z (zrows x n, ld ldz; real or complex) updated in place.
template <class TZ, int NT> __global__ void __launch_bounds__(NT) steqr_kernel(SteqrArgs a) {
  extern __shared__ double smem[];
  const int n = a.n, tid = threadIdx.x;
  double* d = smem;
  double* e = d + n;
  int* ctrl = reinterpret_cast<int*>(e + n);
  void* logb = static_cast<void*>(ctrl + 16);
  for (int i = tid; i < n; i += NT) d[i] = a.d[i];
  for (int i = tid; i < n - 1; i += NT) e[i] = a.e[i];
  __syncthreads();
  TZ* Z = static_cast<TZ*>(a.Z);
  const StMem mem = st_mem(logb, n, ctrl);
  StState S;
  if (tid == 0) st_init(S, n, a.maxsweeps);
  st_solve<NT, TZ>(S, d, e, Z != nullptr, a.zrows, [&](int r, int k) -> TZ& { return Z[r + (int64_t)(k - 1) * a.ldz]; },
                   mem);
  for (int i = tid; i < n; i += NT) a.d[i] = d[i];
  for (int i = tid; i < n - 1; i += NT) a.e[i] = e[i];
  if (tid == 0) {
    if (a.info) a.info[0] = S.info;
    if (a.sweeps) a.sweeps[0] = S.sweeps;
  }
}

size_t heev_tq_lds_bytes(int n, bool cplx) {
  const size_t es = cplx ? 16 : 8;
  const size_t ld = (size_t)(n | 1);
  return ld * n * es + heev_union_bytes(n, es) + 16 * (size_t)n + 16 * 8 + 64;
}

size_t heev_apply_lds_bytes(int n, bool cplx, bool wantz) {
  const size_t es = cplx ? 16 : 8;
  const size_t ld = (size_t)(n | 1);
  return (wantz ? ld * n * es : 0) + 16 * (size_t)n + 32 + 4 * (size_t)((n + 3) / 4 * 4) + st_log_bytes(n) + 64;
}

size_t steqr_lds_bytes(int n) { return 16 * (size_t)n + 16 * 4 + st_log_bytes(n) + 64; }

size_t heev_log_records(int n) { return 2 * (size_t)n * n + 8 * (size_t)n; }

size_t heev_state_bytes() { return (sizeof(StState) + 15) / 16 * 16; }

template <class T, int NT, int RM> static hipError_t launch_heev_t(const HeevArgs& a, bool cplx, hipStream_t s) {
  const size_t l1 = heev_tq_lds_bytes(a.n, cplx);
  (void)hipFuncSetAttribute((const void*)heev_tq_kernel<T, NT, RM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)l1);
  hipLaunchKernelGGL((heev_tq_kernel<T, NT, RM>), dim3(a.count), dim3(NT), l1, s, a);
  if (a.mode == 2) return hipGetLastError();
  const size_t l2 = (size_t)1024 * a.n;
  (void)hipFuncSetAttribute((const void*)steqr_lanes_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)l2);
  hipLaunchKernelGGL(steqr_lanes_kernel, dim3((a.count + 63) / 64), dim3(64), l2, s, a);
  const size_t l3 = heev_apply_lds_bytes(a.n, cplx, a.Z != nullptr);
  (void)hipFuncSetAttribute((const void*)heev_apply_kernel<T, 128>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)l3);
  hipLaunchKernelGGL((heev_apply_kernel<T, 128>), dim3(a.count), dim3(128), l3, s, a);
  return hipGetLastError();
}

hipError_t launch_heev(const HeevArgs& a, bool cplx, hipStream_t s) {
  if (a.count <= 0 || a.n <= 0) return hipSuccess;
  if (a.n > heev_max_n(cplx)) return hipErrorInvalidValue;
  if (cplx) return launch_heev_t<zd, 256, 2>(a, true, s);
  return launch_heev_t<double, 256, 2>(a, false, s);
}

hipError_t launch_steqr(const SteqrArgs& a, bool zcplx, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  if (a.n > steqr_max_n()) return hipErrorInvalidValue;
  const size_t lds = steqr_lds_bytes(a.n);
  if (zcplx) {
    (void)hipFuncSetAttribute((const void*)steqr_kernel<zd, 256>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((steqr_kernel<zd, 256>), dim3(1), dim3(256), lds, s, a);
  } else {
    (void)hipFuncSetAttribute((const void*)steqr_kernel<double, 256>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL((steqr_kernel<double, 256>), dim3(1), dim3(256), lds, s, a);
  }
  return hipGetLastError();
}

}  // namespace gs
