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
// Batches of LDS-resident orders run in three launches: heev_tq_kernel
// (one workgroup per matrix, the matrix in LDS: the tridiagonalization and
// Q), steqr_lanes_kernel (the QL/QR recurrence of 64 matrices per
// wavefront, one per lane, each writing its rotations to a record stream in
// HBM) and heev_apply_kernel (the streams replayed onto Q's rows in LDS, an
// in-place finish for a stream that overflowed, the stable sort, the
// write-out).  The upper-triangle form is reduced as the lower-triangle form
// of J A J (J the reversal permutation): the reference's U reduction
// applies, at each step, the same reflector as the L reduction of J A J with
// its tail in reverse order, so d, e and Q = J Q_L J follow by reversing
// indices (the tail's order only changes the rounding of its norm, which a
// parallel sum changes anyway).
//
// Orders above the LDS range run one matrix at a time over the whole GPU
// (the hb_* kernels: the reduction and Q as grid launches per column, the
// recurrence's record stream replayed onto Q's rows in rounds).
//
// steqr_kernel: the QL/QR iteration alone on (d, e) with Z in global memory
// (gschur!(H, Z) / _gschur!(A::SymTridiagonal, QRIteration(), Z)).
#include <algorithm>

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
// in M[p+2 :, p].  vb: LDS for two reflector vectors of n elements
// (alternating by step), w: LDS vector of n elements.  With two or more
// wavefronts, wavefront 0 updates the trailing block's first column and
// then forms the next reflector from it while the other wavefronts finish
// the rank-2 update (the next step's input column is final once its own
// update is done); the element operations are those of the step-by-step
// order.
template <int NT, class T>
GS_DEV void he_tridiag(T* M, int ld, int n, T* tau, T* vb, T* w, double* d, double* e, double* red) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int NW = NT / 64;
  constexpr bool OVL = NW >= 2;
  // reflector of step p into the vector buffer of p's parity (wavefront 0)
  auto reflect = [&](int p) {
    const int m = n - 1 - p;
    T* x = M + (p + 1) + (int64_t)p * ld;
    T* v = vb + (p & 1) * n;
    const T tp = he_reflector(x, m, lane);
    // each lane reads back the tail entries it scaled (x[1 + lane + 64 j])
    for (int k = 1 + lane; k < m; k += 64) v[k] = x[k];
    if (lane == 0) {
      v[0] = tfrom<T>(1.0);
      tau[p] = tp;
    }
  };
  if (tid == 0 && n > 0) M[0] = tfrom<T>(tre(M[0]));
  __syncthreads();
  if (OVL && wave == 0 && n >= 2) reflect(0);
  for (int p = 0; p <= n - 2; ++p) {
    const int m = n - 1 - p;
    T* x = M + (p + 1) + (int64_t)p * ld;
    T* B = M + (p + 1) + (int64_t)(p + 1) * ld;  // trailing block, m x m
    const T* v = vb + (p & 1) * n;
    if (!OVL && wave == 0) reflect(p);
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
      // rank2update!(B, -1, v, w, Val(:L)); with OVL wavefront 0 takes
      // column 0 alone, then the next reflector
      auto upd = [&](int c) {
        const T vc = v[c], wc = w[c];
        T* col = B + (int64_t)c * ld;
        if (tzero(vc) && tzero(wc)) {
          if (lane == 0) col[c] = tfrom<T>(tre(col[c]));
          return;
        }
        const T t1 = -tconj(wc), t2 = tconj(-vc);
        for (int r = c + lane; r < m; r += 64) {
          const T u = v[r] * t1 + w[r] * t2;
          col[r] = (r == c) ? tfrom<T>(tre(col[r]) + tre(u)) : col[r] + u;
        }
      };
      if (OVL) {
        if (wave == 0) {
          upd(0);
          if (p + 1 <= n - 2) reflect(p + 1);
        } else {
          // columns 1 .. m-1 over wavefronts 1 .. NW-1
          for (int c = wave; c < m; c += NW - 1) upd(c);
        }
      } else {
        for (int c = wave; c < m; c += NW) upd(c);
      }
    } else {
      if (tid == 0) B[0] = tfrom<T>(tre(B[0]));
      if (OVL && wave == 0 && p + 1 <= n - 2) reflect(p + 1);
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
  T* v = tau + n;  // two reflector buffers
  T* w = v + 2 * n;
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

// ------------------------------------------------------------------------
// Orders above the LDS-resident kernels (heev_max_n < n <= heev_big_max_n),
// one matrix at a time over the whole GPU, the same steps as heev_tq_kernel
// + steqr + sorteig!: _hehessenberg! on the lower triangle (hessenberg.jl:
// 27-60), Q by backward accumulation (:150-166), the QL/QR iteration
// (symtridiag.jl:77-427) on Q in HBM, the stable sort.  The working matrix
// stays in HBM in full Hermitian storage (the upper triangle the exact
// conjugate of the lower, the diagonal real), so the product with the
// trailing block is a column-major gemv.  Each reflector step is four
// launches (reflector, gemv partials, combine, rank-2 update), each
// reflector of Q's accumulation two.
template <class T> struct HbArgs {
  int n = 0, p = 0, ks = 1, kchunk = 0;
  T* M = nullptr;    // n x n, ld n
  T* tau = nullptr;  // n
  T* v = nullptr;    // n: the current reflector [1; tail]
  T* w = nullptr;    // n
  T* P = nullptr;    // ks x n gemv partials (ld n)
  T* Q = nullptr;    // n x n, ld n
  T* u = nullptr;    // n: Q accumulation dots
  double* d = nullptr;
  double* e = nullptr;
};

GS_DEV double twsum(double v) { return wave_sum(v); }
GS_DEV zd twsum(zd v) { return wave_zsum(v); }

// M = the lower triangle of A (of J A J for the upper form, J the exchange
// matrix) with its conjugate mirrored above and the diagonal real
template <class T> __global__ void __launch_bounds__(256) hb_load_kernel(HbArgs<T> a, const T* A, int64_t lda, int upper) {
  const int n = a.n;
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int i = (int)(idx % n), j = (int)(idx / n);
    const int li = i >= j ? i : j, lj = i >= j ? j : i;
    T x = upper ? A[(n - 1 - li) + (int64_t)(n - 1 - lj) * lda] : A[li + (int64_t)lj * lda];
    if (i == j) x = tfrom<T>(tre(x));
    else if (i < j) x = tconj(x);
    a.M[idx] = x;
  }
}

// reflector of step p on M[p+1 :, p]; v, tau[p], e[p], d[p]
template <class T> __global__ void __launch_bounds__(64) hb_reflect_kernel(HbArgs<T> a) {
  const int n = a.n, p = a.p, m = n - 1 - p, lane = threadIdx.x;
  T* x = a.M + (p + 1) + (int64_t)p * n;
  const T tp = he_reflector(x, m, lane);
  // each lane reads back the tail entries it scaled
  for (int k = 1 + lane; k < m; k += 64) a.v[k] = x[k];
  if (lane == 0) {
    a.v[0] = tfrom<T>(1.0);
    a.tau[p] = tp;
    a.e[p] = tre(x[0]);
    a.d[p] = tre(a.M[p + (int64_t)p * n]);
  }
}

// P[y][i] = sum over column chunk y of B(i, k) v[k], B the trailing block
template <class T> __global__ void __launch_bounds__(256) hb_gemv_kernel(HbArgs<T> a) {
  __shared__ double raw[4 * 64 * 2];
  T* red = reinterpret_cast<T*>(raw);
  const int n = a.n, p = a.p, m = n - 1 - p;
  if (tzero(a.tau[p])) return;
  const T* B = a.M + (p + 1) + (int64_t)(p + 1) * n;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int k0 = blockIdx.y * a.kchunk, k1 = min(m, k0 + a.kchunk);
  T acc = tfrom<T>(0.0);
  if (i < m)
    for (int k = k0 + wv; k < k1; k += 4) acc = acc + B[i + (int64_t)k * n] * a.v[k];
  red[wv * 64 + lane] = acc;
  __syncthreads();
  if (wv == 0 && i < m)
    a.P[(int64_t)blockIdx.y * n + i] = ((red[lane] + red[64 + lane]) + red[128 + lane]) + red[192 + lane];
}

// w = tau B v; w += -(tau / 2) (w^H v) v
template <class T> __global__ void __launch_bounds__(1024) hb_combine_kernel(HbArgs<T> a) {
  __shared__ double red[16];
  const int n = a.n, p = a.p, m = n - 1 - p, tid = threadIdx.x;
  const T tp = a.tau[p];
  if (tzero(tp)) return;
  T part = tfrom<T>(0.0);
  for (int i = tid; i < m; i += 1024) {
    T s = a.P[i];
    for (int q = 1; q < a.ks; ++q) s = s + a.P[(int64_t)q * n + i];
    const T wi = tp * s;
    a.w[i] = wi;
    part = part + tconj(wi) * a.v[i];
  }
  const T dot = tblock_sum<1024>(part, red);
  const T alpha = -(tp * 0.5) * dot;
  for (int i = tid; i < m; i += 1024) a.w[i] = a.w[i] + alpha * a.v[i];
}

// rank2update!(B, -1, v, w, Val(:L)) on the lower triangle, 32 x 32 tiles,
// each lower tile's conjugate written to its mirror
template <class T> __global__ void __launch_bounds__(256) hb_rank2_kernel(HbArgs<T> a) {
  __shared__ double raw[32 * 33 * 2];
  T* tile = reinterpret_cast<T*>(raw);
  const int n = a.n, p = a.p, m = n - 1 - p;
  const int I0 = blockIdx.x * 32, J0 = blockIdx.y * 32;
  if (I0 < J0) return;
  const T tp = a.tau[p];
  if (tzero(tp)) return;  // B unchanged; its diagonal is real already
  T* B = a.M + (p + 1) + (int64_t)(p + 1) * n;
  const int t = threadIdx.x, rl = t & 31, cg = t >> 5;
  const int r = I0 + rl;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int cl = cg + 8 * s, c = J0 + cl;
    if (r < m && c < m && r >= c) {
      const T vc = a.v[c], wc = a.w[c];
      const T old = B[r + (int64_t)c * n];
      T nw;
      if (tzero(vc) && tzero(wc)) {
        nw = (r == c) ? tfrom<T>(tre(old)) : old;
      } else {
        const T t1 = -tconj(wc), t2 = tconj(-vc);
        const T upd = a.v[r] * t1 + a.w[r] * t2;
        nw = (r == c) ? tfrom<T>(tre(old) + tre(upd)) : old + upd;
      }
      B[r + (int64_t)c * n] = nw;
      tile[cl * 33 + rl] = nw;
    }
  }
  __syncthreads();
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int rl2 = cg + 8 * s, cl2 = t & 31;
    const int rr = I0 + rl2, cc = J0 + cl2;
    if (rr < m && cc < m && rr > cc) B[cc + (int64_t)rr * n] = tconj(tile[cl2 * 33 + rl2]);
  }
}

template <class T> __global__ void __launch_bounds__(64) hb_tail_kernel(HbArgs<T> a) {
  if (threadIdx.x == 0) a.d[a.n - 1] = tre(a.M[(int64_t)(a.n - 1) * (a.n + 1)]);
}

template <class T> __global__ void __launch_bounds__(256) hb_identity_kernel(int n, T* Q) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256)
    Q[idx] = tfrom<T>((idx % n) == (idx / n) ? 1.0 : 0.0);
}

// reflector k = a.p applied to Q(c:, c+1:), c = k + 1: the dots
// u[j] = tau (Q[c, j] + sum_{r > c} conj(v_r) Q[r, j]), one wavefront per column
template <class T> __global__ void __launch_bounds__(256) hq_dots_kernel(HbArgs<T> a) {
  const int n = a.n, k = a.p, c = k + 1;
  const int j = c + 1 + blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (j >= n) return;
  T acc = tfrom<T>(0.0);
  for (int r = c + 1 + lane; r < n; r += 64) acc = acc + tconj(a.M[r + (int64_t)k * n]) * a.Q[r + (int64_t)j * n];
  acc = twsum(acc);
  if (lane == 0) a.u[j] = a.tau[k] * (a.Q[c + (int64_t)j * n] + acc);
}

// Q(r, j) -= u_j [1; v]_r for j > c; column c = e_c - tau [1; v]
template <class T> __global__ void __launch_bounds__(256) hq_update_kernel(HbArgs<T> a) {
  const int n = a.n, k = a.p, c = k + 1;
  const int r = c + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = c + blockIdx.y * 4 + (threadIdx.x >> 6);
  if (r >= n || j >= n) return;
  const T t = a.tau[k];
  T* q = a.Q + r + (int64_t)j * n;
  if (j == c) {
    *q = (r == c) ? tfrom<T>(1.0) - t : -(t * a.M[r + (int64_t)k * n]);
  } else {
    const T h = *q, sj = a.u[j];
    *q = (r == c) ? h - sj : h - sj * a.M[r + (int64_t)k * n];
  }
}

// outputs in the reference's coordinates: d, e (reversed for the upper
// form) and Q (J Q J for the upper form) into Zo (ld ldz)
template <class T> __global__ void __launch_bounds__(256) hb_out_kernel(HbArgs<T> a, int upper, double* dref,
                                                                       double* eref, T* Zo, int64_t ldz) {
  const int n = a.n;
  const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x, G = (int64_t)gridDim.x * 256;
  for (int64_t i = g; i < n; i += G) dref[i] = upper ? a.d[n - 1 - i] : a.d[i];
  for (int64_t i = g; i < n - 1; i += G) eref[i] = upper ? a.e[n - 2 - i] : a.e[i];
  if (Zo) {
    const int64_t nn = (int64_t)n * n;
    for (int64_t idx = g; idx < nn; idx += G) {
      const int r = (int)(idx % n), j = (int)(idx / n);
      Zo[r + (int64_t)j * ldz] = upper ? a.Q[(n - 1 - r) + (int64_t)(n - 1 - j) * n] : a.Q[idx];
    }
  }
}

// The QL/QR recurrence of one large matrix by one thread (d, e staged in
// LDS), from the start or (resume) from the state record; with rec its
// rotations go to a record stream in HBM, and a full stream stops it at a
// step boundary with its state left for the next round.
__global__ void __launch_bounds__(64) hb_produce_kernel(int n, int maxsweeps, double* d, double* e, double2* rec,
                                                       int cap, int* npos, StState* state, int resume) {
  extern __shared__ double smem[];
  double* dl = smem;
  double* el = smem + n;
  for (int i = threadIdx.x; i < n; i += 64) dl[i] = d[i];
  for (int i = threadIdx.x; i < n - 1; i += 64) el[i] = e[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    StState S;
    if (resume) S = *state;
    else st_init(S, n, maxsweeps);
    if (rec) {
      StLogG L;
      L.rec = rec;
      L.pos = 0;
      L.cap = cap;
      st_produce(S, StVec{dl, 1}, StVec{el, 1}, &L);
      *npos = L.pos;
    } else {
      st_produce<StLogG>(S, StVec{dl, 1}, StVec{el, 1}, nullptr);
      *npos = 0;
    }
    *state = S;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < n; i += 64) d[i] = dl[i];
  for (int i = threadIdx.x; i < n - 1; i += 64) e[i] = el[i];
}

// the record stream replayed onto rows 0 .. nrows-1 of Z (ld ldz), one row
// per thread
template <class T> __global__ void __launch_bounds__(64) hb_replay_kernel(int nrows, int64_t ldz, const double2* rec,
                                                                         const int* npos, T* Z) {
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= nrows) return;
  st_replay_row<T>(rec, *npos, [&](int k) -> T& { return Z[r + (int64_t)(k - 1) * ldz]; });
}

__global__ void __launch_bounds__(64) hb_info_kernel(const StState* state, int* info, int* sweeps) {
  if (threadIdx.x == 0) {
    if (info) *info = state->info;
    if (sweeps) *sweeps = state->sweeps;
  }
}

// sorteig! (stable, isless) of d into W; perm[j] = source column
__global__ void __launch_bounds__(1024) hb_sort_kernel(int n, int sort, const double* d, int* perm, double* W) {
  for (int i = threadIdx.x; i < n; i += 1024) {
    int rank = i;
    if (sort) {
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
  for (int j = threadIdx.x; j < n; j += 1024) W[j] = d[perm[j]];
}

template <class T> __global__ void __launch_bounds__(256) hb_permute_kernel(int n, const T* Zs, const int* perm, T* V,
                                                                           int64_t ldv) {
  const int64_t nn = (int64_t)n * n;
  for (int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x; idx < nn; idx += (int64_t)gridDim.x * 256) {
    const int r = (int)(idx % n), j = (int)(idx / n);
    V[r + (int64_t)j * ldv] = Zs[r + (int64_t)perm[j] * n];
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
  // 512-thread tridiagonalization workgroups above order 48 (complex) / 96
  // (real), 256 below (measured: complex n = 64 343k against 333k matrices/s,
  // n = 32 1.31M against 1.69M; real n = 128 85.0k against 73.8k, n = 64
  // 537k against 559k)
  if (cplx) return a.n > 48 ? launch_heev_t<zd, 512, 2>(a, true, s) : launch_heev_t<zd, 256, 2>(a, true, s);
  return a.n > 96 ? launch_heev_t<double, 512, 2>(a, false, s) : launch_heev_t<double, 256, 2>(a, false, s);
}

static int hb_blocks(int64_t elems) { return (int)std::min<int64_t>(4096, (elems + 255) / 256); }

size_t heev_big_ws_bytes(int n, bool cplx) {
  const size_t es = cplx ? 16 : 8, nn = (size_t)n * n;
  auto seg = [](size_t b) { return (b + 255) / 256 * 256; };
  return 3 * seg(nn * es) + 4 * seg((size_t)n * es) + seg(16 * (size_t)n * es) + 4 * seg(8 * (size_t)n) +
         seg(4 * (size_t)n) + seg(16 * heev_log_records(n)) + seg(4) + seg(sizeof(StState));
}

// The QL/QR iteration of one matrix (d, e in device memory, n <= 4096)
// with its rotations replayed onto rows 0 .. zrows-1 of Z (ld ldz; null:
// values only) in rounds of at most cap records: the producer thread, then
// the replay over the whole GPU, until the recurrence is done (the host
// reads the state record after each round; one round unless a stream
// fills).  The element operations are those of steqr_kernel.
template <class T> static hipError_t st_rounds(int n, int maxsweeps, double* d, double* e, T* Z, int zrows,
                                               int64_t ldz, double2* rec, int cap, int* npos, StState* state,
                                               int* info, int* sweeps, hipStream_t s) {
  const bool wantz = Z != nullptr && zrows > 0;
  cap = std::max(cap, n + 2);  // room for any one step: every round progresses
  for (int round = 0;; ++round) {
    hipLaunchKernelGGL(hb_produce_kernel, dim3(1), dim3(64), 16 * (size_t)n, s, n, maxsweeps, d, e,
                       wantz ? rec : nullptr, cap, npos, state, round > 0 ? 1 : 0);
    if (!wantz) break;
    hipLaunchKernelGGL(hb_replay_kernel<T>, dim3((zrows + 63) / 64), dim3(64), 0, s, zrows, ldz, rec, npos, Z);
    StState hs;
    hipError_t e2 = hipMemcpyAsync(&hs, state, sizeof hs, hipMemcpyDeviceToHost, s);
    if (e2 == hipSuccess) e2 = hipStreamSynchronize(s);
    if (e2 != hipSuccess) return e2;
    if (hs.done) break;
  }
  hipLaunchKernelGGL(hb_info_kernel, dim3(1), dim3(64), 0, s, state, info, sweeps);
  return hipGetLastError();
}

size_t steqr_big_ws_bytes(int n) {
  return (16 * heev_log_records(n) + 255) / 256 * 256 + 256 + (sizeof(StState) + 255) / 256 * 256;
}

// gs_steqr above the LDS log's range: the rounds of st_rounds on a.Z
hipError_t launch_steqr_big(const SteqrArgs& a, bool zcplx, void* ws, hipStream_t s) {
  if (a.n <= 0) return hipSuccess;
  if (a.n > heev_big_max_n() || !ws) return hipErrorInvalidValue;
  char* p = static_cast<char*>(ws);
  double2* rec = reinterpret_cast<double2*>(p);
  p += (16 * heev_log_records(a.n) + 255) / 256 * 256;
  int* npos = reinterpret_cast<int*>(p);
  StState* state = reinterpret_cast<StState*>(p + 256);
  const int cap = (int)heev_log_records(a.n);
  if (zcplx)
    return st_rounds<zd>(a.n, a.maxsweeps, a.d, a.e, static_cast<zd*>(a.Z), a.zrows, a.ldz, rec, cap, npos, state,
                         a.info, a.sweeps, s);
  return st_rounds<double>(a.n, a.maxsweeps, a.d, a.e, static_cast<double*>(a.Z), a.zrows, a.ldz, rec, cap, npos,
                           state, a.info, a.sweeps, s);
}

template <class T> static hipError_t launch_heev_big_t(const HeevArgs& a, bool cplx, void* ws, hipStream_t s) {
  const int n = a.n;
  const size_t es = sizeof(T), nn = (size_t)n * n;
  auto seg = [](size_t b) { return (b + 255) / 256 * 256; };
  char* p = static_cast<char*>(ws);
  HbArgs<T> h;
  h.n = n;
  h.M = reinterpret_cast<T*>(p);
  p += seg(nn * es);
  h.Q = reinterpret_cast<T*>(p);
  p += seg(nn * es);
  T* Zr = reinterpret_cast<T*>(p);  // Q in the reference's coordinates
  p += seg(nn * es);
  h.tau = reinterpret_cast<T*>(p);
  p += seg((size_t)n * es);
  h.v = reinterpret_cast<T*>(p);
  p += seg((size_t)n * es);
  h.w = reinterpret_cast<T*>(p);
  p += seg((size_t)n * es);
  h.u = reinterpret_cast<T*>(p);
  p += seg((size_t)n * es);
  h.P = reinterpret_cast<T*>(p);
  p += seg(16 * (size_t)n * es);
  h.d = reinterpret_cast<double*>(p);
  p += seg(8 * (size_t)n);
  h.e = reinterpret_cast<double*>(p);
  p += seg(8 * (size_t)n);
  double* dref = reinterpret_cast<double*>(p);
  p += seg(8 * (size_t)n);
  double* eref = reinterpret_cast<double*>(p);
  p += seg(8 * (size_t)n);
  int* perm = reinterpret_cast<int*>(p);
  p += seg(4 * (size_t)n);
  double2* rec = reinterpret_cast<double2*>(p);
  const int cap = (a.logcap > 0 && (size_t)a.logcap <= heev_log_records(n)) ? a.logcap : (int)heev_log_records(n);
  p += seg(16 * heev_log_records(n));
  int* npos = reinterpret_cast<int*>(p);
  p += seg(4);
  StState* state = reinterpret_cast<StState*>(p);
  const bool wantz = a.Z != nullptr;
  for (int b = 0; b < a.count; ++b) {
    const T* A = static_cast<const T*>(a.A) + (int64_t)b * a.sA;
    hipLaunchKernelGGL(hb_load_kernel<T>, dim3(hb_blocks((int64_t)nn)), dim3(256), 0, s, h, A, a.lda, a.upper);
    for (int st = 0; st + 2 <= n; ++st) {
      const int m = n - 1 - st;
      h.p = st;
      const int nbr = (m + 63) / 64;
      h.ks = std::max(1, std::min(16, (512 + nbr - 1) / nbr));
      h.kchunk = (m + h.ks - 1) / h.ks;
      h.ks = (m + h.kchunk - 1) / h.kchunk;
      const int nt = (m + 31) / 32;
      hipLaunchKernelGGL(hb_reflect_kernel<T>, dim3(1), dim3(64), 0, s, h);
      hipLaunchKernelGGL(hb_gemv_kernel<T>, dim3(nbr, h.ks), dim3(256), 0, s, h);
      hipLaunchKernelGGL(hb_combine_kernel<T>, dim3(1), dim3(1024), 0, s, h);
      hipLaunchKernelGGL(hb_rank2_kernel<T>, dim3(nt, nt), dim3(256), 0, s, h);
    }
    hipLaunchKernelGGL(hb_tail_kernel<T>, dim3(1), dim3(64), 0, s, h);
    const bool formq = wantz;
    if (formq) {
      hipLaunchKernelGGL(hb_identity_kernel<T>, dim3(hb_blocks((int64_t)nn)), dim3(256), 0, s, n, h.Q);
      for (int k = n - 2; k >= 0; --k) {
        const int c = k + 1;
        h.p = k;
        if (n - 1 - c > 0) hipLaunchKernelGGL(hq_dots_kernel<T>, dim3((n - 1 - c + 3) / 4), dim3(256), 0, s, h);
        hipLaunchKernelGGL(hq_update_kernel<T>, dim3((n - c + 63) / 64, (n - c + 3) / 4), dim3(256), 0, s, h);
      }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (a.mode == 2) {
      // tridiagonal form and Q only
      T* Zo = wantz ? static_cast<T*>(a.Z) + (int64_t)b * a.sZ : nullptr;
      hipLaunchKernelGGL(hb_out_kernel<T>, dim3(hb_blocks((int64_t)nn)), dim3(256), 0, s, h, a.upper,
                         a.W + (int64_t)b * a.sW, a.E + (int64_t)b * a.sE, Zo, a.ldz);
      continue;
    }
    hipLaunchKernelGGL(hb_out_kernel<T>, dim3(hb_blocks((int64_t)nn)), dim3(256), 0, s, h, a.upper, dref, eref,
                       wantz ? Zr : nullptr, (int64_t)n);
    int* infob = a.info ? a.info + b : nullptr;
    int* sweepsb = a.sweeps ? a.sweeps + b : nullptr;
    // the recurrence by one thread writing its rotations to HBM, their replay
    // onto Q's rows over the whole GPU (values only: the recurrence alone)
    if ((e = st_rounds<T>(n, a.maxsweeps, dref, eref, wantz ? Zr : nullptr, n, (int64_t)n, rec, cap, npos, state,
                          infob, sweepsb, s)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(hb_sort_kernel, dim3(1), dim3(1024), 0, s, n, a.sort, dref, perm, a.W + (int64_t)b * a.sW);
    if (wantz)
      hipLaunchKernelGGL(hb_permute_kernel<T>, dim3(hb_blocks((int64_t)nn)), dim3(256), 0, s, n, Zr, perm,
                         static_cast<T*>(a.Z) + (int64_t)b * a.sZ, a.ldz);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t launch_heev_big(const HeevArgs& a, bool cplx, void* ws, hipStream_t s) {
  if (a.count <= 0 || a.n <= 0) return hipSuccess;
  if (a.n > heev_big_max_n() || !ws) return hipErrorInvalidValue;
  return cplx ? launch_heev_big_t<zd>(a, true, ws, s) : launch_heev_big_t<double>(a, false, ws, s);
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
