// Reordering and conditioning of a complex Schur form: gordschur!
// (ordschur.jl:23-73), and eigvalscond / subspacesep (ordschur.jl:85-131)
// with the triangular Sylvester solvers (sylvester.jl:24-139) and the
// one-norm estimator (norm1est.jl:21-100).  Off the QR hot path: each kernel
// is one wavefront per matrix; the swaps, the Sylvester recurrences and the
// estimator iterations are sequential, the vector work along them is spread
// over the lanes.
#include <hip/hip_runtime.h>

#include "dev_common.h"
#include "internal.h"

#pragma clang fp contract(off)

namespace gs {

namespace {

// Julia's floatmin2(Float64) = 2^-485 and its reciprocal
constexpr double kSafmn2 = 0x1p-485, kSafmx2 = 0x1p485;

GS_DEV double amax1(zd z) { return fmax(fabs(z.x), fabs(z.y)); }

// LinearAlgebra.givensAlgorithm(f, g) for complex arguments (the xLARTG
// scheme of LAPACK 3.x with power-of-two scaling): cs, sn with
// [cs sn; -conj(sn) cs] [f; g] = [r; 0].
GS_DEV void zgivens(zd f, zd g, double& cs, zd& sn) {
  const double safmin = GS_FMIN;
  double scalepar = fmax(amax1(f), amax1(g));
  zd fs = f, gs = g;
  int count = 0;
  if (scalepar >= kSafmx2) {
    while (true) {
      count += 1;
      fs = fs * kSafmn2;
      gs = gs * kSafmn2;
      scalepar = scalepar * kSafmn2;
      if (scalepar < kSafmx2 || count >= 20) break;
    }
  } else if (scalepar <= kSafmn2) {
    if (g.x == 0.0 && g.y == 0.0) {
      cs = 1.0;
      sn = zr(0.0);
      return;
    }
    while (true) {
      count -= 1;
      fs = fs * kSafmx2;
      gs = gs * kSafmx2;
      scalepar = scalepar * kSafmx2;
      if (scalepar > kSafmn2) break;
    }
  }
  const double f2 = zabs2(fs), g2 = zabs2(gs);
  if (f2 <= fmax(g2, 1.0) * safmin) {
    if (f.x == 0.0 && f.y == 0.0) {
      cs = 0.0;
      const double d = hypot(gs.x, gs.y);
      sn = zmk(gs.x / d, -gs.y / d);
      return;
    }
    const double f2s = hypot(fs.x, fs.y);
    const double g2s = sqrt(g2);
    cs = f2s / g2s;
    zd ff;
    if (amax1(f) > 1.0) {
      const double d = hypot(f.x, f.y);
      ff = zmk(f.x / d, f.y / d);
    } else {
      const double dr = kSafmx2 * f.x, di = kSafmx2 * f.y;
      const double d = hypot(dr, di);
      ff = zmk(dr / d, di / d);
    }
    sn = ff * zmk(gs.x / g2s, -gs.y / g2s);
  } else {
    const double f2s = sqrt(1.0 + g2 / f2);
    const zd r = zmk(f2s * fs.x, f2s * fs.y);
    cs = 1.0 / f2s;
    const double d = f2 + g2;
    sn = zmk(r.x / d, r.y / d) * zconj(gs);
  }
}

GS_DEV zd rdz0(zd v) { return zmk(rdfirst(v.x), rdfirst(v.y)); }

}  // namespace

// gordschur!(T, Z, select): one wavefront per matrix, T and Z in global
// memory.  Each adjacent swap (_trexchange!): lane 0 forms the rotation,
// then the lanes apply lmul!(G, T) to rows k, k+1 and rmul!(T, G') to
// columns k, k+1 over the whole matrix, as the reference does (the rounding
// residue the swaps leave below the diagonal enters later swaps' diagonal
// entries), and rmul!(Z, G') to Z's columns k, k+1.  Finally T below the
// diagonal is zeroed (triu!) and W = diag(T).
__global__ void __launch_bounds__(64) ordschur_kernel(int n, zd* T, int64_t ldt, zd* Z, int64_t ldz, int wantz,
                                                      const int32_t* select, zd* W) {
  const int lane = threadIdx.x;
  auto t = [&](int i, int j) -> zd& { return T[(i - 1) + (int64_t)(j - 1) * ldt]; };
  auto z = [&](int i, int j) -> zd& { return Z[(i - 1) + (int64_t)(j - 1) * ldz]; };
  int ks = 0;
  for (int kk = 1; kk <= n; ++kk) {
    if (!select[kk - 1]) continue;
    ks += 1;
    if (kk == ks) continue;
    // move kk to ks: swaps at k = kk - 1 down to ks
    for (int k = kk - 1; k >= ks; --k) {
      double cs = 0.0;
      zd sn = zr(0.0);
      if (lane == 0) zgivens(t(k, k + 1), t(k + 1, k + 1) - t(k, k), cs, sn);
      cs = rdfirst(cs);
      sn = rdz0(sn);
      __syncthreads();
      for (int j = 1 + lane; j <= n; j += 64) {
        const zd a1 = t(k, j), a2 = t(k + 1, j);
        t(k, j) = cs * a1 + sn * a2;
        t(k + 1, j) = (-zconj(sn)) * a1 + cs * a2;
      }
      __syncthreads();
      // G' = Givens(cs, -sn)
      const zd ms = -sn;
      for (int i = 1 + lane; i <= n; i += 64) {
        const zd a1 = t(i, k), a2 = t(i, k + 1);
        t(i, k) = a1 * cs - a2 * zconj(ms);
        t(i, k + 1) = a1 * ms + a2 * cs;
      }
      if (wantz)
        for (int i = 1 + lane; i <= n; i += 64) {
          const zd a1 = z(i, k), a2 = z(i, k + 1);
          z(i, k) = a1 * cs - a2 * zconj(ms);
          z(i, k + 1) = a1 * ms + a2 * cs;
        }
      __syncthreads();
    }
  }
  for (int j = 1; j <= n; ++j)
    for (int i = j + 1 + lane; i <= n; i += 64) t(i, j) = zr(0.0);
  __syncthreads();
  for (int j = 1 + lane; j <= n; j += 64) W[j - 1] = t(j, j);
}

namespace {

// A = T(1:m, 1:m), B = T(m+1:n, m+1:n) of the triangular T (ld = ldt)
struct Blocks {
  const zd* T;
  int64_t ld;
  int m, n2;
  GS_DEV zd a(int i, int j) const { return T[(i - 1) + (int64_t)(j - 1) * ld]; }
  GS_DEV zd b(int i, int j) const { return T[(m + i - 1) + (int64_t)(m + j - 1) * ld]; }
};

GS_DEV double wmax(double v) { return wave_max(v); }

// smin of the Sylvester solvers: max(small, eps * maxabs(A) * maxabs(B))
GS_DEV double syl_smin(const Blocks& bl) {
  double ma = 0.0, mb = 0.0;
  for (int j = 1; j <= bl.m; ++j)
    for (int i = 1 + (int)threadIdx.x; i <= bl.m; i += 64) ma = fmax(ma, zabs(bl.a(i, j)));
  for (int j = 1; j <= bl.n2; ++j)
    for (int i = 1 + (int)threadIdx.x; i <= bl.n2; i += 64) mb = fmax(mb, zabs(bl.b(i, j)));
  ma = wmax(ma);
  mb = wmax(mb);
  const double small = GS_FMIN * (double)bl.m * (double)bl.n2 / GS_EPS;
  return fmax(small, GS_EPS * ma * mb);
}

// trsylvester! (adj = false) or adjtrsylvester! (adj = true) on X (m x n2,
// ld m) in global memory; returns scale.  Lane-parallel dot products, lane 0
// the scalar update of each entry.
GS_DEV double sylvester(const Blocks& bl, zd* X, bool adj, double smin) {
  const int m = bl.m, n2 = bl.n2, lane = threadIdx.x;
  const double small = GS_FMIN * (double)m * (double)n2 / GS_EPS;
  const double bignum = 1.0 / small;
  auto x = [&](int i, int j) -> zd& { return X[(i - 1) + (int64_t)(j - 1) * m]; };
  double scale = 1.0;
  for (int li = 1; li <= n2; ++li) {
    const int l = adj ? n2 + 1 - li : li;
    for (int ki = 1; ki <= m; ++ki) {
      const int k = adj ? ki : m + 1 - ki;
      zd suml = zr(0.0), sumr = zr(0.0);
      if (!adj) {
        for (int j = k + 1 + lane; j <= m; j += 64) suml = suml + bl.a(k, j) * x(j, l);
        for (int j = 1 + lane; j <= l - 1; j += 64) sumr = sumr + x(k, j) * bl.b(j, l);
      } else {
        for (int j = 1 + lane; j <= k - 1; j += 64) suml = suml + zconj(bl.a(j, k)) * x(j, l);
        for (int j = l + 1 + lane; j <= n2; j += 64) sumr = sumr + zconj(x(k, j)) * bl.b(l, j);
      }
      suml = wave_zsum(suml);
      sumr = wave_zsum(sumr);
      double scaloc = 1.0;
      zd xkl = zr(0.0);
      if (lane == 0) {
        const zd v = adj ? x(k, l) - (suml - zconj(sumr)) : x(k, l) - (suml - sumr);
        zd a11 = adj ? zconj(bl.a(k, k) - bl.b(l, l)) : bl.a(k, k) - bl.b(l, l);
        double da11 = zabs1(a11);
        if (adj ? (da11 < smin) : (da11 <= smin)) {
          a11 = zr(smin);
          da11 = smin;
        }
        const double db = zabs1(v);
        if (da11 < 1.0 && db > 1.0 && db > bignum * da11) scaloc = 1.0 / db;
        const zd vs = v * scaloc;
        xkl = (a11.y == 0.0) ? zmk(vs.x / a11.x, vs.y / a11.x) : vs / a11;
      }
      scaloc = rdfirst(scaloc);
      if (scaloc != 1.0) {
        for (int e = lane; e < m * n2; e += 64) X[e] = X[e] * scaloc;
        scale = scale * scaloc;
      }
      __syncthreads();
      if (lane == 0) x(k, l) = xkl;
      __syncthreads();
    }
  }
  return scale;
}

GS_DEV double norm1v(const zd* x, int n) {
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 64) s = s + zabs(x[i]);
  return wave_sum(s);
}

// _findamax of a complex vector: the first index of the largest |x|^2
GS_DEV double findamax(const zd* x, int n, int& i0) {
  double best = -1.0;
  int bi = 0x7fffffff;
  for (int i = threadIdx.x; i < n; i += 64) {
    const double a = zabs2(x[i]);
    if (a > best) {
      best = a;
      bi = i;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double ob = __shfl_xor(best, o, 64);
    const int oi = __shfl_xor(bi, o, 64);
    if (ob > best || (ob == best && oi < bi)) {
      best = ob;
      bi = oi;
    }
  }
  i0 = bi;
  return sqrt(best);
}

GS_DEV void sgn(zd* x, int n) {
  const double tiny = GS_FMIN;
  for (int i = threadIdx.x; i < n; i += 64) {
    const double a = zabs(x[i]);
    x[i] = a > tiny ? zmk(x[i].x / a, x[i].y / a) : zr(1.0);
  }
  __syncthreads();
}

}  // namespace

// eigvalscond and subspacesep of the triangular T (n x n) split at nsub
// (0 < nsub < n): out[0] = s, out[1] = sep.  work holds 3 m (n - m)
// complex entries (m = nsub): the Sylvester right-hand side / solution and
// the estimator's vectors x, y.
__global__ void __launch_bounds__(64) eigcond_kernel(int n, const zd* T, int64_t ldt, int nsub, zd* work,
                                                     double* out) {
  const int lane = threadIdx.x;
  const Blocks bl{T, ldt, nsub, n - nsub};
  const int m = nsub, n2 = n - nsub, N = m * n2;
  zd* X = work;
  zd* x = work + N;
  zd* y = work + 2 * N;
  const double smin = syl_smin(bl);
  // eigvalscond: T11 R - R T22 = scale T12
  for (int e = lane; e < N; e += 64) X[e] = T[(e % m) + (int64_t)(m + e / m) * ldt];
  __syncthreads();
  double scale = sylvester(bl, X, false, smin);
  double rn = 0.0;
  for (int e = lane; e < N; e += 64) rn = rn + zabs2(X[e]);
  rn = sqrt(wave_sum(rn));
  const double s = rn > 0.0 ? scale / (sqrt(scale * scale / rn + rn) * sqrt(rn)) : 1.0;
  // subspacesep: scale / norm1est of the Sylvester operator
  for (int e = lane; e < N; e += 64) x[e] = zr(1.0 / (double)N);
  __syncthreads();
  scale = sylvester(bl, x, false, smin);
  double est;
  if (N == 1) {
    est = zabs(x[0]);
  } else {
    double gamma = norm1v(x, N);
    sgn(x, N);
    scale = sylvester(bl, x, true, smin);
    int j0;
    findamax(x, N, j0);
    for (int iter = 1; iter <= 5; ++iter) {
      for (int e = lane; e < N; e += 64) x[e] = zr(e == j0 ? 1.0 : 0.0);
      __syncthreads();
      scale = sylvester(bl, x, false, smin);
      for (int e = lane; e < N; e += 64) y[e] = x[e];
      __syncthreads();
      const double oldg = gamma;
      gamma = norm1v(y, N);
      if (gamma <= oldg) break;
      sgn(x, N);
      scale = sylvester(bl, x, true, smin);
      const int jlast = j0;
      const double ax0 = findamax(x, N, j0);
      if (zabs(x[jlast]) == ax0) break;
    }
    for (int e = lane; e < N; e += 64) {
      const double sg = (e & 1) ? -1.0 : 1.0;
      x[e] = zr(sg * (1.0 + (double)e / (double)(N - 1)));
    }
    __syncthreads();
    scale = sylvester(bl, x, false, smin);
    const double tt = 2.0 * norm1v(x, N) / (3.0 * (double)N);
    est = fmax(tt, gamma);
  }
  if (lane == 0) {
    out[0] = s;
    out[1] = scale / est;
  }
}

hipError_t launch_ordschur(int n, double* T, int64_t ldt, double* Z, int64_t ldz, bool wantz, const int32_t* select,
                           double* W, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(ordschur_kernel, dim3(1), dim3(64), 0, s, n, reinterpret_cast<zd*>(T), ldt,
                     reinterpret_cast<zd*>(Z), ldz, wantz ? 1 : 0, select, reinterpret_cast<zd*>(W));
  return hipGetLastError();
}

hipError_t launch_eigcond(int n, const double* T, int64_t ldt, int nsub, double* work, double* out, hipStream_t s) {
  hipLaunchKernelGGL(eigcond_kernel, dim3(1), dim3(64), 0, s, n, reinterpret_cast<const zd*>(T), ldt, nsub,
                     reinterpret_cast<zd*>(work), out);
  return hipGetLastError();
}

}  // namespace gs
