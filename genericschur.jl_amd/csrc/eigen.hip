// Kernels around the Schur factorization for `eigen!` (pirates.jl:63-160):
// balancing (balance.jl:33-199) and its back-transform of eigenvectors
// (balance.jl:201-261), the real-to-complex conversion of a standard-form
// Schur factorization (triangularize, triang.jl:9-43), the eigenvector
// normalisation _enormalize! (util.jl:572-592) and the reciprocal
// eigenvalue condition numbers of eigen!'s `jce` option (pirates.jl:112-120).
//
// None of these is on the QR hot path: each is O(n^2) per sweep of its
// own loop, one workgroup per matrix (balancing, triangularize) or one
// wavefront per column (back-transform, normalisation, conditions).  The
// arithmetic is written without contraction so that the oracle's
// restatement (compiled with -ffp-contract=off) gives the same bits where
// the reference's operation order is fixed.
#include "dev_common.h"
#include "internal.h"

#pragma clang fp contract(off)

namespace gs {

namespace {

template <int NT> GS_DEV double blk_max(double v, double* red) {
  v = wave_max(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) r = fmax(r, red[i]);
  return r;
}

template <int NT> GS_DEV double blk_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = red[0];
#pragma unroll
  for (int i = 1; i < NT / 64; ++i) r += red[i];
  return r;
}

GS_DEV double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Julia's max / min: a NaN operand gives NaN (fmax / fmin drop it)
GS_DEV double jmax(double a, double b) { return (isnan(a) || isnan(b)) ? a + b : fmax(a, b); }
GS_DEV double jmin(double a, double b) { return (isnan(a) || isnan(b)) ? a + b : fmin(a, b); }

// element helpers for real (double) and complex (zd) matrices
GS_DEV bool nonzero(double x) { return x != 0.0; }
GS_DEV bool nonzero(zd x) { return x.x != 0.0 || x.y != 0.0; }
GS_DEV double cmax(double x) { return fabs(x); }
GS_DEV double cmax(zd x) { return fmax(fabs(x.x), fabs(x.y)); }
GS_DEV double sqsum(double x, double s) {
  const double y = x / s;
  return y * y;
}
GS_DEV double sqsum(zd x, double s) {
  const double a = x.x / s, b = x.y / s;
  return a * a + b * b;
}
// the max-abs measure of _findamax (norm1est.jl:76-100): |x| for real
// entries, sqrt(max |x|^2) (taken after the maximum) for complex ones
GS_DEV double amax_term(double x) { return fabs(x); }
GS_DEV double amax_term(zd x) { return x.x * x.x + x.y * x.y; }
GS_DEV double amax_final(double m, double) { return m; }
GS_DEV double amax_final(double m, zd) { return sqrt(m); }
GS_DEV double scl(double x, double s) { return x * s; }
GS_DEV zd scl(zd x, double s) { return zmk(x.x * s, x.y * s); }

// 2-norm and _findamax measure of m entries x[0], x[inc], ... (block-wide)
template <int NT, class E> GS_DEV void norm_amax(const E* x, int64_t inc, int m, double* red, double& nrm,
                                                 double& am) {
  double mx = 0.0, a2 = 0.0;
  for (int i = threadIdx.x; i < m; i += NT) {
    const E v = x[i * inc];
    mx = fmax(mx, cmax(v));
    a2 = fmax(a2, amax_term(v));
  }
  mx = blk_max<NT>(mx, red);
  am = amax_final(blk_max<NT>(a2, red), E());
  if (mx == 0.0 || !(mx <= 1.79769313486231570815e308)) {
    nrm = mx;  // zero, inf or nan
    return;
  }
  double ss = 0.0;
  for (int i = threadIdx.x; i < m; i += NT) ss += sqsum(x[i * inc], mx);
  nrm = mx * sqrt(blk_sum<NT>(ss, red));
}

// balance!(A; scale, permute) of one matrix (balance.jl:33-199), in place
// in global memory.  out[0..3] = ilo, ihi, trivial, status (0, or -3 for a
// NaN met while scaling, -4 when the scaling sweeps did not settle within
// 10000 sweeps); sp[0..n-1] the permutation record (1-based row/column
// indices as in the reference), D the scale factors.
template <class E, int NT> __global__ void __launch_bounds__(NT) balance_kernel(E* A, int64_t lda, int n, int scale,
                                                                               int permute, int32_t* out,
                                                                               int32_t* sp, double* D) {
  __shared__ double red[NT / 64];
  const int tid = threadIdx.x;
  auto at = [&](int i, int j) -> E& { return A[(int64_t)(i - 1) + (int64_t)(j - 1) * lda]; };
  for (int i = tid; i < n; i += NT) {
    D[i] = 1.0;
    sp[i] = 0;
  }
  int ilo = 1, ihi = n, status = 0;
  bool trivial = true;
  // rows js, ms of columns 1..rlast, then rows js, ms of columns cfirst..n
  auto swap2 = [&](int js, int ms, int rlast, int cfirst) {
    for (int i = 1 + tid; i <= rlast; i += NT) {
      const E t = at(i, js);
      at(i, js) = at(i, ms);
      at(i, ms) = t;
    }
    __syncthreads();
    for (int i = cfirst + tid; i <= n; i += NT) {
      const E t = at(js, i);
      at(js, i) = at(ms, i);
      at(ms, i) = t;
    }
    __syncthreads();
  };
  if (permute) {
    ihi = n + 1;
    int js = 0;
    while (ihi > 1) {
      ihi -= 1;
      bool exch = false;
      int ms = 0;
      for (int j = ihi; j >= 1; --j) {
        bool nz = false;
        for (int i = 1 + tid; i <= ihi; i += NT)
          if (i != j && nonzero(at(j, i))) nz = true;
        nz = __syncthreads_or(nz);
        js = j;
        exch = !nz;
        if (exch) {
          ms = ihi;
          break;
        }
      }
      if (!exch) break;
      if (tid == 0) sp[ms - 1] = js;
      if (js != ms) {
        trivial = false;
        swap2(js, ms, ihi, ilo);
      }
    }
    if (ihi > 1) {
      ilo = 0;
      while (ilo < n) {
        ilo += 1;
        bool exch = false;
        int ms = 0;
        js = 0;
        for (int j = ilo; j <= ihi; ++j) {
          bool nz = false;
          for (int i = ilo + tid; i <= ihi; i += NT)
            if (i != j && nonzero(at(i, j))) nz = true;
          nz = __syncthreads_or(nz);
          js = j;
          exch = !nz;
          if (exch) {
            ms = ilo;
            break;
          }
        }
        if (!exch) break;
        if (tid == 0) sp[ms - 1] = js;
        if (ms != js) {
          trivial = false;
          swap2(js, ms, ihi, ilo);
        }
      }
    }
  }
  if (scale) {
    const double beta = 2.0, factor = 0.95;
    const double sfmin1 = GS_FMIN / GS_EPS;
    const double sfmin2 = sfmin1 * beta;
    const double sfmax2 = 1.0 / sfmin2;
    bool converged = false;
    for (int sweep = 0; !converged && status == 0; ++sweep) {
      if (sweep == 10000) {
        status = -4;
        break;
      }
      converged = true;
      for (int i = ilo; i <= ihi && status == 0; ++i) {
        const int m = ihi - ilo + 1;
        double c, r, ca, ra;
        norm_amax<NT>(&at(ilo, i), 1, m, red, c, ca);
        norm_amax<NT>(&at(i, ilo), lda, m, red, r, ra);
        // a NaN norm would make the reference's sweep loop run forever (no
        // comparison with it ever settles); it is the reference's
        // "NaN encountered while balancing" error instead
        if (isnan(c) || isnan(r)) {
          status = -3;
          break;
        }
        if (c == 0.0 || r == 0.0) continue;
        double g = r / beta;
        const double s = c + r;
        double f = 1.0;
        while (c < r / beta) {
          if (c >= g || jmax(f, jmax(c, ca)) >= sfmax2 || jmin(r, jmin(g, ra)) <= sfmin2) break;
          if (isnan(c + f + ca + r + g + ra)) {
            status = -3;
            break;
          }
          f *= beta;
          c *= beta;
          ca *= beta;
          r /= beta;
          g /= beta;
          ra /= beta;
        }
        if (status != 0) break;
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
        if (f != 1.0) trivial = false;
        if (c + r >= factor * s) continue;
        converged = false;
        if (tid == 0) D[i - 1] *= f;
        const double fi = 1.0 / f;
        for (int k = ilo + tid; k <= n; k += NT) at(i, k) = scl(at(i, k), fi);
        __syncthreads();
        for (int k = 1 + tid; k <= ihi; k += NT) at(k, i) = scl(at(k, i), f);
        __syncthreads();
      }
    }
  }
  if (tid == 0) {
    out[0] = ilo;
    out[1] = ihi;
    out[2] = trivial ? 1 : 0;
    out[3] = status;
  }
}

// lmul!(B, V) (inverse = 0: right eigenvectors) or ldiv!(B, V) (inverse = 1:
// left eigenvectors) on the m complex columns of V (balance.jl:201-261):
// one thread per column, rows scaled by D (or 1 / D) when ilo != ihi, then
// the recorded row exchanges in the reference's order.
__global__ void __launch_bounds__(256) balback_kernel(int n, int m, const int32_t* iit, const int32_t* sp,
                                                      const double* D, zd* V, int64_t ldv, int inverse) {
  const int j = blockIdx.x * 256 + threadIdx.x;
  if (j >= m) return;
  const int ilo = iit[0], ihi = iit[1];
  if (iit[2]) return;  // trivial
  zd* v = V + (int64_t)j * ldv;
  if (ilo != ihi)
    for (int i = 0; i < n; ++i) v[i] = scl(v[i], inverse ? 1.0 / D[i] : D[i]);
  for (int r = ilo - 1; r >= 1; --r) {
    const int q = sp[r - 1];
    if (q == r || q < 1 || q > n) continue;
    const zd t = v[r - 1];
    v[r - 1] = v[q - 1];
    v[q - 1] = t;
  }
  for (int r = ihi + 1; r <= n; ++r) {
    const int q = sp[r - 1];
    if (q == r || q < 1 || q > n) continue;
    const zd t = v[r - 1];
    v[r - 1] = v[q - 1];
    v[q - 1] = t;
  }
}

// triangularize (triang.jl:9-43): a standard-form real Schur factorization
// (Tr, Zr) to a complex triangular one (Tc, Zc).  Each 2 x 2 block is
// rotated by the complex Givens rotation G = [c, s; -conj(s), c] with
// c = sqrt|T[j-1,j]| / r, s = -i sqrt|T[j,j-1]| / r, r = hypot of the two:
// T <- G T G^H, Z <- Z G^H, blocks from the bottom up as in the reference
// (a block's rotation of rows and that of columns further right do not
// commute in rounding).  W receives diag(Tc).
template <int NT> __global__ void __launch_bounds__(NT) triang_kernel(int n, const double* Tr, int64_t ldt,
                                                                     const double* Zr, int64_t ldz, zd* Tc,
                                                                     int64_t ldtc, zd* Zc, int64_t ldzc, zd* W) {
  const int tid = threadIdx.x;
  for (int64_t e = tid; e < (int64_t)n * n; e += NT) {
    const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
    Tc[r + c * ldtc] = zr(Tr[r + c * ldt]);
    if (Zr) Zc[r + c * ldzc] = zr(Zr[r + c * ldz]);
  }
  __syncthreads();
  for (int j = n; j >= 2; --j) {
    const double sub = Tr[(j - 1) + (int64_t)(j - 2) * ldt];
    if (sub == 0.0) continue;
    const double s = sqrt(fabs(sub));
    const double c = sqrt(fabs(Tr[(j - 2) + (int64_t)(j - 1) * ldt]));
    const double r = hypot(s, c);
    const zd gc = zr(c / r);
    const zd gs = zmk(0.0, -(s / r));
    for (int k = 1 + tid; k <= n; k += NT) {
      zd* p1 = Tc + (j - 2) + (int64_t)(k - 1) * ldtc;
      zd* p2 = p1 + 1;
      const zd a1 = *p1, a2 = *p2;
      *p1 = gc * a1 + gs * a2;
      *p2 = -zconj(gs) * a1 + gc * a2;
    }
    __syncthreads();
    for (int k = 1 + tid; k <= n; k += NT) {
      zd* p1 = Tc + (k - 1) + (int64_t)(j - 2) * ldtc;
      zd* p2 = p1 + ldtc;
      const zd a1 = *p1, a2 = *p2;
      *p1 = a1 * zconj(gc) + a2 * zconj(gs);
      *p2 = -a1 * gs + a2 * zconj(gc);
      if (Zr) {
        zd* q1 = Zc + (k - 1) + (int64_t)(j - 2) * ldzc;
        zd* q2 = q1 + ldzc;
        const zd b1 = *q1, b2 = *q2;
        *q1 = b1 * zconj(gc) + b2 * zconj(gs);
        *q2 = -b1 * gs + b2 * zconj(gc);
      }
    }
    __syncthreads();
  }
  for (int64_t e = tid; e < (int64_t)n * n; e += NT) {
    const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
    if (r > c) Tc[r + c * ldtc] = zr(0.0);
  }
  __syncthreads();
  for (int i = tid; i < n; i += NT) W[i] = Tc[i + (int64_t)i * ldtc];
}

// _enormalize! (util.jl:572-592) on the m columns of V, one wavefront per
// column: scale to unit 2-norm with the first entry of largest modulus
// rotated to be real and positive.
__global__ void __launch_bounds__(256) enorm_kernel(int n, int m, zd* V, int64_t ldv) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= m) return;
  zd* v = V + (int64_t)j * ldv;
  double mx = 0.0, t = -1.0;
  int i0 = n;
  for (int i = lane; i < n; i += 64) {
    const zd x = v[i];
    mx = fmax(mx, fmax(fabs(x.x), fabs(x.y)));
    const double u = x.x * x.x + x.y * x.y;
    if (u > t) {
      t = u;
      i0 = i;
    }
  }
  mx = wave_max(mx);
  // the first index of the largest |v_i|^2 (ties to the smaller index)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double tu = __shfl_xor(t, o, 64);
    const int iu = __shfl_xor(i0, o, 64);
    if (tu > t || (tu == t && iu < i0)) {
      t = tu;
      i0 = iu;
    }
  }
  double ss = 0.0;
  if (mx > 0.0)
    for (int i = lane; i < n; i += 64) ss += sqsum(v[i], mx);
  const double nrm = mx * sqrt(wsum(ss));
  const double s = 1.0 / nrm;
  const zd vi0 = v[i0];
  const zd tt = (s * zconj(vi0)) / sqrt(t);
  for (int i = lane; i < n; i += 64) {
    const zd x = v[i] * tt;
    v[i] = (i == i0) ? zr(x.x) : x;
  }
}

// rconde[j] = |vl_j^H vr_j| / ||vl_j|| / ||vr_j|| (pirates.jl:112-120), one
// wavefront per column.
__global__ void __launch_bounds__(256) rcond_kernel(int n, int m, const zd* VL, int64_t ldl, const zd* VR,
                                                    int64_t ldr, double* rc) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= m) return;
  const zd* l = VL + (int64_t)j * ldl;
  const zd* r = VR + (int64_t)j * ldr;
  double ml = 0.0, mr = 0.0;
  for (int i = lane; i < n; i += 64) {
    ml = fmax(ml, fmax(fabs(l[i].x), fabs(l[i].y)));
    mr = fmax(mr, fmax(fabs(r[i].x), fabs(r[i].y)));
  }
  ml = wave_max(ml);
  mr = wave_max(mr);
  double dx = 0.0, dy = 0.0, sl = 0.0, sr = 0.0;
  for (int i = lane; i < n; i += 64) {
    // conj(l) r on the scaled entries
    const zd a = zmk(l[i].x / ml, l[i].y / ml), b = zmk(r[i].x / mr, r[i].y / mr);
    dx += a.x * b.x + a.y * b.y;
    dy += a.x * b.y - a.y * b.x;
    sl += a.x * a.x + a.y * a.y;
    sr += b.x * b.x + b.y * b.y;
  }
  dx = wsum(dx);
  dy = wsum(dy);
  sl = wsum(sl);
  sr = wsum(sr);
  if (lane == 0) rc[j] = hypot(dx, dy) / sqrt(sl) / sqrt(sr);
}

}  // namespace

hipError_t launch_balance(bool cplx, double* A, int64_t lda, int n, int scale, int permute, int32_t* out, int32_t* sp,
                          double* D, hipStream_t s) {
  if (cplx)
    hipLaunchKernelGGL((balance_kernel<zd, 256>), dim3(1), dim3(256), 0, s, reinterpret_cast<zd*>(A), lda, n, scale,
                       permute, out, sp, D);
  else
    hipLaunchKernelGGL((balance_kernel<double, 256>), dim3(1), dim3(256), 0, s, A, lda, n, scale, permute, out, sp,
                       D);
  return hipGetLastError();
}

hipError_t launch_balback(int n, int m, const int32_t* iit, const int32_t* sp, const double* D, double* V,
                          int64_t ldv, int inverse, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(balback_kernel, dim3((m + 255) / 256), dim3(256), 0, s, n, m, iit, sp, D,
                     reinterpret_cast<zd*>(V), ldv, inverse);
  return hipGetLastError();
}

hipError_t launch_triangularize(int n, const double* T, int64_t ldt, const double* Z, int64_t ldz, double* Tc,
                                int64_t ldtc, double* Zc, int64_t ldzc, double* W, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(triang_kernel<256>, dim3(1), dim3(256), 0, s, n, T, ldt, Z, ldz, reinterpret_cast<zd*>(Tc),
                     ldtc, reinterpret_cast<zd*>(Zc), ldzc, reinterpret_cast<zd*>(W));
  return hipGetLastError();
}

hipError_t launch_enormalize(int n, int m, double* V, int64_t ldv, hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(enorm_kernel, dim3((m + 3) / 4), dim3(256), 0, s, n, m, reinterpret_cast<zd*>(V), ldv);
  return hipGetLastError();
}

hipError_t launch_eig_rcond(int n, int m, const double* VL, int64_t ldl, const double* VR, int64_t ldr, double* rc,
                            hipStream_t s) {
  if (m <= 0 || n <= 0) return hipSuccess;
  hipLaunchKernelGGL(rcond_kernel, dim3((m + 3) / 4), dim3(256), 0, s, n, m, reinterpret_cast<const zd*>(VL), ldl,
                     reinterpret_cast<const zd*>(VR), ldr, rc);
  return hipGetLastError();
}

}  // namespace gs
