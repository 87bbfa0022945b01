// Complex Schur decomposition on gfx950: Householder Hessenberg reduction,
// formation of Z, and single-shift Francis QR with deflation, following
// GenericSchur.jl's complex path:
//   driver        GenericSchur.jl:350-372   (scale, Hessenberg, Z, QR, unscale)
//   QR driver     GenericSchur.jl:194-335   (zlahqr deflation and shifts)
//   sweep         GenericSchur.jl:374-504   (singleShiftQR!)
//   Hessenberg    hessenberg.jl:3-17, householder.jl:56-172
//   Q formation   hessenberg.jl:150-166
//   scaling       util.jl:14-80
//
// Batched matrices take two or three launches per size bin: zhq_kernel
// reduces A to Hessenberg form and forms Z with the whole matrix resident
// in LDS (one workgroup per matrix); zqrw2_kernel runs the QR iterations
// with H and Z in HBM, two wavefronts per matrix (the bulge chase in a
// 32 x 32 LDS window, and the chunk's reflectors applied to the rest of H
// and Z as chains); for bins above n = 80 it logs the Z updates instead and
// zlog_kernel applies the log with Z in LDS.  zqr_kernel (GS_QR_BAND) keeps
// H's band j >= i-2 in LDS instead.  The single-matrix kernel keeps H in
// HBM.
//
// Every thread runs the scalar control logic (deflation tests, shifts,
// reflectors) redundantly on values read from H; searches are parallel over
// threads followed by a butterfly maximum, so all threads take the same
// branches.
#include <algorithm>
#include <cstdlib>

#include "dev_common.h"
#include "formq.h"
#include "internal.h"
#include "layout.h"

namespace gs {

#define HH(i, j) H[lay.idx((i), (j))]
#define ZZ(i, j) Zg[((i) - 1) + (int64_t)((j) - 1) * ldz]

// Reflector on x = H[p+1 : n-1, p] (householder.jl:56-102) by one
// wavefront (its reductions need no workgroup barrier): tau[p], beta in
// H[p+1, p], the scaled tail below.
GS_DEV void z_hess_reflect(zd* H, int64_t ld, int n, int p, zd* tau) {
  const int lane = threadIdx.x & 63;
  const double sfmin = GS_FMIN / GS_EPS;
  const int r0 = p + 2;
  double amax = 0.0;
  for (int r = r0 + lane; r < n; r += 64) {
    const zd v = H[r + p * ld];
    amax = fmax(amax, fmax(fabs(v.x), fabs(v.y)));
  }
  amax = wave_max(amax);
  double xnorm = 0.0;
  if (amax > 0.0) {
    const double ra = 1.0 / amax;
    double s = 0.0;
    for (int r = r0 + lane; r < n; r += 64) {
      const zd v = H[r + p * ld];
      const double a = v.x * ra, b = v.y * ra;
      s += a * a + b * b;
    }
    s = wave_sum(s);
    xnorm = amax * sqrt(s);
  }
  zd alpha = H[(p + 1) + p * ld];
  zd t_ = zr(0.0);
  if (!(xnorm == 0.0 && alpha.y == 0.0)) {
    double ar = alpha.x, ai = alpha.y;
    double beta = -copysign(hypot3(ar, ai, xnorm), ar);
    int kount = 0;
    if (fabs(beta) < sfmin) {
      const double rsfmin = 1.0 / sfmin;
      bool smallb = true;
      while (smallb) {
        kount += 1;
        wsync();
        for (int r = r0 + lane; r < n; r += 64) H[r + p * ld] = H[r + p * ld] * rsfmin;
        beta *= rsfmin;
        ar *= rsfmin;
        ai *= rsfmin;
        smallb = (fabs(beta) < sfmin) && (kount < 20);
      }
      wsync();
      double am2 = 0.0;
      for (int r = r0 + lane; r < n; r += 64) {
        const zd v = H[r + p * ld];
        am2 = fmax(am2, fmax(fabs(v.x), fabs(v.y)));
      }
      am2 = wave_max(am2);
      xnorm = 0.0;
      if (am2 > 0.0) {
        const double ra = 1.0 / am2;
        double s = 0.0;
        for (int r = r0 + lane; r < n; r += 64) {
          const zd v = H[r + p * ld];
          const double a = v.x * ra, b = v.y * ra;
          s += a * a + b * b;
        }
        s = wave_sum(s);
        xnorm = am2 * sqrt(s);
      }
      alpha = zmk(ar, ai);
      beta = -copysign(hypot3(ar, ai, xnorm), ar);
    }
    t_ = zmk((beta - ar) / beta, -ai / beta);
    const zd sc = zdiv(zr(1.0), alpha - zr(beta));
    for (int r = r0 + lane; r < n; r += 64) H[r + p * ld] = H[r + p * ld] * sc;
    for (int j = 0; j < kount; ++j) beta *= sfmin;
    if (lane == 0) H[(p + 1) + p * ld] = zr(beta);
  }
  if (lane == 0) tau[p] = t_;
}

// ------------------------------------------------------------------------
// LDS bytes of the batched Hessenberg kernel for one matrix of order n:
// H (leading dimension n | 1), tau, two work vectors, and, when they fit in
// the 160 KiB of a CU, four partial-sum vectors for the fused left update
// and right dot products (zhq_colwise); then the reduction slots.
__host__ __device__ inline size_t zhq_need_bytes(int n, int nt, bool cw) {
  const size_t ld = (size_t)(n | 1), m = (size_t)(n > 0 ? n : 1);
  return 16 * (ld * (size_t)n + 3 * m + (cw ? 4 * m : 0)) + 8 * (size_t)(nt / 64 + 1);
}
// decided for the largest workgroup (1024 threads), so that every
// workgroup size takes the same path for a given order
__host__ __device__ inline bool zhq_colwise(int n) { return zhq_need_bytes(n, 1024, true) <= 160 * 1024; }

// ------------------------------------------------------------------------
// Hessenberg reduction in place (hessenberg.jl:3-17), full layout, 0-based.
// Reflector p is stored below the subdiagonal of column p, tau[p] its scalar.
// RM > 0: n <= 64 RM (LDS-resident matrices), so each lane keeps the
// reflector and row factors of its RM rows in registers in the rank-1
// updates; RM = 0: any n.
template <int NT, int RM = 0> GS_DEV void z_hessenberg(zd* H, int64_t ld, int n, zd* tau, double* red, zd* wk) {
  (void)red;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  // RM > 0: wavefront 0 computes reflector p + 1 (column p + 1, final once
  // the right update's dot products are done) while the other wavefronts
  // apply reflector p's right update to columns p + 2 .. n - 1
  constexpr bool OVL = RM > 0 && NT >= 128;
  // reflector p by wavefront 0 alone
  auto reflect = [&](const int p) { z_hess_reflect(H, ld, n, p, tau); };
  if (OVL) {
    if (wave == 0 && n >= 2) reflect(0);
  }
  for (int p = 0; p <= n - 2; ++p) {
    const int r0 = p + 2;  // first tail row
    if (!OVL && wave == 0) reflect(p);
    __syncthreads();
    const zd t_ = tau[p];
    const zd ct = zconj(t_);
    // lmul!(H', A[p+1:n, p+1:n])   (householder.jl:157-172):
    // va_j = ct (H[p+1, j] + sum_r conj(v_r) H[r, j]); column j -= va_j [1; v]
    // (the dot products leave row p+1 alone, so the lane holding va_j's sum
    // finishes it at once)
    zdots4e<NT>(p + 1, n, r0, n, [&](int r, int j) { return zconj(H[r + p * ld]) * H[r + (int64_t)j * ld]; },
                [&](int j, zd acc) { wk[j] = ct * (H[(p + 1) + (int64_t)j * ld] + acc); });
    __syncthreads();
    const zd mt = -t_;
    if (RM > 0 && zhq_colwise(n)) {
      // The left update fused with the right update's dot products, by
      // units of (column class g, row block h): columns c = r0 + g + 4 k
      // (g < 4), rows r = lane + 64 h (h < 2 above n = 64, else h = 0),
      // with v_r in a register and each column's va_c and v_c uniform
      // loads.  Per element: one read and one write of H.  Unit (g, h)
      // sums its rows over its columns in increasing c into
      // wk[2 n + g n + r]; the four classes are then added in class order,
      // so the result does not depend on the workgroup size.
      constexpr int NWA = NT / 64;
      zd* xw = wk + n;
      zd* xp = wk + 2 * n;
      const int nunits = n > 64 ? 8 : 4;
      for (int u = wave; u < nunits; u += NWA) {
        const int g = u & 3, r = lane + 64 * (u >> 2);
        const bool upd = r < n && r >= p + 1;
        const zd vr = (r == p + 1) ? zr(1.0) : (upd ? H[r + p * ld] : zr(0.0));
        zd xa = zr(0.0);
        if (r < n) {
#pragma unroll 4
          for (int c = r0 + g; c < n; c += 4) {
            const zd va = wk[c];
            const zd vc = H[c + p * ld];
            zd h = H[r + (int64_t)c * ld];
            if (upd) {
              h = h - va * vr;
              H[r + (int64_t)c * ld] = h;
            }
            xa = xa + h * vc;
          }
          xp[g * n + r] = xa;
        }
      }
      __syncthreads();
      for (int r = tid; r < n; r += NT) {
        const zd acc = ((xp[r] + xp[n + r]) + xp[2 * n + r]) + xp[3 * n + r];
        zd h1 = H[r + (int64_t)(p + 1) * ld];
        if (r >= p + 1) h1 = h1 - wk[p + 1] * ((r == p + 1) ? zr(1.0) : H[r + p * ld]);
        const zd x = h1 + acc;
        H[r + (int64_t)(p + 1) * ld] = h1 - t_ * x;
        xw[r] = x * mt;
      }
    } else if constexpr (RM > 0) {
      // The left update fused with the right update's dot products: the
      // thread that sums row r over columns c >= p + 2 applies
      // H[r, c] -= va_c v_r to each element as it reads it (rows r >= p + 1),
      // so the trailing block is read once for both; the lane that finishes
      // x_r applies the left update to H[r, p + 1] itself.  x_r goes to
      // wk[n + r], since wk[0 .. n) still holds the va_c.
      zd* xw = wk + n;
      const int s4 = tid & 3, il = tid >> 2;
      const int seg = (n - r0 + 3) >> 2;
      const int ks = r0 + s4 * seg, ke = min(ks + seg, n);
      for (int ib = 0; ib < n; ib += NT / 4) {
        const int r = ib + il;
        zd acc = zr(0.0);
        if (r < n) {
          if (r >= p + 1) {
            const zd vr = (r == p + 1) ? zr(1.0) : H[r + p * ld];
#pragma unroll 4
            for (int c = ks; c < ke; ++c) {
              const zd h = H[r + (int64_t)c * ld] - wk[c] * vr;
              H[r + (int64_t)c * ld] = h;
              acc = acc + h * H[c + p * ld];
            }
          } else {
#pragma unroll 4
            for (int c = ks; c < ke; ++c) acc = acc + H[r + (int64_t)c * ld] * H[c + p * ld];
          }
        }
        acc.x += __shfl_xor(acc.x, 1, 64);
        acc.y += __shfl_xor(acc.y, 1, 64);
        acc.x += __shfl_xor(acc.x, 2, 64);
        acc.y += __shfl_xor(acc.y, 2, 64);
        if (s4 == 0 && r < n) {
          zd h1 = H[r + (int64_t)(p + 1) * ld];
          if (r >= p + 1) h1 = h1 - wk[p + 1] * ((r == p + 1) ? zr(1.0) : H[r + p * ld]);
          const zd x = h1 + acc;
          H[r + (int64_t)(p + 1) * ld] = h1 - t_ * x;
          xw[r] = x * mt;
        }
      }
    } else {
#pragma unroll 4
      for (int j = p + 1 + wave; j < n; j += NT / 64) {
        const zd va = wk[j];
        for (int r = p + 1 + lane; r < n; r += 64) {
          const zd h = H[r + (int64_t)j * ld];
          H[r + (int64_t)j * ld] = (r == p + 1) ? h - va : h - va * H[r + p * ld];
        }
      }
      __syncthreads();
      // rmul!(A[:, p+1:n], H, x)     (householder.jl:140-155, util.jl:103-114):
      // x_r = H[r, p+1] + sum_c H[r, c] v_c; H[r, p+1] -= tau x_r;
      // H[r, c] += (-tau x_r) conj(v_c) (the dot products leave column p+1
      // alone, so the lane holding x_r's sum finishes it at once)
      zdots4e<NT>(0, n, r0, n, [&](int c, int r) { return H[r + (int64_t)c * ld] * H[c + p * ld]; },
                  [&](int r, zd acc) {
                    const zd x = H[r + (int64_t)(p + 1) * ld] + acc;
                    H[r + (int64_t)(p + 1) * ld] = H[r + (int64_t)(p + 1) * ld] - t_ * x;
                    wk[r] = x * mt;
                  });
    }
    __syncthreads();
    if constexpr (RM > 0) {
      if (OVL && wave == 0) {
        if (p + 1 <= n - 2) reflect(p + 1);
      } else {
      zd xr[RM];
#pragma unroll
      for (int q = 0; q < RM; ++q) xr[q] = (lane + 64 * q < n) ? wk[n + lane + 64 * q] : zr(0.0);
      constexpr int NW = OVL ? NT / 64 - 1 : NT / 64;
#pragma unroll 4
      for (int c = r0 + wave - (OVL ? 1 : 0); c < n; c += NW) {
        const zd vc = zconj(H[c + p * ld]);
#pragma unroll
        for (int q = 0; q < RM; ++q) {
          const int r = lane + 64 * q;
          if (r < n) H[r + (int64_t)c * ld] = H[r + (int64_t)c * ld] + xr[q] * vc;
        }
      }
      }
    } else {
#pragma unroll 4
      for (int c = r0 + wave; c < n; c += NT / 64) {
        const zd vc = zconj(H[c + p * ld]);
        for (int r = lane; r < n; r += 64) H[r + (int64_t)c * ld] = H[r + (int64_t)c * ld] + wk[r] * vc;
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------
// Load (full layout, leading dimension ld), optional subdiagonal check and
// _scale! (util.jl:14-29; norm(A, Inf) is the largest entry modulus).
// Returns -2 for a non-real subdiagonal (checksd), else 0; scl receives
// (cscale, anrm) when A was scaled, (0, 0) otherwise.
template <int NT> GS_DEV int z_prep(const zd* Ag, int64_t lda, zd* H, int64_t ld, int n, int flags, double* red,
                                    double& cscale, double& anrm) {
  const int tid = threadIdx.x;
  const bool scale = (flags & 0x02) != 0;
  const bool hessin = (flags & 0x04) != 0;
  const bool checksd = (flags & 0x08) != 0;
  if ((const void*)Ag != (const void*)H) {
    for (int64_t e = tid; e < (int64_t)n * n; e += NT) {
      const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
      H[r + c * ld] = Ag[r + c * lda];
    }
  }
  __syncthreads();
  cscale = 0.0;
  anrm = 0.0;
  if (hessin && checksd) {
    bool bad = false;
    for (int j = 1; j <= n - 1; ++j) bad |= (H[j + (int64_t)(j - 1) * ld].y != 0.0);
    if (bad) return -2;
  }
  if (scale) {
    double m = 0.0;
    for (int64_t e = tid; e < (int64_t)n * n; e += NT) {
      const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
      m = fmax(m, zabs(H[r + c * ld]));
    }
    const double an = block_max<NT>(m, red);
    const double smlnum = sqrt(GS_FMIN) / GS_EPS;
    const double bignum = 1.0 / smlnum;
    double cs = 0.0;
    if (an > 0.0 && an < smlnum) cs = smlnum;
    else if (an > bignum) cs = bignum;
    if (cs != 0.0) {
      cscale = cs;
      anrm = an;
      double mul[4];
      const int nm = safescale_muls(an, cs, mul);
      for (int64_t e = tid; e < (int64_t)n * n; e += NT) {
        const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
        zd v = H[r + c * ld];
        for (int q = 0; q < nm; ++q) v = v * mul[q];
        H[r + c * ld] = v;
      }
      __syncthreads();
    }
  }
  return 0;
}

// ------------------------------------------------------------------------
// Deferred application of one sweep's rotations to Z (columns k0..k1+1), as
// one chain per row; `ph` is the phase of the shifted-start scaling (applied
// to columns k0 and k0+2..k1+1), `tl` the tail phase of column k1+1.  The row
// segment is loaded CH entries at a time, one chunk ahead of the rotations.
template <int NT> GS_DEV void z_apply_chain(zd* Zg, int64_t ldz, int n, const zd* rot, int k0, int k1, bool has_ph,
                                            zd ph, bool has_tl, zd tl) {
  constexpr int CH = 8;
  const zd cph = zconj(ph);
  for (int r = 1 + threadIdx.x; r <= n; r += NT) {
    zd carry = ZZ(r, k0);
    zd nxt[CH];
#pragma unroll
    for (int q = 0; q < CH; ++q)
      if (k0 + q <= k1) nxt[q] = ZZ(r, k0 + q + 1);
    for (int kb = k0; kb <= k1; kb += CH) {
      zd cur[CH];
#pragma unroll
      for (int q = 0; q < CH; ++q) cur[q] = nxt[q];
#pragma unroll
      for (int q = 0; q < CH; ++q)
        if (kb + CH + q <= k1) nxt[q] = ZZ(r, kb + CH + q + 1);
#pragma unroll
      for (int q = 0; q < CH; ++q) {
        const int k = kb + q;
        if (k <= k1) {
          const zd tau1 = rot[2 * (k - k0)];
          const zd v2 = rot[2 * (k - k0) + 1];
          const double tau2 = (tau1 * v2).x;
          zd b = cur[q];
          if (has_ph && k + 1 >= k0 + 2) b = b * cph;
          const zd ss = tau1 * carry + tau2 * b;
          zd znew = carry - ss;
          if (has_ph && k == k0) znew = znew * cph;
          ZZ(r, k) = znew;
          carry = b - ss * zconj(v2);
        }
      }
    }
    if (has_tl) carry = carry * tl;
    ZZ(r, k1 + 1) = carry;
  }
}

// ------------------------------------------------------------------------
// singleShiftQR! (GenericSchur.jl:374-504) on H with deferred Z.
template <int NT, class L, bool FAST> GS_DEV void z_sweep(zd* H, L lay, int n, zd* Zg, int64_t ldz, bool wantz, zd* rot,
                                               int* redi, zd shift, int istart, int iend, int64_t& nstep) {
  const int tid = threadIdx.x;
  const double ulp = GS_EPS;
  // start row: the lowest mm in (istart, iend) whose start would make
  // H[mm, mm-1] negligible (two consecutive small subdiagonals)
  int found = -1;
  for (int mm = iend - 1 - tid; mm >= istart + 1; mm -= NT) {
    const zd h11 = HH(mm, mm);
    const zd h22 = HH(mm + 1, mm + 1);
    zd h11s = h11 - shift;
    double h21 = HH(mm + 1, mm).x;
    const double s = zabs1(h11s) + fabs(h21);
    h11s = h11s / s;
    h21 = h21 / s;
    const double h10 = HH(mm, mm - 1).x;
    if (fabs(h10) * fabs(h21) <= ulp * (zabs1(h11s) * (zabs1(h11) + zabs1(h22)))) {
      found = mm;
      break;
    }
  }
  found = block_max_int<NT>(found, redi);
  const int istart1 = (found > 0) ? found : istart;
  zd v0, v1;
  {
    const zd h11 = HH(istart1, istart1);
    zd h11s = h11 - shift;
    double h21 = HH(istart1 + 1, istart1).x;
    const double s = zabs1(h11s) + fabs(h21);
    h11s = h11s / s;
    h21 = h21 / s;
    v0 = h11s;
    v1 = zr(h21);
  }
  bool has_ph = false;
  zd ph = zr(1.0);
  // FAST: the next step's reflector input (H[k+1,k], H[k+2,k]) is taken
  // from the registers of the lanes that computed it in the right update
  // (single-wavefront blocks), the reflector uses fewer divisions, and the
  // block barriers are wavefront ordering points
  constexpr bool RL = FAST && NT == 64;
  bool reload = true;
  zd nv0 = zr(0.0), nv1 = zr(0.0);
  for (int k = istart1; k <= iend - 1; ++k) {
    if (k > istart1) {
      if (RL && !reload) {
        v0 = nv0;
        v1 = nv1;
      } else {
        v0 = HH(k, k - 1);
        v1 = HH(k + 1, k - 1);
      }
    }
    const zd tau1 = FAST ? zreflector2_fast(v0, v1) : zreflector2(v0, v1);
    if (FAST) bsync_full<NT>();
    else bsync<NT>();
    if (k > istart1 && tid == 0) {
      HH(k, k - 1) = v0;
      HH(k + 1, k - 1) = zr(0.0);
    }
    const zd v2 = v1;
    const double tau2 = (tau1 * v2).x;
    const zd ct1 = zconj(tau1);
    // left: rows k, k+1 over columns k..n
    for (int j = k + tid; j <= n; j += NT) {
      const zd a = HH(k, j), b = HH(k + 1, j);
      const zd ss = ct1 * a + tau2 * b;
      HH(k, j) = a - ss;
      HH(k + 1, j) = b - ss * v2;
    }
    if (FAST) bsync_full<NT>();
    else __syncthreads();
    // right: columns k, k+1 over rows 1..min(k+2, iend)
    const int jm = min(k + 2, iend);
    const zd cv2 = zconj(v2);
    zd cap0 = zr(0.0), cap1 = zr(0.0);
    for (int j = 1 + tid; j <= jm; j += NT) {
      const zd a = HH(j, k), b = HH(j, k + 1);
      const zd ss = tau1 * a + tau2 * b;
      const zd na = a - ss;
      HH(j, k) = na;
      HH(j, k + 1) = b - ss * cv2;
      if (RL) {
        if (j == k + 1) cap0 = na;
        if (j == k + 2) cap1 = na;
      }
    }
    if (wantz && tid == 0) {
      rot[2 * (k - istart1)] = tau1;
      rot[2 * (k - istart1) + 1] = v2;
    }
    if (RL) {
      const int l0 = k & 63, l1 = (k + 1) & 63;
      nv0 = zmk(rdlane(cap0.x, l0), rdlane(cap0.y, l0));
      nv1 = zmk(rdlane(cap1.x, l1), rdlane(cap1.y, l1));
      reload = false;
    }
    if (FAST) bsync_full<NT>();
    else __syncthreads();
    if (k == istart1 && istart1 > istart) {
      reload = true;  // H[istart1+1, istart1] is rescaled below
      zd t = zr(1.0) - tau1;
      t = t / zabs(t);
      has_ph = true;
      ph = t;
      const zd ctt = zconj(t);
      if (tid == 0) {
        HH(istart1 + 1, istart1) = HH(istart1 + 1, istart1) * ctt;
        if (istart1 + 2 <= iend) HH(istart1 + 2, istart1 + 1) = HH(istart1 + 2, istart1 + 1) * t;
      }
      __syncthreads();
      // rows j in S right of the diagonal times t, then columns j in S
      // above the diagonal times conj(t); S = istart1..iend minus istart1+1
      for (int c = 1 + tid; c <= n; c += NT) {
        const bool cin = (c >= istart1 && c <= iend && c != istart1 + 1);
        const int rmax = min(c - 1, iend);
        for (int r = 1; r <= rmax; ++r) {
          const bool rin = (r >= istart1 && r != istart1 + 1);
          if (!rin && !cin) continue;
          zd x = HH(r, c);
          if (rin) x = x * t;
          if (cin) x = x * ctt;
          HH(r, c) = x;
        }
      }
      __syncthreads();
    }
  }
  nstep += iend - istart1;
  // ensure reality of the tail subdiagonal
  zd t = HH(iend, iend - 1);
  bool has_tl = false;
  if (t.y != 0.0) {
    const double rtv = zabs(t);
    t = t / rtv;
    has_tl = true;
    bsync<NT>();
    if (tid == 0) HH(iend, iend - 1) = zr(rtv);
    const zd ctt = zconj(t);
    for (int c = iend + 1 + tid; c <= n; c += NT) HH(iend, c) = HH(iend, c) * ctt;
    for (int r = 1 + tid; r <= iend - 1; r += NT) HH(r, iend) = HH(r, iend) * t;
  }
  if (wantz) z_apply_chain<NT>(Zg, ldz, n, rot, istart1, iend - 1, has_ph, ph, has_tl, t);
  __syncthreads();
}

// ------------------------------------------------------------------------
// Windowed sweep for the batched QR kernel (one wavefront per matrix, H and
// Z in HBM).  The bulge chase of a sweep runs in chunks of up to ZKC steps;
// a chunk's window of H (rows and columns k0-1 .. kend+2) is chased in LDS
// and the chunk's reflectors, logged in LDS, are then applied to the rest of
// H (columns right of the window, rows above it) and to Z as one chain per
// row or column.  Each element receives the reflectors in the same order
// and with the same expressions as in z_sweep, and the phase and tail
// scalings of singleShiftQR! (GenericSchur.jl:461-500) are applied to all
// of H and Z at the same points of the sweep, so the result equals
// z_sweep's.  The LDS footprint is the window only, so many matrices share
// a CU and hide each other's latency.
constexpr int ZW = 32;       // window rows and columns
constexpr int ZKC = ZW - 3;  // bulge steps per chunk
// LDS pitch of a window column: odd, so that the lanes of a row access
// (left update, lane = column) fall into different banks
constexpr int ZWL = ZW + 1;

// Logged reflectors of steps ka..kb (lg[2(k-ka)] = tau1, lg[2(k-ka)+1] = v2,
// kb - ka < ZKC) applied to the line x(t) = p[(t - 1) st]: LEFT updates rows
// k, k+1 of a column of H (GenericSchur.jl:437-446), otherwise columns k, k+1
// of a row of H or Z (:447-460).  The segment x(ka) .. x(kb+1) moves through
// registers in groups of ZCG elements, two groups in flight: a group's loads
// are issued as soon as the group two places before it has been stored, so a
// chain of a full chunk waits on memory about twice, not once per element.
// tail2 (optional): receives the last two elements x(kb), x(kb+1) in place
// of the line.
constexpr int ZCG = 8;
template <bool LEFT> GS_DEV void zchain(zd* p, int64_t st, const zd* lg, int ka, int kb, zd* tail2 = nullptr) {
  constexpr int NG = (ZKC + 1 + ZCG - 1) / ZCG;
  const int s = __builtin_amdgcn_readfirstlane(kb - ka + 1);  // uniform: branches on SCC
  ka = __builtin_amdgcn_readfirstlane(ka);
  // log offset through an opaque (volatile asm) SGPR: the log is re-read per
  // line (LDS broadcasts) rather than held in registers across the callers'
  // loops over lines
  int lb = 0;
  asm volatile("" : "+s"(lb));
  const zd* lgq = lg + lb;
  zd* q0 = p + (int64_t)(ka - 1) * st;
  zd x[NG * ZCG];
#pragma unroll
  for (int g = 0; g < 2 && g < NG; ++g)
    if (g * ZCG <= s) {
#pragma unroll
      for (int i = 0; i < ZCG; ++i) x[g * ZCG + i] = q0[(int64_t)min(g * ZCG + i, s) * st];
    }
#pragma unroll
  for (int g = 0; g < NG; ++g) {
    if (g * ZCG > s) break;
#pragma unroll
    for (int i = 0; i < ZCG; ++i) {
      const int q = g * ZCG + i;
      if (q < s) {
        const zd tau1 = lgq[2 * q], v2 = lgq[2 * q + 1];
        const double tau2 = (tau1 * v2).x;
        const zd a = x[q], b = x[q + 1];
        if (LEFT) {
          const zd ss = zconj(tau1) * a + tau2 * b;
          x[q] = a - ss;
          x[q + 1] = b - ss * v2;
        } else {
          zd x0 = a, x1 = b;
          zrot_right_fixed(x0, x1, tau1, v2);
          x[q] = x0;
          x[q + 1] = x1;
        }
      }
    }
#pragma unroll
    for (int i = 0; i < ZCG; ++i) {
      const int q = g * ZCG + i;
      if (tail2 && q >= s - 1 && q <= s)
        tail2[q - (s - 1)] = x[q];
      else if (q <= s)
        q0[(int64_t)q * st] = x[q];
    }
    __builtin_amdgcn_sched_barrier(0);
    if (g + 2 < NG && (g + 2) * ZCG <= s) {
#pragma unroll
      for (int i = 0; i < ZCG; ++i) x[(g + 2) * ZCG + i] = q0[(int64_t)min((g + 2) * ZCG + i, s) * st];
    }
  }
}

// zchain<true> for a full chunk (kb = ka + ZKC - 1) on a column of H whose
// rows kb, kb + 1 go to tail2 (the next window) instead of HBM.  Every count
// and index is a compile-time constant, so the compiler waits on memory once
// per group rather than before every store.
template <bool TAIL = true> GS_DEV void zchain_full(zd* p, const zd* lg, int ka, zd* tail2) {
  constexpr int S = ZKC, NG = (S + 1 + ZCG - 1) / ZCG;
  ka = __builtin_amdgcn_readfirstlane(ka);
  int lb = 0;
  asm volatile("" : "+s"(lb));
  const zd* lgq = lg + lb;
  zd* q0 = p + (ka - 1);
  zd x[NG * ZCG];
#pragma unroll
  for (int e = 0; e < 2 * ZCG; ++e)
    if (e <= S) x[e] = q0[e];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
#pragma unroll
    for (int i = 0; i < ZCG; ++i) {
      const int q = g * ZCG + i;
      if (q < S) {
        const zd tau1 = lgq[2 * q], v2 = lgq[2 * q + 1];
        const double tau2 = (tau1 * v2).x;
        const zd a = x[q], b = x[q + 1];
        const zd ss = zconj(tau1) * a + tau2 * b;
        x[q] = a - ss;
        x[q + 1] = b - ss * v2;
      }
    }
#pragma unroll
    for (int i = 0; i < ZCG; ++i) {
      const int q = g * ZCG + i;
      if (q < S - 1 || (!TAIL && q <= S))
        q0[q] = x[q];
      else if (q <= S)
        tail2[q - (S - 1)] = x[q];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < ZCG; ++i) {
      const int e = (g + 2) * ZCG + i;
      if (e <= S) x[e] = q0[e];
    }
  }
}

// zchain<false> (a row of H or Z, x(t) = p[(t - 1) st]) for a full chunk:
// fixed shape, as zchain_full.
template <int G> GS_DEV void zchain_full_row(zd* p, int64_t st, const zd* lg, int ka) {
  constexpr int ZCG = G, S = ZKC, NG = (S + 1 + ZCG - 1) / ZCG;
  ka = __builtin_amdgcn_readfirstlane(ka);
  int lb = 0;
  asm volatile("" : "+s"(lb));
  const zd* lgq = lg + lb;
  zd* q0 = p + (int64_t)(ka - 1) * st;
  zd x[NG * ZCG];
#pragma unroll
  for (int e = 0; e < 2 * ZCG; ++e)
    if (e <= S) x[e] = q0[(int64_t)e * st];
#pragma unroll
  for (int g = 0; g < NG; ++g) {
#pragma unroll
    for (int i = 0; i < ZCG; ++i) {
      const int q = g * ZCG + i;
      if (q < S) zrot_right_fixed(x[q], x[q + 1], lgq[2 * q], lgq[2 * q + 1]);
    }
#pragma unroll
    for (int i = 0; i < ZCG; ++i) {
      const int q = g * ZCG + i;
      if (q <= S) q0[(int64_t)q * st] = x[q];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < ZCG; ++i) {
      const int e = (g + 2) * ZCG + i;
      if (e <= S) x[e] = q0[(int64_t)e * st];
    }
  }
}

// Z log of the windowed QR kernel.  Z is not read while the sweeps run, so
// instead of updating Z in HBM after every chunk the kernel records each
// sweep and zlog_kernel applies the records to Z, held in LDS, afterwards:
// the same operations on every element of Z in the same order.  Record of a
// sweep, in complex entries: (istart1, iend); ctt of the phase scaling after
// the first step, or 0; (tau1, v2) per step; t of the tail scaling, or 0.
// When a sweep's record would not fit, the kernel first applies the records
// so far to Z itself (zlog_row on Z in HBM) and starts the log again.
struct ZLog {
  zd* g;        // null: Z updated in place
  int64_t cap;  // capacity in complex entries (at least one sweep's record)
  int64_t pos;  // entries in use
};

// Applies log records [0, len) to one row of Z, x(c) = row[(c - 1) st]: per
// sweep the steps as one chain (GenericSchur.jl:447-460 on Z), the phase
// scaling after the first step (:461-482) and the tail scaling (:486-500).
GS_DEV void zlog_row(zd* row, int64_t st, const zd* __restrict__ lg, int64_t len) {
  int64_t pos = 0;
  while (pos < len) {
    const zd h0 = lg[pos];
    const int i1 = (int)h0.x, ie = (int)h0.y;
    const zd ctt = lg[pos + 1];
    const zd* e = lg + pos + 2;
    zd c0 = row[(int64_t)(i1 - 1) * st];
    for (int k = i1; k < ie; ++k) {
      zd x1 = row[(int64_t)k * st];
      zrot_right_fixed(c0, x1, e[2 * (k - i1)], e[2 * (k - i1) + 1]);
      row[(int64_t)(k - 1) * st] = c0;
      c0 = x1;
      if (k == i1 && !(ctt.x == 0.0 && ctt.y == 0.0)) {
        // columns i1 .. ie except i1 + 1 (whose value rides in c0)
        row[(int64_t)(i1 - 1) * st] = zmul_fixed(row[(int64_t)(i1 - 1) * st], ctt);
        for (int c = i1 + 2; c <= ie; ++c) row[(int64_t)(c - 1) * st] = zmul_fixed(row[(int64_t)(c - 1) * st], ctt);
      }
    }
    row[(int64_t)(ie - 1) * st] = c0;
    const zd t = e[2 * (ie - i1)];
    if (!(t.x == 0.0 && t.y == 0.0)) row[(int64_t)(ie - 1) * st] = zmul_fixed(row[(int64_t)(ie - 1) * st], t);
    pos += 3 + 2 * (ie - i1);
  }
}

template <bool FAST, int NT = 128> GS_DEV void z_sweep_win2(zd* H, int64_t ld, int n, zd* Zg, int64_t ldz, bool wantz, zd* win,
                                             zd* lgbuf, int* redi, zd shift, int istart, int iend, int64_t& nstep,
                                             ZLog& zl) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int CS = NT - 64;  // threads of the chain wavefronts
  const int cw = tid - 64;
  const double ulp = GS_EPS;
  const FullLayout lay{ld};
  int found = -1;
  for (int mm = iend - 1 - tid; mm >= istart + 1; mm -= NT) {
    const zd h11 = HH(mm, mm);
    const zd h22 = HH(mm + 1, mm + 1);
    zd h11s = h11 - shift;
    double h21 = HH(mm + 1, mm).x;
    const double s = zabs1(h11s) + fabs(h21);
    h11s = h11s / s;
    h21 = h21 / s;
    const double h10 = HH(mm, mm - 1).x;
    if (fabs(h10) * fabs(h21) <= ulp * (zabs1(h11s) * (zabs1(h11) + zabs1(h22)))) {
      found = mm;
      break;
    }
  }
  found = block_max_int<NT>(found, redi);
  const int istart1 = (found > 0) ? found : istart;
  zd v0, v1;
  {
    const zd h11 = HH(istart1, istart1);
    zd h11s = h11 - shift;
    double h21 = HH(istart1 + 1, istart1).x;
    const double s = zabs1(h11s) + fabs(h21);
    h11s = h11s / s;
    h21 = h21 / s;
    v0 = h11s;
    v1 = zr(h21);
  }
  bool phase_pending = istart1 > istart;
  // Z log: room for this sweep's record, then its header
  const bool zlg = wantz && zl.g;
  const int64_t need = 3 + 2 * (int64_t)(iend - istart1);
  if (zlg && zl.pos + need > zl.cap) {
    __syncthreads();
    for (int row = 1 + tid; row <= n; row += NT) zlog_row(Zg + (row - 1), ldz, zl.g, zl.pos);
    __syncthreads();
    zl.pos = 0;
  }
  zd* rec = zlg ? zl.g + zl.pos : nullptr;
  if (zlg && tid == 0) {
    rec[0] = zmk((double)istart1, (double)iend);
    rec[1] = zr(0.0);
  }
  int pk0 = 1, pkend = 0, pwe = 0, pwb = 1;  // chunk whose chains are pending (none)
  int k0 = istart1;
  for (int c = 0;; ++c) {
    const bool chase = k0 <= iend - 1;
    const bool prev = pk0 <= pkend;
    if (!chase && !prev) break;
    const int kend = chase ? (phase_pending ? k0 : min(k0 + ZKC - 1, iend - 1)) : 0;
    const int wb = chase ? max(k0 - 1, 1) : 0, we = chase ? min(kend + 2, iend) : 0;
    zd* lg = lgbuf + (c & 1) * 2 * ZKC;
    const zd* plg = lgbuf + ((c + 1) & 1) * 2 * ZKC;
    if (wave == 0) {
      if (chase) {
        // the window's loads are issued first; while they are in flight the
        // previous chunk's reflectors are applied to the window's columns
        // right of the previous window (rows pk0 .. pkend + 1 in HBM), whose
        // last two rows (the window's rows wb, wb + 1) go to the window
        const int W = we - wb + 1;
        {
          // window columns straight into LDS (lane r: row wb + r), the rows
          // that the chains below provide left out
          const int Wu = __builtin_amdgcn_readfirstlane(W);
          const int cs = prev ? min(max(pwe + 1 - wb, 0), Wu) : Wu;
          const zd* gp = H + (wb - 1) + (int64_t)(wb - 1) * ld + lane;
          if (prev) {
            // the previous chunk is a full one (only a sweep's last chunk is shorter)
            const int col = pwe + 1 + lane;
            if (col <= we) zchain_full(H + (int64_t)(col - 1) * ld, plg, pk0, win + (pkend - wb) + (col - wb) * ZWL);
          }
          if (lane < Wu)
            for (int c = 0; c < cs; ++c)
              __builtin_amdgcn_global_load_lds((const void*)(gp + (int64_t)c * ld),
                                               (__attribute__((address_space(3))) void*)(win + c * ZWL), 16, 0, 0);
          if (lane < Wu && lane >= 2)
            for (int c = cs; c < Wu; ++c)
              __builtin_amdgcn_global_load_lds((const void*)(gp + (int64_t)c * ld),
                                               (__attribute__((address_space(3))) void*)(win + c * ZWL), 16, 0, 0);
          __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the window is in LDS
        }
        wsync();
        auto WIN = [&](int r, int cc) -> zd& { return win[(r - wb) + (cc - wb) * ZWL]; };
        // Step k in lanes: the row pass (rows k, k + 1) on column k + lane;
        // the column pass (columns k, k + 1) on row k + lane for lanes 0..2
        // and on row wb + lane - 3 (rows above k) for lanes 3 and up.  All
        // operands but the column pass's rows k and k + 1 are final before
        // the step's reflector, so they are loaded first (the wavefront's
        // LDS accesses complete in order); rows k and k + 1 come from the
        // row pass's lanes 0 and 1 by a quad permutation, and the next
        // reflector's inputs from the column pass's lanes 1 and 2, so no LDS
        // round trip lies between one reflector and the next.  Lanes with
        // nothing to do load in-window entries and store to a dummy area.
        zd* const dmy = lgbuf + 4 * ZKC;
        const bool hi3 = lane >= 3;
        const int cidx = hi3 ? lane - 3 : lane;
        const int laneZ = lane * ZWL;
        if (k0 > istart1) {
          v0 = WIN(k0, k0 - 1);
          v1 = WIN(k0 + 1, k0 - 1);
        }
        // the reflector's inputs, the same in every lane: kept in scalar
        // registers from one step to the next
        zd nv0 = zmk(rdfirst(v0.x), rdfirst(v0.y)), nv1 = zmk(rdfirst(v1.x), rdfirst(v1.y));
        zd* wk = &WIN(k0, k0);
        zd* lgk = lg;
        double betal = 0.0;  // lane s: beta of step k0 + s
        for (int k = k0; k <= kend; ++k, wk += ZWL + 1, lgk += 2) {
          const int kw = k - wb;
          const bool rv = lane <= we - k;
          zd* const rp = wk + (rv ? laneZ : 0);
          zd ra = rp[0], rb = rp[1];
          // column pass rows relative to k: lane (lanes 0..2), lane - 3 - kw (3 and up)
          const bool cv = cidx < (hi3 ? kw : (k + 2 <= iend ? 3 : 2));
          zd* const cp = wk + (cv ? (hi3 ? cidx - kw : lane) : 0);
          zd ca = cp[0], cb = cp[ZWL];
          v0 = nv0;
          v1 = nv1;
          const zd tau1 = FAST ? zreflector2_wave(v0, v1) : zreflector2(v0, v1);
          const zd v2 = v1;
          const double tau2 = tau1.x * v2.x - tau1.y * v2.y;
          {
            // ss = conj(tau1) ra + tau2 rb; ra -= ss; rb -= ss v2
            const double sx = fma(tau1.x, ra.x, fma(tau1.y, ra.y, tau2 * rb.x));
            const double sy = fma(tau1.x, ra.y, fma(-tau1.y, ra.x, tau2 * rb.y));
            ra = zmk(ra.x - sx, ra.y - sy);
            rb = zmk(fma(-sx, v2.x, fma(sy, v2.y, rb.x)), fma(-sx, v2.y, fma(-sy, v2.x, rb.y)));
          }
          {
            // lane 0 takes lane 1's (k, k + 1), lane 1 takes lane 0's (k + 1, k)
            const zd x10 = zmk(qperm<0xF5>(ra.x), qperm<0xF5>(ra.y));
            const zd x01 = zmk(qperm<0xA0>(rb.x), qperm<0xA0>(rb.y));
            const bool l0 = lane == 0, l1 = lane == 1;
            ca = zmk(l0 ? ra.x : (l1 ? x01.x : ca.x), l0 ? ra.y : (l1 ? x01.y : ca.y));
            cb = zmk(l0 ? x10.x : (l1 ? rb.x : cb.x), l0 ? x10.y : (l1 ? rb.y : cb.y));
          }
          {
            // ss = tau1 ca + tau2 cb; ca -= ss; cb -= ss conj(v2)
            const double sx = fma(tau1.x, ca.x, fma(-tau1.y, ca.y, tau2 * cb.x));
            const double sy = fma(tau1.x, ca.y, fma(tau1.y, ca.x, tau2 * cb.y));
            ca = zmk(ca.x - sx, ca.y - sy);
            cb = zmk(fma(-sx, v2.x, fma(-sy, v2.y, cb.x)), fma(-sy, v2.x, fma(sx, v2.y, cb.y)));
          }
          nv0 = zmk(rdlane(ca.x, 1), rdlane(ca.y, 1));
          nv1 = zmk(rdlane(ca.x, 2), rdlane(ca.y, 2));
          zd* const rq = (rv && lane >= 2) ? rp : dmy;
          rq[0] = ra;
          rq[1] = rb;
          zd* const cq = cv ? cp : dmy;
          cq[0] = ca;
          cq[ZWL] = cb;
          betal = lane == k - k0 ? v0.x : betal;
          // lane 0: the step's log entry; the other lanes store to the dummy area
          zd* const lq = lane == 0 ? lgk : dmy;
          lq[0] = tau1;
          lq[1] = v2;
          wsync();
        }
        // the annihilated columns: (k, k - 1) = beta, (k + 1, k - 1) = 0 (no
        // later step of the chunk reads them)
        {
          const int k = k0 + lane;
          if (k <= kend && k > istart1) {
            WIN(k, k - 1) = zr(betal);
            WIN(k + 1, k - 1) = zr(0.0);
          }
          wsync();
        }
        win_store<zd, ZW, ZWL>(H + (wb - 1) + (int64_t)(wb - 1) * ld, ld, W, win, lane);
        if (zlg && lane < 2 * (kend - k0 + 1)) rec[2 + 2 * (k0 - istart1) + lane] = lg[lane];
      }
    } else if (prev) {
      // the chain wavefronts (1 .. NT / 64 - 1) share the lines: thread cw
      // of CS takes every CS-th column, row above the window and row of Z
      const int lcol0 = (chase ? we : pwe) + 1;
      if (chase) {  // with no chase left (the sweep's last chains) wavefront 0 takes the columns
        if (pkend - pk0 + 1 == ZKC)
          for (int col = lcol0 + cw; col <= n; col += CS) zchain_full<false>(H + (int64_t)(col - 1) * ld, plg, pk0, nullptr);
        else
          for (int col = lcol0 + cw; col <= n; col += CS) zchain<true>(H + (int64_t)(col - 1) * ld, 1, plg, pk0, pkend);
      }
      if (chase && pkend - pk0 + 1 == ZKC) {
        for (int row = 1 + cw; row <= pwb - 1; row += CS) zchain_full_row<4>(H + (row - 1), ld, plg, pk0);
        if (wantz && !zlg)
          for (int row = 1 + cw; row <= n; row += CS) zchain_full_row<4>(Zg + (row - 1), ldz, plg, pk0);
      } else {
        for (int row = 1 + cw; row <= pwb - 1; row += CS) zchain<false>(H + (row - 1), ld, plg, pk0, pkend);
        if (wantz && !zlg)
          for (int row = 1 + cw; row <= n; row += CS) zchain<false>(Zg + (row - 1), ldz, plg, pk0, pkend);
      }
    }
    if (wave == 0 && !chase && prev)
      for (int col = pwe + 1 + lane; col <= n; col += 64) zchain<true>(H + (int64_t)(col - 1) * ld, 1, plg, pk0, pkend);
    __syncthreads();
    if (!chase) break;
    if (phase_pending) {
      // the first step's chains, then the phase scaling (GenericSchur.jl:461-482)
      phase_pending = false;
      if (wave == 0) {
        for (int col = we + 1 + lane; col <= n; col += 64) zchain<true>(H + (int64_t)(col - 1) * ld, 1, lg, k0, kend);
      } else {
        for (int row = 1 + cw; row <= wb - 1; row += CS) zchain<false>(H + (row - 1), ld, lg, k0, kend);
        if (wantz && !zlg)
          for (int row = 1 + cw; row <= n; row += CS) zchain<false>(Zg + (row - 1), ldz, lg, k0, kend);
      }
      __syncthreads();
      zd t = zr(1.0) - lg[0];
      t = t / zabs(t);
      const zd ctt = zconj(t);
      if (tid == 0) {
        HH(istart1 + 1, istart1) = HH(istart1 + 1, istart1) * ctt;
        if (istart1 + 2 <= iend) HH(istart1 + 2, istart1 + 1) = HH(istart1 + 2, istart1 + 1) * t;
      }
      __syncthreads();
      for (int cc = 1; cc <= n; ++cc) {
        const bool cin = (cc >= istart1 && cc <= iend && cc != istart1 + 1);
        const int rmax = min(cc - 1, iend);
        for (int r = 1 + tid; r <= rmax; r += NT) {
          const bool rin = (r >= istart1 && r != istart1 + 1);
          if (!rin && !cin) continue;
          zd x = HH(r, cc);
          if (rin) x = x * t;
          if (cin) x = x * ctt;
          HH(r, cc) = x;
        }
      }
      if (zlg) {
        if (tid == 0) rec[1] = ctt;
      } else if (wantz) {
        for (int r = 1 + tid; r <= n; r += NT)
          for (int cc = istart1; cc <= iend; ++cc)
            if (cc != istart1 + 1) ZZ(r, cc) = zmul_fixed(ZZ(r, cc), ctt);
      }
      __syncthreads();
      pk0 = 1;
      pkend = 0;
    } else {
      pk0 = k0;
      pkend = kend;
      pwe = we;
      pwb = wb;
    }
    k0 = kend + 1;
  }
  nstep += iend - istart1;
  // tail realness (GenericSchur.jl:486-500)
  zd t = HH(iend, iend - 1);
  const bool tail = t.y != 0.0;
  if (tail) {
    const double rtv = zabs(t);
    t = t / rtv;
    __syncthreads();
    if (tid == 0) HH(iend, iend - 1) = zr(rtv);
    const zd ctt = zconj(t);
    for (int cc = iend + 1 + tid; cc <= n; cc += NT) HH(iend, cc) = HH(iend, cc) * ctt;
    for (int r = 1 + tid; r <= iend - 1; r += NT) HH(r, iend) = HH(r, iend) * t;
    if (wantz && !zlg)
      for (int r = 1 + tid; r <= n; r += NT) ZZ(r, iend) = zmul_fixed(ZZ(r, iend), t);
  }
  if (zlg) {
    if (tid == 0) rec[need - 1] = tail ? t : zr(0.0);
    zl.pos += need;
  }
  __syncthreads();
}

// ------------------------------------------------------------------------
// gschur!(H::Hessenberg{Complex}, Z) (GenericSchur.jl:194-335).
// Returns 0, or the positive iteration count when maxiter was exceeded.
template <int NT, class L, bool FAST, int WIN = 0> GS_DEV int64_t z_qr(zd* H, L lay, int n, zd* Zg, int64_t ldz, bool wantz, zd* rot,
                                               int* redi, int64_t maxiter, int64_t maxinner, int64_t& nsweep,
                                               int64_t& nstep, ZLog* zl = nullptr) {
  const int tid = threadIdx.x;
  const double ulp = GS_EPS;
  const double smallnum = GS_FMIN * ((double)n / ulp);
  const double half = 0.5;
  const double threeq = 0.75;
  int64_t it = 0;
  int iend = n;
  while (iend >= 1) {
    int istart = 1;
    for (int64_t its = 0; its <= maxinner; ++its) {
      it += 1;
      if (it > maxiter) return it;
      // lowest negligible subdiagonal in istart..iend-1
      int found = -1;
      for (int is = iend - 1 - tid; is >= istart; is -= NT) {
        bool split = zabs1(HH(is + 1, is)) <= smallnum;
        if (!split) {
          double tst = zabs1(HH(is, is)) + zabs1(HH(is + 1, is + 1));
          if (tst == 0.0) {
            if (is - 1 >= 1) tst += fabs(HH(is, is - 1).x);
            if (is + 2 <= n) tst += fabs(HH(is + 2, is + 1).x);
          }
          if (fabs(HH(is + 1, is).x) <= ulp * tst) {
            const double ab = fmax(zabs1(HH(is + 1, is)), zabs1(HH(is, is + 1)));
            const double ba = fmin(zabs1(HH(is + 1, is)), zabs1(HH(is, is + 1)));
            const double aa = fmax(zabs1(HH(is + 1, is + 1)), zabs1(HH(is, is) - HH(is + 1, is + 1)));
            const double bb = fmin(zabs1(HH(is + 1, is + 1)), zabs1(HH(is, is) - HH(is + 1, is + 1)));
            const double s = aa + ab;
            split = ba * (ab / s) <= fmax(smallnum, ulp * (bb * (aa / s)));
          }
        }
        if (split) {
          found = is;
          break;
        }
      }
      found = block_max_int<NT>(found, redi);
      if (found >= 0) istart = found + 1;
      if (istart > 1) {
        bsync<NT>();
        if (tid == 0) HH(istart, istart - 1) = zr(0.0);
        __syncthreads();
      }
      if (istart >= iend) {
        iend -= 1;
        break;
      }
      zd t;
      if (its % 30 == 10) {
        const double s = threeq * fabs(HH(istart + 1, istart).x);
        t = zr(s) + HH(istart, istart);
      } else if (its % 30 == 20) {
        const double s = threeq * fabs(HH(iend, iend - 1).x);
        t = zr(s) + HH(iend, iend);
      } else {
        t = HH(iend, iend);
        const zd u = zsqrt(HH(iend - 1, iend)) * zsqrt(HH(iend, iend - 1));
        double s = zabs1(u);
        if (s != 0.0) {
          const zd x = half * (HH(iend - 1, iend - 1) - t);
          const double sx = zabs1(x);
          s = fmax(s, zabs1(x));
          const zd xs = x / s, us = u / s;
          zd y = s * zsqrt(xs * xs + us * us);
          if (sx > 0.0) {
            const zd xx = x / sx;
            if (xx.x * y.x + xx.y * y.y < 0.0) y = -y;
          }
          t = t - u * (u / (x + y));
        }
      }
      nsweep += 1;
      if constexpr (WIN == 2)
        z_sweep_win2<FAST, NT>(H, lay.ld, n, Zg, ldz, wantz, rot, rot + ZWL * ZW, redi, t, istart, iend, nstep, *zl);
      else
        z_sweep<NT, L, FAST>(H, lay, n, Zg, ldz, wantz, rot, redi, t, istart, iend, nstep);
    }
  }
  return 0;
}

// T = triu(H) (GenericSchur.jl:334), values = diag(T), unscaled with
// safescale!(T, cscale, anrm) when A was scaled (GenericSchur.jl:367-370).
template <int NT, class L> GS_DEV void z_finish(const zd* H, L lay, int n, zd* Tg, int64_t ldt, zd* Wg,
                                                double cscale, double anrm) {
  double umul[4];
  int nu = 0;
  if (cscale != 0.0) nu = safescale_muls(cscale, anrm, umul);
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
    const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
    zd v = (r > c) ? zr(0.0) : H[lay.idx(r + 1, c + 1)];
    for (int q = 0; q < nu; ++q) v = v * umul[q];
    Tg[r + c * ldt] = v;
    if (r == c) Wg[r] = v;
  }
}

// ------------------------------------------------------------------------
// Batched kernel 1: load, scale, Hessenberg reduction and Z formation with
// the whole matrix in LDS.  Writes H (zero below the subdiagonal) to T and,
// with wantZ, Q to Z.
template <int NT> __global__ void __launch_bounds__(NT) zhq_kernel(BatchArgs a) {
  extern __shared__ double smem[];
  const int b = blockIdx.x;
  if (b >= a.count) return;
  const int m = a.order[a.first + b];
  const int n = a.n[m];
  const int64_t o = a.off[m];
  if (n <= 0) {
    if (threadIdx.x == 0) a.info[m] = (n == 0) ? 0 : -1;
    return;
  }
  const int64_t ld = (int64_t)(n | 1);
  zd* H = reinterpret_cast<zd*>(smem);
  zd* tau = H + ld * n;
  zd* wk = tau + n;
  double* red = reinterpret_cast<double*>(wk + 2 * n + (zhq_colwise(n) ? 4 * n : 0));
  const zd* A = reinterpret_cast<const zd*>(a.A) + o;
  zd* T = reinterpret_cast<zd*>(a.T) + o;
  zd* Z = reinterpret_cast<zd*>(a.Z) + o;
  const bool wantz = (a.flags & 0x01) != 0;
  const bool hessin = (a.flags & 0x04) != 0;
  double cscale, anrm;
  const int st = z_prep<NT>(A, n, H, ld, n, a.flags, red, cscale, anrm);
  if (threadIdx.x == 0) {
    a.info[m] = st;
    a.scl[2 * m] = cscale;
    a.scl[2 * m + 1] = anrm;
  }
  if (st != 0) return;
  if (!hessin) z_hessenberg<NT, 2>(H, ld, n, tau, red, wk);
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
    const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
    T[e] = (r > c + 1) ? zr(0.0) : H[r + c * ld];
  }
  if (wantz && !hessin) {
    __syncthreads();
    z_form_q_inplace<NT, 2>(H, ld, n, tau, wk);
    for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
      const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
      Z[e] = H[r + c * ld];
    }
  }
}

// Batched kernel 2: QR iterations with the band of H in LDS, Z in HBM.
template <int NT, bool FAST> __global__ void __launch_bounds__(NT) zqr_kernel(BatchArgs a) {
  extern __shared__ double smem[];
  const int b = blockIdx.x;
  if (b >= a.count) return;
  const int m = a.order[a.first + b];
  const int n = a.n[m];
  if (n <= 0) return;
  if (a.info[m] != 0) return;  // rejected by zhq_kernel
  const int64_t o = a.off[m];
  BandLayout<2> lay;
  lay.init(n);
  zd* H = reinterpret_cast<zd*>(smem);
  zd* rot = H + BandLayout<2>::size(n);
  int* redi = reinterpret_cast<int*>(rot + 2 * n);
  zd* T = reinterpret_cast<zd*>(a.T) + o;
  zd* Z = reinterpret_cast<zd*>(a.Z) + o;
  zd* W = reinterpret_cast<zd*>(a.W) + a.woff[m];
  const bool wantz = (a.flags & 0x01) != 0;
  // load the band (T holds H with zeros below the subdiagonal)
  for (int j = 1; j <= n; ++j) {
    const int len = min(j + 2, n);
    const int64_t cb = lay.col(j);
    for (int i = 1 + threadIdx.x; i <= len; i += NT) H[cb + i - 1] = T[(i - 1) + (int64_t)(j - 1) * n];
  }
  __syncthreads();
  int64_t maxiter = a.maxiter > 0 ? a.maxiter : 100 * (int64_t)n;
  int64_t maxinner = a.maxinner > 0 ? a.maxinner : 30 * (int64_t)n;
  int64_t nsweep = 0, nstep = 0;
  const int64_t st = z_qr<NT, BandLayout<2>, FAST>(H, lay, n, Z, (int64_t)n, wantz, rot, redi, maxiter, maxinner, nsweep, nstep);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (a.stats) {
      a.stats[2 * m] = nsweep;
      a.stats[2 * m + 1] = nstep;
    }
    a.info[m] = (int32_t)(st > 0x7fffffff ? 0x7fffffff : st);
  }
  if (st != 0) return;
  z_finish<NT>(H, lay, n, T, (int64_t)n, W, a.scl[2 * m], a.scl[2 * m + 1]);
}

// Two-wavefront windowed QR kernel: the chase of chunk c overlaps the chains
// of chunk c-1 (z_sweep_win2); the control logic runs on both wavefronts.
// QR iterations of matrix m of a batch (the body of zqrw2_kernel and of its
// work-queue form); every thread of the workgroup calls it.
template <bool FAST, int NT>
GS_DEV void zqrw2_matrix(const BatchArgs& a, int m, zd* win, int* redi) {
  const int n = a.n[m];
  if (a.zlog && threadIdx.x == 0) a.llen[m] = 0;
  if (n <= 0) return;
  if (a.info[m] != 0) return;
  const int64_t o = a.off[m];
  zd* T = reinterpret_cast<zd*>(a.T) + o;
  zd* Z = reinterpret_cast<zd*>(a.Z) + o;
  zd* W = reinterpret_cast<zd*>(a.W) + a.woff[m];
  const bool wantz = (a.flags & 0x01) != 0;
  const FullLayout lay{(int64_t)n};
  int64_t maxiter = a.maxiter > 0 ? a.maxiter : 100 * (int64_t)n;
  int64_t maxinner = a.maxinner > 0 ? a.maxinner : 30 * (int64_t)n;
  int64_t nsweep = 0, nstep = 0;
  ZLog zl{nullptr, 0, 0};
  if (a.zlog && wantz && n >= a.zlog_min) {
    zl.g = reinterpret_cast<zd*>(a.zlog) + a.loff[m];
    zl.cap = a.loff[m + 1] - a.loff[m];
  }
  const int64_t st = z_qr<NT, FullLayout, FAST, 2>(T, lay, n, Z, (int64_t)n, wantz, win, redi, maxiter, maxinner,
                                                    nsweep, nstep, &zl);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (zl.g) a.llen[m] = zl.pos;
    if (a.stats) {
      a.stats[2 * m] = nsweep;
      a.stats[2 * m + 1] = nstep;
    }
    a.info[m] = (int32_t)(st > 0x7fffffff ? 0x7fffffff : st);
  }
  if (st != 0) return;
  z_finish<NT>(T, lay, n, T, (int64_t)n, W, a.scl[2 * m], a.scl[2 * m + 1]);
}

template <bool FAST, int NT = 128>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) zqrw2_kernel(BatchArgs a) {
  __shared__ zd win[ZWL * ZW + 4 * ZKC + ZWL + 1];
  __shared__ int redi[NT / 64 > 4 ? NT / 64 : 4];
  const int b = blockIdx.x;
  if (b >= a.count) return;
  // the chasing wavefront (0) issues ahead of the chain wavefronts sharing
  // its SIMD (flag 0x4000 turns this off for comparisons)
  if (!(a.flags & 0x4000) && threadIdx.x < 64) __builtin_amdgcn_s_setprio(2);
  zqrw2_matrix<FAST, NT>(a, a.order[a.first + b], win, redi);
}

// Work-queue form: a fixed number of workgroups, each taking the next
// matrix of order[first .. first + count) from the counter a.qctr until the
// queue is empty (every workgroup exits once the counter passes count).  The
// order is largest first, so the long matrices start first and the short
// ones fill the slots the long ones leave, across size bins.
template <bool FAST, int NT = 128>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4))) zqrw2_queue_kernel(BatchArgs a) {
  __shared__ zd win[ZWL * ZW + 4 * ZKC + ZWL + 1];
  __shared__ int redi[NT / 64 > 4 ? NT / 64 : 4];
  __shared__ int slot;
  if (!(a.flags & 0x4000) && threadIdx.x < 64) __builtin_amdgcn_s_setprio(2);
  for (;;) {
    if (threadIdx.x == 0) slot = atomicAdd(a.qctr, 1);
    __syncthreads();
    const int q = slot;
    __syncthreads();
    if (q >= a.count) break;
    zqrw2_matrix<FAST, NT>(a, a.order[a.first + q], win, redi);
    __syncthreads();
  }
}

// ------------------------------------------------------------------------
// Single-matrix kernel: H in HBM (the T output buffer), any n.
template <int NT> __global__ void __launch_bounds__(NT) zschur_global_kernel(SingleArgs a) {
  __shared__ double red[NT / 64 + 1];
  __shared__ int redi[NT / 64 + 1];
  const int n = (int)a.n;
  zd* A = reinterpret_cast<zd*>(a.A);
  zd* T = reinterpret_cast<zd*>(a.T);
  zd* Z = reinterpret_cast<zd*>(a.Z);
  zd* W = reinterpret_cast<zd*>(a.W);
  zd* tau = reinterpret_cast<zd*>(a.ws);
  zd* rot = tau + n;
  const int64_t ldt = a.ldt, ldz = a.ldz;
  const bool wantz = (a.flags & 0x01) != 0;
  const bool hessin = (a.flags & 0x04) != 0;
  double cscale, anrm;
  const int st0 = z_prep<NT>(A, a.lda, T, ldt, n, a.flags, red, cscale, anrm);
  if (st0 != 0) {
    if (threadIdx.x == 0) *a.info = st0;
    return;
  }
  if (!hessin) {
    z_hessenberg<NT>(T, ldt, n, tau, red, tau + 4 * n);
    if (wantz) {
      // hold H in Z, form Q in place in T, then swap the two
      for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
        const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
        Z[r + c * ldz] = T[r + c * ldt];
      }
      __syncthreads();
      z_form_q_inplace<NT>(T, ldt, n, tau, tau + 4 * n);
      for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
        const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
        const zd q = T[r + c * ldt];
        const zd h = Z[r + c * ldz];
        Z[r + c * ldz] = q;
        T[r + c * ldt] = h;
      }
      __syncthreads();
    }
  }
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
    const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
    if (r > c + 1) T[r + c * ldt] = zr(0.0);
  }
  __syncthreads();
  FullLayout lay{ldt};
  const int64_t maxiter = a.maxiter > 0 ? a.maxiter : 100 * (int64_t)n;
  const int64_t maxinner = a.maxinner > 0 ? a.maxinner : 30 * (int64_t)n;
  int64_t nsweep = 0, nstep = 0;
  const int64_t st = z_qr<NT, FullLayout, true>(T, lay, n, Z, ldz, wantz, rot, redi, maxiter, maxinner, nsweep, nstep);
  __syncthreads();
  if (threadIdx.x == 0) {
    if (a.stats) {
      a.stats[0] = nsweep;
      a.stats[1] = nstep;
    }
    *a.info = (int32_t)(st > 0x7fffffff ? 0x7fffffff : st);
  }
  if (st != 0) return;
  z_finish<NT>(T, lay, n, T, ldt, W, cscale, anrm);
}

// Hessenberg reduction and Q formation of one matrix in global memory, for
// the windowed QR kernel above the LDS-resident orders: the scaled A (or
// the Hessenberg input) as H in T, zero below the subdiagonal, Q (or the
// caller's Z) in Z, (cscale, anrm) of _scale! in scl[0], scl[1] and the
// status in info.  The same steps as zschur_global_kernel before its QR
// phase; zqrw2_kernel then runs the sweeps with T and Z in HBM.
template <int NT, bool PREP_ONLY> __global__ void __launch_bounds__(NT) zhq_global_kernel(SingleArgs a, double* scl) {
  __shared__ double red[NT / 64 + 1];
  const int n = (int)a.n;
  zd* A = reinterpret_cast<zd*>(a.A);
  zd* T = reinterpret_cast<zd*>(a.T);
  zd* Z = reinterpret_cast<zd*>(a.Z);
  zd* tau = reinterpret_cast<zd*>(a.ws);
  const int64_t ldt = a.ldt, ldz = a.ldz;
  const bool wantz = (a.flags & 0x01) != 0;
  const bool hessin = (a.flags & 0x04) != 0;
  double cscale, anrm;
  const int st0 = z_prep<NT>(A, a.lda, T, ldt, n, a.flags, red, cscale, anrm);
  if (threadIdx.x == 0) {
    *a.info = st0;
    scl[0] = cscale;
    scl[1] = anrm;
  }
  if (st0 != 0 || PREP_ONLY) return;
  if (!hessin) {
    z_hessenberg<NT>(T, ldt, n, tau, red, tau + 4 * n);
    if (wantz) {
      for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
        const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
        Z[r + c * ldz] = T[r + c * ldt];
      }
      __syncthreads();
      z_form_q_inplace<NT>(T, ldt, n, tau, tau + 4 * n);
      for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
        const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
        const zd q = T[r + c * ldt];
        const zd h = Z[r + c * ldz];
        Z[r + c * ldz] = q;
        T[r + c * ldt] = h;
      }
      __syncthreads();
    }
  }
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += NT) {
    const int c = (int)(e / n), r = (int)(e - (int64_t)c * n);
    if (r > c + 1) T[r + c * ldt] = zr(0.0);
  }
}

// The Z log of one matrix above the LDS range replayed onto Z's rows in
// HBM (ld ld), one row per thread over the whole GPU: zlog_row's
// operations, taken in groups of eight steps whose row entries are read
// before any of the group's writes (as zlog_kernel does in LDS), so that a
// group waits for memory once.
__global__ void __launch_bounds__(64) zlog_rows_kernel(int n, zd* Z, int64_t ld, const zd* __restrict__ lg,
                                                       const int64_t* llen) {
  const int r = blockIdx.x * 64 + threadIdx.x;
  if (r >= n) return;
  zd* row = Z + r;
  const int64_t len = *llen;
  for (int64_t pos = 0; pos < len;) {
    const zd h0 = lg[pos];
    const int i1 = (int)h0.x, ie = (int)h0.y;
    const zd ctt = lg[pos + 1];
    const zd* e = lg + pos + 2;
    zd c0 = row[(int64_t)(i1 - 1) * ld];
    {  // first step, then the phase scaling when logged
      zd x1 = row[(int64_t)i1 * ld];
      zrot_right_fixed(c0, x1, e[0], e[1]);
      row[(int64_t)(i1 - 1) * ld] = c0;
      c0 = x1;
      if (!(ctt.x == 0.0 && ctt.y == 0.0)) {
        row[(int64_t)(i1 - 1) * ld] = zmul_fixed(row[(int64_t)(i1 - 1) * ld], ctt);
        for (int c = i1 + 2; c <= ie; ++c) row[(int64_t)(c - 1) * ld] = zmul_fixed(row[(int64_t)(c - 1) * ld], ctt);
      }
    }
    int k = i1 + 1;
    for (; k + 7 < ie; k += 8) {
      zd xa[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) xa[j] = row[(int64_t)(k + j) * ld];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        zrot_right_fixed(c0, xa[j], e[2 * (k + j - i1)], e[2 * (k + j - i1) + 1]);
        row[(int64_t)(k + j - 1) * ld] = c0;
        c0 = xa[j];
      }
    }
    for (; k < ie; ++k) {
      zd x1 = row[(int64_t)k * ld];
      zrot_right_fixed(c0, x1, e[2 * (k - i1)], e[2 * (k - i1) + 1]);
      row[(int64_t)(k - 1) * ld] = c0;
      c0 = x1;
    }
    row[(int64_t)(ie - 1) * ld] = c0;
    const zd t = e[2 * (ie - i1)];
    if (!(t.x == 0.0 && t.y == 0.0)) row[(int64_t)(ie - 1) * ld] = zmul_fixed(row[(int64_t)(ie - 1) * ld], t);
    pos += 3 + 2 * (int64_t)(ie - i1);
  }
}

hipError_t launch_zlog_rows(int n, double* Z, int64_t ld, const double* lg, const int64_t* llen, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zlog_rows_kernel, dim3((n + 63) / 64), dim3(64), 0, s, n, reinterpret_cast<zd*>(Z), ld,
                     reinterpret_cast<const zd*>(lg), llen);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
// The same reduction and Q for orders from kZhbMinN up, over the whole GPU:
// per column p the reflector (one wavefront), the left update's column dots
// and rank-1 update, the right update's gemv partials (64-row blocks times
// column chunks), their sum with column p+1's update, and the right rank-1
// update of columns p+2..; then Q's backward accumulation (z_form_q_inplace's
// steps, two launches per reflector).  The element operations are those of
// z_hessenberg and z_form_q_inplace; only the dot products' summation
// order differs.  Every launch first checks the status z_prep wrote.
constexpr int kZhbMinN = 192;
struct ZhbArgs {
  int n = 0, p = 0, ks = 1, kchunk = 0;
  zd* H = nullptr;
  int64_t ld = 0;
  zd* Q = nullptr;
  int64_t ldq = 0;
  zd* tau = nullptr;
  zd* wk = nullptr;   // n
  zd* wk2 = nullptr;  // n
  zd* P = nullptr;    // 16 x n
  const int32_t* info = nullptr;
};

__global__ void __launch_bounds__(64) zhb_reflect_kernel(ZhbArgs a) {
  if (*a.info != 0) return;
  z_hess_reflect(a.H, a.ld, a.n, a.p, a.tau);
}

// lmul!(H', A[p+1:n, p+1:n]): va_j = conj(tau) (H[p+1, j] + sum_{r > p+1}
// conj(v_r) H[r, j]), one wavefront per column j > p
__global__ void __launch_bounds__(256) zhb_ldots_kernel(ZhbArgs a) {
  if (*a.info != 0) return;
  const int n = a.n, p = a.p, lane = threadIdx.x & 63;
  const int j = p + 1 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  const zd* H = a.H;
  const int64_t ld = a.ld;
  zd acc = zr(0.0);
  for (int r = p + 2 + lane; r < n; r += 64) acc = acc + zconj(H[r + p * ld]) * H[r + (int64_t)j * ld];
  acc = wave_zsum(acc);
  if (lane == 0) a.wk[j] = zconj(a.tau[p]) * (H[(p + 1) + (int64_t)j * ld] + acc);
}

// column j -= va_j [1; v] over rows p+1 ..
__global__ void __launch_bounds__(256) zhb_lrank1_kernel(ZhbArgs a) {
  if (*a.info != 0) return;
  const int n = a.n, p = a.p;
  const int r = p + 1 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = p + 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
  if (r >= n || j >= n) return;
  zd* H = a.H;
  const int64_t ld = a.ld;
  const zd h = H[r + (int64_t)j * ld], va = a.wk[j];
  H[r + (int64_t)j * ld] = (r == p + 1) ? h - va : h - va * H[r + p * ld];
}

// rmul!(A[:, p+1:n], H): P[y][r] = sum over chunk y of columns c > p+1 of
// H[r, c] v_c, all rows
__global__ void __launch_bounds__(256) zhb_rgemv_kernel(ZhbArgs a) {
  __shared__ double raw[4 * 64 * 2];
  zd* red = reinterpret_cast<zd*>(raw);
  if (*a.info != 0) return;
  const int n = a.n, p = a.p, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int r = blockIdx.x * 64 + lane;
  const int c0 = p + 2 + blockIdx.y * a.kchunk, c1 = min(n, c0 + a.kchunk);
  const zd* H = a.H;
  const int64_t ld = a.ld;
  zd acc = zr(0.0);
  if (r < n)
    for (int c = c0 + wv; c < c1; c += 4) acc = acc + H[r + (int64_t)c * ld] * H[c + p * ld];
  red[wv * 64 + lane] = acc;
  __syncthreads();
  if (wv == 0 && r < n) a.P[(int64_t)blockIdx.y * n + r] = ((red[lane] + red[64 + lane]) + red[128 + lane]) + red[192 + lane];
}

// x_r = H[r, p+1] + sum_y P[y][r]; H[r, p+1] -= tau x_r; wk2[r] = -tau x_r
__global__ void __launch_bounds__(256) zhb_rcombine_kernel(ZhbArgs a) {
  if (*a.info != 0) return;
  const int n = a.n, p = a.p;
  const int r = blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  zd acc = zr(0.0);
  for (int y = 0; y < a.ks; ++y) acc = acc + a.P[(int64_t)y * n + r];
  zd* h = a.H + r + (int64_t)(p + 1) * a.ld;
  const zd t = a.tau[p];
  const zd x = *h + acc;
  *h = *h - t * x;
  a.wk2[r] = x * (-t);
}

// H[r, c] += wk2[r] conj(v_c), all rows, columns p+2 ..
__global__ void __launch_bounds__(256) zhb_rrank1_kernel(ZhbArgs a) {
  if (*a.info != 0) return;
  const int n = a.n, p = a.p;
  const int r = blockIdx.x * 64 + (threadIdx.x & 63);
  const int c = p + 2 + blockIdx.y * 4 + (threadIdx.x >> 6);
  if (r >= n || c >= n) return;
  zd* H = a.H;
  const int64_t ld = a.ld;
  H[r + (int64_t)c * ld] = H[r + (int64_t)c * ld] + a.wk2[r] * zconj(H[c + p * ld]);
}

__global__ void __launch_bounds__(256) zhb_identity_kernel(ZhbArgs a) {
  if (*a.info != 0) return;
  const int64_t nn = (int64_t)a.n * a.n;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nn; e += (int64_t)gridDim.x * 256) {
    const int c = (int)(e / a.n), r = (int)(e - (int64_t)c * a.n);
    a.Q[r + c * a.ldq] = zr(r == c ? 1.0 : 0.0);
  }
}

// reflector k = a.p applied to Q(c:, c+1:), c = k + 1: s_j = tau (Q[c, j] +
// sum_{r > c} conj(v_r) Q[r, j]), one wavefront per column
__global__ void __launch_bounds__(256) zhb_qdots_kernel(ZhbArgs a) {
  if (*a.info != 0) return;
  const int n = a.n, k = a.p, c = k + 1, lane = threadIdx.x & 63;
  const int j = c + 1 + blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  zd acc = zr(0.0);
  for (int r = c + 1 + lane; r < n; r += 64) acc = acc + zconj(a.H[r + k * a.ld]) * a.Q[r + (int64_t)j * a.ldq];
  acc = wave_zsum(acc);
  if (lane == 0) a.wk[j] = a.tau[k] * (a.Q[c + (int64_t)j * a.ldq] + acc);
}

// column j > c -= s_j [1; v]; column c = e_c - tau [1; v]
__global__ void __launch_bounds__(256) zhb_qupdate_kernel(ZhbArgs a) {
  if (*a.info != 0) return;
  const int n = a.n, k = a.p, c = k + 1;
  const int r = c + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = c + blockIdx.y * 4 + (threadIdx.x >> 6);
  if (r >= n || j >= n) return;
  const zd t = a.tau[k];
  zd* q = a.Q + r + (int64_t)j * a.ldq;
  if (j == c) {
    *q = (r == c) ? zr(1.0) - t : -(t * a.H[r + k * a.ld]);
  } else {
    const zd h = *q, sj = a.wk[j];
    *q = (r == c) ? h - sj : h - sj * a.H[r + k * a.ld];
  }
}

// T = triu(H, -1)
__global__ void __launch_bounds__(256) zhb_zero_kernel(ZhbArgs a) {
  if (*a.info != 0) return;
  const int64_t nn = (int64_t)a.n * a.n;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < nn; e += (int64_t)gridDim.x * 256) {
    const int c = (int)(e / a.n), r = (int)(e - (int64_t)c * a.n);
    if (r > c + 1) a.H[r + c * a.ld] = zr(0.0);
  }
}

static hipError_t launch_zhb(const SingleArgs& g, hipStream_t s) {
  const int n = (int)g.n;
  ZhbArgs a;
  a.n = n;
  a.H = reinterpret_cast<zd*>(g.T);
  a.ld = g.ldt;
  a.Q = reinterpret_cast<zd*>(g.Z);
  a.ldq = g.ldz;
  a.tau = reinterpret_cast<zd*>(g.ws);
  a.wk = a.tau + n;
  a.wk2 = a.wk + n;
  a.P = a.wk2 + n;
  a.info = g.info;
  const int nrb = (n + 63) / 64;
  const int gridn = std::min(4096, (int)(((int64_t)n * n + 255) / 256));
  for (int p = 0; p + 2 <= n; ++p) {
    a.p = p;
    const int m = n - 1 - p;   // rows / columns p+1 ..
    const int m2 = n - 2 - p;  // columns p+2 ..
    hipLaunchKernelGGL(zhb_reflect_kernel, dim3(1), dim3(64), 0, s, a);
    hipLaunchKernelGGL(zhb_ldots_kernel, dim3((m + 3) / 4), dim3(256), 0, s, a);
    hipLaunchKernelGGL(zhb_lrank1_kernel, dim3((m + 63) / 64, (m + 3) / 4), dim3(256), 0, s, a);
    a.ks = 1;
    a.kchunk = 0;
    if (m2 > 0) {
      a.ks = std::max(1, std::min(16, (512 + nrb - 1) / nrb));
      a.kchunk = (m2 + a.ks - 1) / a.ks;
      a.ks = (m2 + a.kchunk - 1) / a.kchunk;
    }
    hipLaunchKernelGGL(zhb_rgemv_kernel, dim3(nrb, a.ks), dim3(256), 0, s, a);
    hipLaunchKernelGGL(zhb_rcombine_kernel, dim3((n + 255) / 256), dim3(256), 0, s, a);
    if (m2 > 0) hipLaunchKernelGGL(zhb_rrank1_kernel, dim3(nrb, (m2 + 3) / 4), dim3(256), 0, s, a);
  }
  if (g.flags & 0x01) {
    hipLaunchKernelGGL(zhb_identity_kernel, dim3(gridn), dim3(256), 0, s, a);
    for (int k = n - 2; k >= 0; --k) {
      const int c = k + 1;
      a.p = k;
      if (n - 1 - c > 0) hipLaunchKernelGGL(zhb_qdots_kernel, dim3((n - 1 - c + 3) / 4), dim3(256), 0, s, a);
      hipLaunchKernelGGL(zhb_qupdate_kernel, dim3((n - c + 63) / 64, (n - c + 3) / 4), dim3(256), 0, s, a);
    }
  }
  hipLaunchKernelGGL(zhb_zero_kernel, dim3(gridn), dim3(256), 0, s, a);
  return hipGetLastError();
}

size_t zhq_global_ws_doubles(int64_t n) { return 2 * (19 * (size_t)n) + 8; }

hipError_t launch_zhq_global(const SingleArgs& a, double* scl, hipStream_t s) {
  const bool hessin = (a.flags & 0x04) != 0;
  if (!hessin && a.n >= kZhbMinN) {
    hipLaunchKernelGGL((zhq_global_kernel<1024, true>), dim3(1), dim3(1024), 0, s, a, scl);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_zhb(a, s);
  }
  hipLaunchKernelGGL((zhq_global_kernel<1024, false>), dim3(1), dim3(1024), 0, s, a, scl);
  return hipGetLastError();
}

// ------------------------------------------------------------------------
size_t zhq_lds_bytes(int n, int nt) {
  // the most any order up to n needs (the partial sums drop out above the
  // orders where they fit)
  size_t b = zhq_need_bytes(n, nt, false);
  for (int k = 1; k <= n; ++k)
    if (zhq_colwise(k)) b = std::max(b, zhq_need_bytes(k, nt, true));
  return b;
}
size_t zqr_lds_bytes(int n, int nt) {
  return 16 * ((size_t)BandLayout<2>::size(n) + 2 * (size_t)(n > 0 ? n : 1)) + 4 * (size_t)(nt / 64 + 1);
}

int zschur_max_lds_n() {
  // largest n whose Hessenberg-phase footprint fits the 160 KiB of LDS of a
  // gfx950 CU
  int n = 1;
  while (zhq_lds_bytes(n + 1, 256) <= 160 * 1024) ++n;
  return n;
}

template <int NT> static hipError_t launch_zhq_nt(const BatchArgs& a, size_t lds, hipStream_t s) {
  (void)hipFuncSetAttribute((const void*)zhq_kernel<NT>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(zhq_kernel<NT>, dim3(a.count), dim3(NT), lds, s, a);
  return hipGetLastError();
}
template <int NT> static hipError_t launch_zqr_nt(const BatchArgs& a, size_t lds, hipStream_t s) {
  if (a.flags & 0x1000) {  // reference-form sweep (kept for A/B measurements)
    (void)hipFuncSetAttribute((const void*)zqr_kernel<NT, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL((zqr_kernel<NT, false>), dim3(a.count), dim3(NT), lds, s, a);
  } else {
    (void)hipFuncSetAttribute((const void*)zqr_kernel<NT, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL((zqr_kernel<NT, true>), dim3(a.count), dim3(NT), lds, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_zhq(const BatchArgs& a, int nmax, int nt, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  const size_t lds = zhq_lds_bytes(nmax, nt);
  switch (nt) {
    case 64: return launch_zhq_nt<64>(a, lds, s);
    case 128: return launch_zhq_nt<128>(a, lds, s);
    case 256: return launch_zhq_nt<256>(a, lds, s);
    case 512: return launch_zhq_nt<512>(a, lds, s);
    case 1024: return launch_zhq_nt<1024>(a, lds, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_zqr(const BatchArgs& a, int nmax, int nt, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  const size_t lds = zqr_lds_bytes(nmax, nt);
  switch (nt) {
    case 64: return launch_zqr_nt<64>(a, lds, s);
    case 128: return launch_zqr_nt<128>(a, lds, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_zqrw(const BatchArgs& a, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  if (a.flags & 0x1000)
    hipLaunchKernelGGL((zqrw2_kernel<false>), dim3(a.count), dim3(128), 0, s, a);
  else
    hipLaunchKernelGGL((zqrw2_kernel<true>), dim3(a.count), dim3(128), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_zqrw_queue(const BatchArgs& a, int slots, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  const int grid = std::max(1, std::min(slots, a.count));
  if (a.flags & 0x1000)
    hipLaunchKernelGGL((zqrw2_queue_kernel<false>), dim3(grid), dim3(128), 0, s, a);
  else
    hipLaunchKernelGGL((zqrw2_queue_kernel<true>), dim3(grid), dim3(128), 0, s, a);
  return hipGetLastError();
}

__global__ void fill_i32_kernel(int* p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

hipError_t launch_fill_i32(int* p, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_i32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

// One matrix above the LDS-resident orders: wavefront 0 chases, the other
// nt / 64 - 1 wavefronts of the workgroup share the chains (H's rows above
// and columns right of the window, and Z), whose count grows with n.
hipError_t launch_zqrw_single(const BatchArgs& a, int nt, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  switch (nt) {
    case 128: hipLaunchKernelGGL((zqrw2_kernel<true, 128>), dim3(a.count), dim3(128), 0, s, a); break;
    case 256: hipLaunchKernelGGL((zqrw2_kernel<true, 256>), dim3(a.count), dim3(256), 0, s, a); break;
    case 512: hipLaunchKernelGGL((zqrw2_kernel<true, 512>), dim3(a.count), dim3(512), 0, s, a); break;
    case 1024: hipLaunchKernelGGL((zqrw2_kernel<true, 1024>), dim3(a.count), dim3(1024), 0, s, a); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// After a bin's QR launch: each matrix's Z log applied to Z by one
// workgroup of ceil(n / 64) wavefronts, all of Z held in LDS (row r, column
// c at zs[(c - 1) n + r]), each thread replaying the log on one row with
// zlog_row's operations.  Each wavefront stages a sweep's record in its own
// LDS buffer and loads the next record into registers while it replays the
// current one, so the wavefronts never wait for each other.
constexpr int ZLREC = 200;  // >= 3 + 2 (n - 1) for every batched complex order
__global__ void __launch_bounds__(128) zlog_kernel(BatchArgs a) {
  extern __shared__ zd zs[];
  __shared__ zd rbuf[2][ZLREC];  // a separate array: its reads do not wait on the row writes
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int b = blockIdx.x;
  if (b >= a.count) return;
  const int m = a.order[a.first + b];
  const int n = a.n[m];
  if (n <= 0 || wave * 64 >= n) return;
  const int64_t len = a.llen[m];
  if (len <= 0) return;
  zd* Z = reinterpret_cast<zd*>(a.Z) + a.off[m];
  const zd* lg = reinterpret_cast<const zd*>(a.zlog) + a.loff[m];
  const int r = min(tid, n - 1);  // threads past the last row repeat it, unstored
  const bool live = tid < n;
  zd t[4];
  auto fetch = [&](int64_t p) {
#pragma unroll
    for (int j = 0; j < 4; ++j) t[j] = lg[min(p + lane + 64 * j, len - 1)];
  };
  fetch(0);
  for (int c0 = 0; c0 < n; c0 += 8) {
    zd v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = Z[r + (int64_t)min(c0 + i, n - 1) * n];
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (c0 + i < n && live) zs[(c0 + i) * n + r] = v[i];
  }
  zd* row = zs + r;
  const int st = n;
  zd* rec = rbuf[wave];
  for (int64_t pos = 0; pos < len;) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if (lane + 64 * j < ZLREC) rec[lane + 64 * j] = t[j];
    wsync();
    const zd h0 = rec[0];
    const int i1 = __builtin_amdgcn_readfirstlane((int)h0.x), ie = __builtin_amdgcn_readfirstlane((int)h0.y);
    const int64_t rl = 3 + 2 * (int64_t)(ie - i1);
    if (pos + rl < len) fetch(pos + rl);
    if (live) {
      const zd ctt = rec[1];
      const zd* e = rec + 2;
      zd c0 = row[(i1 - 1) * st];
      {  // first step, then the phase scaling when logged
        zd x1 = row[i1 * st];
        zrot_right_fixed(c0, x1, e[0], e[1]);
        row[(i1 - 1) * st] = c0;
        c0 = x1;
        if (!(ctt.x == 0.0 && ctt.y == 0.0)) {
          row[(i1 - 1) * st] = zmul_fixed(row[(i1 - 1) * st], ctt);
#pragma unroll 4
          for (int c = i1 + 2; c <= ie; ++c) row[(c - 1) * st] = zmul_fixed(row[(c - 1) * st], ctt);
        }
      }
      // groups of eight steps: the group's row entries (which its own steps
      // do not write first) are read before any of its writes, so a group
      // waits for LDS once
      int k = i1 + 1;
      for (; k + 7 < ie; k += 8) {
        zd xa[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) xa[j] = row[(k + j) * st];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          zrot_right_fixed(c0, xa[j], e[2 * (k + j - i1)], e[2 * (k + j - i1) + 1]);
          row[(k + j - 1) * st] = c0;
          c0 = xa[j];
        }
      }
      for (; k < ie; ++k) {
        zd x1 = row[k * st];
        zrot_right_fixed(c0, x1, e[2 * (k - i1)], e[2 * (k - i1) + 1]);
        row[(k - 1) * st] = c0;
        c0 = x1;
      }
      row[(ie - 1) * st] = c0;
      const zd tt = e[2 * (ie - i1)];
      if (!(tt.x == 0.0 && tt.y == 0.0)) row[(ie - 1) * st] = zmul_fixed(row[(ie - 1) * st], tt);
    }
    wsync();  // the record buffer is rewritten for the next sweep
    pos += rl;
  }
  if (live) {
    for (int c0 = 0; c0 < n; c0 += 8) {
      zd v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = zs[min(c0 + i, n - 1) * n + r];
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (c0 + i < n) Z[r + (int64_t)(c0 + i) * n] = v[i];
    }
  }
}

static size_t zlog_lds_bytes(int nmax) { return (size_t)nmax * nmax * sizeof(zd); }

hipError_t launch_zlog(const BatchArgs& a, int nmax, hipStream_t s) {
  if (a.count <= 0 || !a.zlog) return hipSuccess;
  if (3 + 2 * (nmax - 1) > ZLREC || nmax > 128) return hipErrorInvalidValue;
  const int waves = (nmax + 63) / 64;
  const size_t lds = zlog_lds_bytes(nmax);
  (void)hipFuncSetAttribute((const void*)zlog_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL(zlog_kernel, dim3(a.count), dim3(64 * waves), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_zschur_global(const SingleArgs& a, int nt, hipStream_t s) {
  switch (nt) {
    case 64:
      hipLaunchKernelGGL(zschur_global_kernel<64>, dim3(1), dim3(64), 0, s, a);
      break;
    case 256:
      hipLaunchKernelGGL(zschur_global_kernel<256>, dim3(1), dim3(256), 0, s, a);
      break;
    case 1024:
      hipLaunchKernelGGL(zschur_global_kernel<1024>, dim3(1), dim3(1024), 0, s, a);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace gs
