// Eigenvalue vector-Jacobian product on gfx950 (no counterpart in the
// reference; the backward of the eigenvalue rrule).
//
// For A = X Lambda X^{-1} with right eigenvectors x_k and left eigenvectors
// y_k (A^H y_k = conj(lambda_k) y_k), a real loss L(lambda) with cotangent
// gbar_k = dL/dRe(lambda_k) + i dL/dIm(lambda_k) per eigenvalue slot has
//     dL = Re sum_k conj(gbar_k) y_k^H dA x_k / (y_k^H x_k),
// so the gradient with respect to a real A is
//     Abar = Re( sum_k c_k y_k x_k^H ),   c_k = gbar_k / conj(y_k^H x_k),
// i.e. Re(X^{-H} diag(gbar) X^H).  The eigenvectors are recomputed from the
// Schur factors by the eigenvector kernels (normalised, any scaling cancels
// in c_k y_k x_k^H); this kernel forms c_k and the rank-n sum, one
// workgroup per matrix, with the two n x n complex factors staged in LDS.
#include "dev_common.h"
#include "internal.h"

namespace gs {

template <int NT> __global__ void __launch_bounds__(NT) vjp_assemble_kernel(VjpArgs a) {
  extern __shared__ double smem[];
  const int b = blockIdx.x;
  if (b >= a.count) return;
  const int m = a.order ? a.order[b] : b;
  const int n = a.n[m];
  if (n <= 0) {
    if (threadIdx.x == 0) a.info[m] = 0;
    return;
  }
  const int64_t o = a.off[m];
  const zd* VR = reinterpret_cast<const zd*>(a.VR) + o;
  const zd* VL = reinterpret_cast<const zd*>(a.VL) + o;
  const zd* g = reinterpret_cast<const zd*>(a.gbar) + a.woff[m];
  double* Ab = a.Abar + o;
  zd* sR = reinterpret_cast<zd*>(smem);  // n x n, ld = n + 1
  const int ld = n + 1;
  zd* sL = sR + (size_t)ld * n;
  zd* c = sL + (size_t)ld * n;
  int* bad = reinterpret_cast<int*>(c + n);
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += NT) {
    const int col = e / n, row = e - col * n;
    sR[row + col * ld] = VR[e];
    sL[row + col * ld] = VL[e];
  }
  if (tid == 0) *bad = 0;
  __syncthreads();
  // c_k = gbar_k / conj(y_k^H x_k), one wavefront per column k
  const int lane = tid & 63, w = tid >> 6;
  for (int k = w; k < n; k += NT / 64) {
    zd s = zr(0.0);
    for (int i = lane; i < n; i += 64) s = s + zconj(sL[i + k * ld]) * sR[i + k * ld];
    s = wave_zsum(s);
    if (lane == 0) {
      const zd d = zconj(s);
      if (d.x == 0.0 && d.y == 0.0) {
        *bad = 1;
        c[k] = zr(0.0);
      } else {
        c[k] = g[k] / d;
      }
    }
  }
  __syncthreads();
  // Abar[r, s] = Re sum_k c_k yL[r, k] conj(xR[s, k])
  for (int e = tid; e < n * n; e += NT) {
    const int s = e / n, r = e - s * n;
    double acc = 0.0;
    for (int k = 0; k < n; ++k) {
      const zd t = c[k] * sL[r + k * ld];
      const zd x = sR[s + k * ld];
      acc += t.x * x.x + t.y * x.y;  // Re(t * conj(x))
    }
    Ab[e] = acc;
  }
  if (tid == 0) a.info[m] = *bad;
}

size_t vjp_lds_bytes(int n) { return 16 * (2 * (size_t)(n + 1) * n + (size_t)n) + 16; }

hipError_t launch_vjp_assemble(const VjpArgs& a, int nmax, hipStream_t s) {
  if (a.count <= 0) return hipSuccess;
  const size_t lds = vjp_lds_bytes(nmax);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  (void)hipFuncSetAttribute((const void*)vjp_assemble_kernel<256>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds);
  hipLaunchKernelGGL(vjp_assemble_kernel<256>, dim3(a.count), dim3(256), lds, s, a);
  return hipGetLastError();
}

}  // namespace gs
