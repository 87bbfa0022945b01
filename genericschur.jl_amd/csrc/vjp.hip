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

// Orders whose two factors do not fit LDS: c_k for one matrix, one
// wavefront per column, with the factors split into planes for a real GEMM:
// P = [Re(Y C) | Im(Y C)] and Q = [Re X | Im X] (n x 2n each, column-major,
// ld = n), so that Abar = Re(Y C X^H) = P Q^T.  info = 1 when some
// y_k^H x_k vanished.
__global__ void __launch_bounds__(256) vjp_planes_kernel(int n, const zd* VR, const zd* VL, const zd* g,
                                                         double* P, double* Q, int32_t* info) {
  const int lane = threadIdx.x & 63;
  const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (k >= n) return;
  const zd* x = VR + (int64_t)k * n;
  const zd* y = VL + (int64_t)k * n;
  zd s = zr(0.0);
  for (int i = lane; i < n; i += 64) s = s + zconj(y[i]) * x[i];
  s = wave_zsum(s);
  const zd d = zconj(s);
  const bool zero = d.x == 0.0 && d.y == 0.0;
  const zd c = zero ? zr(0.0) : g[k] / d;
  if (zero && lane == 0) *info = 1;
  const int64_t nn = (int64_t)n * n;
  for (int i = lane; i < n; i += 64) {
    const zd t = c * y[i];
    P[i + (int64_t)k * n] = t.x;
    P[nn + i + (int64_t)k * n] = t.y;
    Q[i + (int64_t)k * n] = x[i].x;
    Q[nn + i + (int64_t)k * n] = x[i].y;
  }
}

hipError_t launch_vjp_planes(int n, const double* VR, const double* VL, const double* g, double* P, double* Q,
                             int32_t* info, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(vjp_planes_kernel, dim3((n + 3) / 4), dim3(256), 0, s, n, reinterpret_cast<const zd*>(VR),
                     reinterpret_cast<const zd*>(VL), reinterpret_cast<const zd*>(g), P, Q, info);
  return hipGetLastError();
}

__global__ void vjp_info_clear_kernel(int32_t* info, int count) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < count) info[i] = 0;
}

hipError_t launch_vjp_info_clear(int32_t* info, int count, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(vjp_info_clear_kernel, dim3((count + 255) / 256), dim3(256), 0, s, info, count);
  return hipGetLastError();
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
