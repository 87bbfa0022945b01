// Host entry points of the eigen! pieces around the Schur factorization
// (gs_balance_*, gs_balance_apply, gs_triangularize_d, gs_enormalize,
// gs_eig_rcond) and of the reordering and conditioning of a complex Schur
// form (gs_ordschur_z, gs_ordschur_d, gs_eigcond_z), see include/gschur_mi355x.h.  Each is one matrix (or one
// set of columns) on the handle's stream; host pointers are staged through
// the handle's buffers, device pointers (GS_DEVICE_PTRS) used directly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/gschur_mi355x.h"
#include "handle.h"
#include "internal.h"

using namespace gs;

namespace {

#define ECHK(call)                                                  \
  do {                                                              \
    hipError_t e_ = (call);                                         \
    if (e_ != hipSuccess) return handle_fail_hip(h, e_, #call);     \
  } while (0)

struct Lock {
  std::lock_guard<std::mutex> lk;
  explicit Lock(gs_handle* h) : lk(handle_mutex(h)) {}
};

// grow-only staging buffer i of the handle (the eigen! entry points use
// buffers 8 .. 11)
void* stage(gs_handle* h, int i, size_t bytes, hipError_t& e) {
  DevBuf& b = handle_buf(h, 8 + i);
  e = b.ensure(std::max<size_t>(bytes, 64));
  return b.p;
}

}  // namespace

extern "C" {

static int balance_impl(gs_handle* h, bool cplx, int64_t n, double* A, int64_t lda, int scale, int permute,
                        int32_t* out4, int32_t* sp, double* D, int flags) {
  if (!h) return GS_E_ARG;
  if (n < 0 || n > INT32_MAX || lda < std::max<int64_t>(n, 1)) return handle_fail(h, GS_E_DIM, "bad dimension");
  if (!out4 || (n > 0 && (!A || !sp || !D))) return handle_fail(h, GS_E_ARG, "null pointer");
  Lock lk(h);
  ECHK(hipSetDevice(handle_device(h)));
  if (n == 0) {
    out4[0] = 1;
    out4[1] = 0;
    out4[2] = 1;
    out4[3] = 0;
    return GS_OK;
  }
  hipStream_t s = handle_stream(h);
  const size_t es = cplx ? 16 : 8;
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  double* dA = A;
  int64_t dlda = lda;
  int32_t *dout = out4, *dsp = sp;
  double* dD = D;
  hipError_t e;
  if (!dev) {
    dA = static_cast<double*>(stage(h, 0, es * (size_t)n * n, e));
    ECHK(e);
    char* meta = static_cast<char*>(stage(h, 1, 16 + 4 * (size_t)n + 8 * (size_t)n + 64, e));
    ECHK(e);
    dout = reinterpret_cast<int32_t*>(meta);
    dsp = dout + 4;
    dD = reinterpret_cast<double*>(meta + 16 + ((4 * (size_t)n + 7) & ~size_t(7)));
    dlda = n;
    ECHK(copy_cols(dA, es * n, A, es * lda, es * n, n, hipMemcpyHostToDevice, s));
  }
  ECHK(launch_balance(cplx, dA, dlda, (int)n, scale, permute, dout, dsp, dD, s));
  if (!dev) {
    ECHK(copy_cols(A, es * lda, dA, es * n, es * n, n, hipMemcpyDeviceToHost, s));
    ECHK(hipMemcpyAsync(out4, dout, 16, hipMemcpyDeviceToHost, s));
    ECHK(hipMemcpyAsync(sp, dsp, 4 * n, hipMemcpyDeviceToHost, s));
    ECHK(hipMemcpyAsync(D, dD, 8 * n, hipMemcpyDeviceToHost, s));
  }
  if (!dev || !(flags & GS_ASYNC)) ECHK(hipStreamSynchronize(s));
  return GS_OK;
}

int gs_balance_z(gs_handle* h, int64_t n, double* A, int64_t lda, int scale, int permute, int32_t* out4, int32_t* sp,
                 double* D, int flags) {
  return balance_impl(h, true, n, A, lda, scale, permute, out4, sp, D, flags);
}

int gs_balance_d(gs_handle* h, int64_t n, double* A, int64_t lda, int scale, int permute, int32_t* out4, int32_t* sp,
                 double* D, int flags) {
  return balance_impl(h, false, n, A, lda, scale, permute, out4, sp, D, flags);
}

int gs_balance_apply(gs_handle* h, int64_t n, int64_t m, const int32_t* bal3, const int32_t* sp, const double* D,
                     double* V, int64_t ldv, int inverse, int flags) {
  if (!h) return GS_E_ARG;
  if (n < 0 || m < 0 || n > INT32_MAX || m > INT32_MAX || ldv < std::max<int64_t>(n, 1))
    return handle_fail(h, GS_E_DIM, "bad dimension");
  if (n == 0 || m == 0) return GS_OK;
  if (!bal3 || !sp || !D || !V) return handle_fail(h, GS_E_ARG, "null pointer");
  if (bal3[0] < 1 || bal3[1] > n || bal3[0] > std::max<int32_t>(bal3[1], 1))
    return handle_fail(h, GS_E_ARG, "balancer bounds outside 1..n");
  for (int64_t i = 0; i < n && !(flags & GS_DEVICE_PTRS); ++i)
    if ((i < bal3[0] - 1 || i >= bal3[1]) && (sp[i] < 1 || sp[i] > n))
      return handle_fail(h, GS_E_ARG, "balancer permutation entry outside 1..n");
  Lock lk(h);
  ECHK(hipSetDevice(handle_device(h)));
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  const int32_t *dbal = bal3, *dsp = sp;
  const double* dD = D;
  double* dV = V;
  int64_t dldv = ldv;
  hipError_t e;
  if (!dev) {
    dV = static_cast<double*>(stage(h, 0, 16 * (size_t)n * m, e));
    ECHK(e);
    char* meta = static_cast<char*>(stage(h, 1, 16 + 4 * (size_t)n + 8 * (size_t)n + 64, e));
    ECHK(e);
    int32_t* mb = reinterpret_cast<int32_t*>(meta);
    double* md = reinterpret_cast<double*>(meta + 16 + ((4 * (size_t)n + 7) & ~size_t(7)));
    ECHK(hipMemcpyAsync(mb, bal3, 12, hipMemcpyHostToDevice, s));
    ECHK(hipMemcpyAsync(mb + 4, sp, 4 * n, hipMemcpyHostToDevice, s));
    ECHK(hipMemcpyAsync(md, D, 8 * n, hipMemcpyHostToDevice, s));
    ECHK(copy_cols(dV, 16 * n, V, 16 * ldv, 16 * n, m, hipMemcpyHostToDevice, s));
    dbal = mb;
    dsp = mb + 4;
    dD = md;
    dldv = n;
  }
  ECHK(launch_balback((int)n, (int)m, dbal, dsp, dD, dV, dldv, inverse, s));
  if (!dev) ECHK(copy_cols(V, 16 * ldv, dV, 16 * n, 16 * n, m, hipMemcpyDeviceToHost, s));
  if (!dev || !(flags & GS_ASYNC)) ECHK(hipStreamSynchronize(s));
  return GS_OK;
}

int gs_triangularize_d(gs_handle* h, int64_t n, const double* T, int64_t ldt, const double* Z, int64_t ldz,
                       double* Tc, int64_t ldtc, double* Zc, int64_t ldzc, double* W, int flags) {
  if (!h) return GS_E_ARG;
  const int64_t n1 = std::max<int64_t>(n, 1);
  if (n < 0 || n > INT32_MAX || ldt < n1 || ldtc < n1 || (Z && (ldz < n1 || ldzc < n1)))
    return handle_fail(h, GS_E_DIM, "bad dimension");
  if (n == 0) return GS_OK;
  if (!T || !Tc || !W || (Z && !Zc)) return handle_fail(h, GS_E_ARG, "null pointer");
  Lock lk(h);
  ECHK(hipSetDevice(handle_device(h)));
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  const double *dT = T, *dZ = Z;
  double *dTc = Tc, *dZc = Zc, *dW = W;
  int64_t a = ldt, b = ldz, c = ldtc, d = ldzc;
  const size_t nn = (size_t)n * n;
  hipError_t e;
  if (!dev) {
    double* base = static_cast<double*>(stage(h, 0, 8 * (2 * nn + 4 * nn + 2 * (size_t)n), e));
    ECHK(e);
    double* sT = base;
    double* sZ = base + nn;
    dTc = base + 2 * nn;
    dZc = base + 4 * nn;
    dW = base + 6 * nn;
    ECHK(copy_cols(sT, 8 * n, T, 8 * ldt, 8 * n, n, hipMemcpyHostToDevice, s));
    if (Z) ECHK(copy_cols(sZ, 8 * n, Z, 8 * ldz, 8 * n, n, hipMemcpyHostToDevice, s));
    dT = sT;
    dZ = Z ? sZ : nullptr;
    a = b = c = d = n;
  }
  if (!Z) dZc = dTc;  // not written
  ECHK(launch_triangularize((int)n, dT, a, dZ, b, dTc, c, dZ ? dZc : nullptr, d, dW, s));
  if (!dev) {
    ECHK(copy_cols(Tc, 16 * ldtc, dTc, 16 * n, 16 * n, n, hipMemcpyDeviceToHost, s));
    if (Z) ECHK(copy_cols(Zc, 16 * ldzc, dZc, 16 * n, 16 * n, n, hipMemcpyDeviceToHost, s));
    ECHK(hipMemcpyAsync(W, dW, 16 * n, hipMemcpyDeviceToHost, s));
  }
  if (!dev || !(flags & GS_ASYNC)) ECHK(hipStreamSynchronize(s));
  return GS_OK;
}

int gs_enormalize(gs_handle* h, int64_t n, int64_t m, double* V, int64_t ldv, int flags) {
  if (!h) return GS_E_ARG;
  if (n < 0 || m < 0 || n > INT32_MAX || m > INT32_MAX || ldv < std::max<int64_t>(n, 1))
    return handle_fail(h, GS_E_DIM, "bad dimension");
  if (n == 0 || m == 0) return GS_OK;
  if (!V) return handle_fail(h, GS_E_ARG, "null pointer");
  Lock lk(h);
  ECHK(hipSetDevice(handle_device(h)));
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  double* dV = V;
  int64_t dl = ldv;
  hipError_t e;
  if (!dev) {
    dV = static_cast<double*>(stage(h, 0, 16 * (size_t)n * m, e));
    ECHK(e);
    ECHK(copy_cols(dV, 16 * n, V, 16 * ldv, 16 * n, m, hipMemcpyHostToDevice, s));
    dl = n;
  }
  ECHK(launch_enormalize((int)n, (int)m, dV, dl, s));
  if (!dev) ECHK(copy_cols(V, 16 * ldv, dV, 16 * n, 16 * n, m, hipMemcpyDeviceToHost, s));
  if (!dev || !(flags & GS_ASYNC)) ECHK(hipStreamSynchronize(s));
  return GS_OK;
}

int gs_eig_rcond(gs_handle* h, int64_t n, int64_t m, const double* VL, int64_t ldvl, const double* VR, int64_t ldvr,
                 double* rconde, int flags) {
  if (!h) return GS_E_ARG;
  if (n < 0 || m < 0 || n > INT32_MAX || m > INT32_MAX || ldvl < std::max<int64_t>(n, 1) ||
      ldvr < std::max<int64_t>(n, 1))
    return handle_fail(h, GS_E_DIM, "bad dimension");
  if (n == 0 || m == 0) return GS_OK;
  if (!VL || !VR || !rconde) return handle_fail(h, GS_E_ARG, "null pointer");
  Lock lk(h);
  ECHK(hipSetDevice(handle_device(h)));
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  const double *dL = VL, *dR = VR;
  double* dC = rconde;
  int64_t a = ldvl, b = ldvr;
  hipError_t e;
  if (!dev) {
    double* base = static_cast<double*>(stage(h, 0, 8 * (4 * (size_t)n * m + (size_t)m), e));
    ECHK(e);
    double* sL = base;
    double* sR = base + 2 * (size_t)n * m;
    dC = base + 4 * (size_t)n * m;
    ECHK(copy_cols(sL, 16 * n, VL, 16 * ldvl, 16 * n, m, hipMemcpyHostToDevice, s));
    ECHK(copy_cols(sR, 16 * n, VR, 16 * ldvr, 16 * n, m, hipMemcpyHostToDevice, s));
    dL = sL;
    dR = sR;
    a = b = n;
  }
  ECHK(launch_eig_rcond((int)n, (int)m, dL, a, dR, b, dC, s));
  if (!dev) ECHK(hipMemcpyAsync(rconde, dC, 8 * m, hipMemcpyDeviceToHost, s));
  if (!dev || !(flags & GS_ASYNC)) ECHK(hipStreamSynchronize(s));
  return GS_OK;
}

int gs_ordschur_z(gs_handle* h, int64_t n, double* T, int64_t ldt, double* Z, int64_t ldz, const int32_t* select,
                  double* W, int flags) {
  if (!h) return GS_E_ARG;
  const int64_t n1 = std::max<int64_t>(n, 1);
  if (n < 0 || n > INT32_MAX || ldt < n1 || (Z && ldz < n1)) return handle_fail(h, GS_E_DIM, "bad dimension");
  if (n == 0) return GS_OK;
  if (!T || !select || !W) return handle_fail(h, GS_E_ARG, "null pointer");
  Lock lk(h);
  ECHK(hipSetDevice(handle_device(h)));
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  const size_t nn = (size_t)n * n;
  hipError_t e;
  // staging: select (and, with host pointers, T, Z, W)
  char* base = static_cast<char*>(stage(h, 0, 4 * (size_t)n + 256 + (dev ? 0 : 16 * (2 * nn + (size_t)n)), e));
  ECHK(e);
  int32_t* dsel = reinterpret_cast<int32_t*>(base);
  double* dT = T;
  double* dZ = Z;
  double* dW = W;
  int64_t a = ldt, b = ldz;
  ECHK(hipMemcpyAsync(dsel, select, 4 * (size_t)n, hipMemcpyHostToDevice, s));
  if (!dev) {
    double* sb = reinterpret_cast<double*>(base + ((4 * (size_t)n + 255) / 256) * 256);
    dT = sb;
    dZ = Z ? sb + 2 * nn : nullptr;
    dW = sb + 4 * nn;
    ECHK(copy_cols(dT, 16 * n, T, 16 * ldt, 16 * n, n, hipMemcpyHostToDevice, s));
    if (Z) ECHK(copy_cols(dZ, 16 * n, Z, 16 * ldz, 16 * n, n, hipMemcpyHostToDevice, s));
    a = b = n;
  }
  ECHK(launch_ordschur((int)n, dT, a, dZ, b, Z != nullptr, dsel, dW, s));
  if (!dev) {
    ECHK(copy_cols(T, 16 * ldt, dT, 16 * n, 16 * n, n, hipMemcpyDeviceToHost, s));
    if (Z) ECHK(copy_cols(Z, 16 * ldz, dZ, 16 * n, 16 * n, n, hipMemcpyDeviceToHost, s));
    ECHK(hipMemcpyAsync(W, dW, 16 * n, hipMemcpyDeviceToHost, s));
  }
  // select is read from the staging buffer until the launch is done
  ECHK(hipStreamSynchronize(s));
  return GS_OK;
}

int gs_ordschur_d(gs_handle* h, int64_t n, double* T, int64_t ldt, double* Z, int64_t ldz, const int32_t* select,
                  double* W, int32_t* info, int flags) {
  if (!h) return GS_E_ARG;
  const int64_t n1 = std::max<int64_t>(n, 1);
  if (n < 0 || n > INT32_MAX || ldt < n1 || (Z && ldz < n1)) return handle_fail(h, GS_E_DIM, "bad dimension");
  if (!info) return handle_fail(h, GS_E_ARG, "null pointer");
  info[0] = 0;
  if (n == 0) return GS_OK;
  if (!T || !select || !W) return handle_fail(h, GS_E_ARG, "null pointer");
  Lock lk(h);
  ECHK(hipSetDevice(handle_device(h)));
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  const size_t nn = (size_t)n * n;
  hipError_t e;
  // staging: select and info (and, with host pointers, T, Z, W)
  const size_t hdr = ((4 * (size_t)n + 4 + 255) / 256) * 256;
  char* base = static_cast<char*>(stage(h, 0, hdr + (dev ? 0 : 8 * (2 * nn + 2 * (size_t)n)), e));
  ECHK(e);
  int32_t* dsel = reinterpret_cast<int32_t*>(base);
  int32_t* dinfo = dsel + n;
  double* dT = T;
  double* dZ = Z;
  double* dW = W;
  int64_t a = ldt, b = ldz;
  ECHK(hipMemcpyAsync(dsel, select, 4 * (size_t)n, hipMemcpyHostToDevice, s));
  if (!dev) {
    double* sb = reinterpret_cast<double*>(base + hdr);
    dT = sb;
    dZ = Z ? sb + nn : nullptr;
    dW = sb + 2 * nn;
    ECHK(copy_cols(dT, 8 * n, T, 8 * ldt, 8 * n, n, hipMemcpyHostToDevice, s));
    if (Z) ECHK(copy_cols(dZ, 8 * n, Z, 8 * ldz, 8 * n, n, hipMemcpyHostToDevice, s));
    a = b = n;
  }
  ECHK(launch_ordschur_d((int)n, dT, a, dZ, b, Z != nullptr, dsel, dW, dinfo, s));
  if (!dev) {
    ECHK(copy_cols(T, 8 * ldt, dT, 8 * n, 8 * n, n, hipMemcpyDeviceToHost, s));
    if (Z) ECHK(copy_cols(Z, 8 * ldz, dZ, 8 * n, 8 * n, n, hipMemcpyDeviceToHost, s));
    ECHK(hipMemcpyAsync(W, dW, 16 * (size_t)n, hipMemcpyDeviceToHost, s));
  }
  ECHK(hipMemcpyAsync(info, dinfo, 4, hipMemcpyDeviceToHost, s));
  ECHK(hipStreamSynchronize(s));
  return GS_OK;
}

int gs_eigcond_z(gs_handle* h, int64_t n, const double* T, int64_t ldt, int64_t nsub, double* out2, int flags) {
  if (!h) return GS_E_ARG;
  if (n < 2 || n > INT32_MAX || ldt < n) return handle_fail(h, GS_E_DIM, "bad dimension");
  if (nsub < 1 || nsub >= n) return handle_fail(h, GS_E_ARG, "nsub must lie in 1 .. n-1");
  if (!T || !out2) return handle_fail(h, GS_E_ARG, "null pointer");
  Lock lk(h);
  ECHK(hipSetDevice(handle_device(h)));
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  const size_t nn = (size_t)n * n, N = (size_t)nsub * (size_t)(n - nsub);
  hipError_t e;
  double* base = static_cast<double*>(stage(h, 0, 16 * (3 * N + (dev ? 0 : nn)) + 64, e));
  ECHK(e);
  double* work = base;
  double* dout = base + 6 * N;
  const double* dT = T;
  int64_t a = ldt;
  if (!dev) {
    double* sT = dout + 8;
    ECHK(copy_cols(sT, 16 * n, T, 16 * ldt, 16 * n, n, hipMemcpyHostToDevice, s));
    dT = sT;
    a = n;
  }
  ECHK(launch_eigcond((int)n, dT, a, (int)nsub, work, dout, s));
  ECHK(hipMemcpyAsync(out2, dout, 16, hipMemcpyDeviceToHost, s));
  ECHK(hipStreamSynchronize(s));
  return GS_OK;
}

}  // extern "C"
