// C entry points of the Hermitian / real symmetric eigenproblem (gs_heev_z,
// gs_heev_d, gs_hetrd_z, gs_hetrd_d, gs_steqr), declared in
// include/gschur_mi355x.h.  Host pointers are staged through handle buffer
// 12, device pointers (GS_DEVICE_PTRS) used directly; buffer 13 holds the
// eigensolver's workspace (e, the per-matrix rotation streams and states).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "../../include/gschur_mi355x.h"
#include "handle.h"
#include "internal.h"

using namespace gs;

namespace {

#define HCHK(call)                                              \
  do {                                                          \
    hipError_t e_ = (call);                                     \
    if (e_ != hipSuccess) return handle_fail_hip(h, e_, #call); \
  } while (0)

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

int heev_impl(gs_handle* h, bool cplx, int64_t count, int64_t n, const double* A, int64_t lda, int64_t sA,
              int upper, double* W, double* V, int64_t ldv, int64_t sV, int32_t* info, int flags) {
  if (!h) return GS_E_ARG;
  if (count < 0 || count > INT32_MAX || n < 0) return handle_fail(h, GS_E_DIM, "bad dimension");
  if (count == 0 || n == 0) return GS_OK;
  const bool wantv = (flags & GS_WANTZ) != 0;
  if (lda < n || (count > 1 && sA < lda * n) || (wantv && (ldv < n || (count > 1 && sV < ldv * n))))
    return handle_fail(h, GS_E_DIM, "bad leading dimension or stride");
  if (!A || !W || (wantv && !V)) return handle_fail(h, GS_E_ARG, "null pointer");
  if (n > heev_big_max_n()) return handle_fail(h, GS_E_UNSUPPORTED, "order above gs_heev_max_n");
  // above the LDS-resident orders: one matrix at a time over the whole GPU
  const bool big = n > heev_max_n(cplx);
  std::lock_guard<std::mutex> lk(handle_mutex(h));
  HCHK(hipSetDevice(handle_device(h)));
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  const size_t es = cplx ? 16 : 8, nn = (size_t)n * n, cnt = (size_t)count;
  HeevArgs a;
  a.count = (int)count;
  a.n = (int)n;
  a.upper = upper ? 1 : 0;
  a.mode = 0;
  a.sort = (flags & GS_SORT) ? 1 : 0;
  a.maxsweeps = 30 * (int)n;
  a.sW = n;
  // workspace: e (count x n), stream lengths, states, streams (with
  // vectors); for the large orders the working matrices of one matrix
  void* bigws = nullptr;
  if (big) {
    DevBuf& ws = handle_buf(h, 13);
    HCHK(ws.ensure(heev_big_ws_bytes((int)n, cplx)));
    bigws = ws.p;
    // testing: a shorter stream forces the in-place finish
    if (const char* ev = std::getenv("GSCHUR_HEEV_LOGCAP")) a.logcap = (int)std::strtoul(ev, nullptr, 10);
  } else {
    const size_t bE = up256(cnt * (size_t)n * 8), bP = up256(cnt * 4), bS = up256(cnt * heev_state_bytes());
    size_t cap = heev_log_records((int)n);
    // testing: a shorter stream forces the in-place finish of phase 3
    if (const char* ev = std::getenv("GSCHUR_HEEV_LOGCAP")) cap = std::max<size_t>(1, std::strtoul(ev, nullptr, 10));
    const size_t bL = wantv ? cnt * cap * 16 : 0;
    DevBuf& ws = handle_buf(h, 13);
    HCHK(ws.ensure(bE + bP + bS + bL));
    char* q = static_cast<char*>(ws.p);
    a.E = reinterpret_cast<double*>(q);
    a.sE = n;
    a.logpos = reinterpret_cast<int*>(q + bE);
    a.state = q + bE + bP;
    a.log = wantv ? static_cast<void*>(q + bE + bP + bS) : nullptr;
    a.logcap = (int)cap;
  }
  if (dev) {
    a.A = A;
    a.lda = lda;
    a.sA = sA;
    a.W = W;
    a.Z = wantv ? V : nullptr;
    a.ldz = ldv;
    a.sZ = sV;
    a.info = info;
  } else {
    const size_t bA = up256(cnt * nn * es), bW = up256(cnt * n * 8), bV = wantv ? up256(cnt * nn * es) : 0;
    const size_t bI = up256(cnt * 4);
    DevBuf& buf = handle_buf(h, 12);
    HCHK(buf.ensure(bA + bW + bV + bI));
    char* p = static_cast<char*>(buf.p);
    double* dA = reinterpret_cast<double*>(p);
    double* dW = reinterpret_cast<double*>(p + bA);
    double* dV = wantv ? reinterpret_cast<double*>(p + bA + bW) : nullptr;
    int32_t* dI = reinterpret_cast<int32_t*>(p + bA + bW + bV);
    const size_t ed = es / 8;  // doubles per element
    if (lda == n && (count == 1 || sA == (int64_t)nn)) {
      HCHK(hipMemcpyAsync(dA, A, cnt * nn * es, hipMemcpyHostToDevice, s));
    } else {
      for (int64_t k = 0; k < count; ++k)
        HCHK(copy_cols(dA + k * nn * ed, n * es, A + k * sA * ed, lda * es, n * es, n, hipMemcpyHostToDevice,
                              s));
    }
    a.A = dA;
    a.lda = n;
    a.sA = (int64_t)nn;
    a.W = dW;
    a.Z = dV;
    a.ldz = n;
    a.sZ = (int64_t)nn;
    a.info = dI;
    HCHK(big ? launch_heev_big(a, cplx, bigws, s) : launch_heev(a, cplx, s));
    HCHK(hipMemcpyAsync(W, dW, cnt * n * 8, hipMemcpyDeviceToHost, s));
    if (wantv) {
      if (ldv == n && (count == 1 || sV == (int64_t)nn)) {
        HCHK(hipMemcpyAsync(V, dV, cnt * nn * es, hipMemcpyDeviceToHost, s));
      } else {
        for (int64_t k = 0; k < count; ++k)
          HCHK(copy_cols(V + k * sV * ed, ldv * es, dV + k * nn * ed, n * es, n * es, n,
                                hipMemcpyDeviceToHost, s));
      }
    }
    if (info) HCHK(hipMemcpyAsync(info, dI, cnt * 4, hipMemcpyDeviceToHost, s));
    HCHK(hipStreamSynchronize(s));
    return GS_OK;
  }
  HCHK(big ? launch_heev_big(a, cplx, bigws, s) : launch_heev(a, cplx, s));
  if (!(flags & GS_ASYNC)) HCHK(hipStreamSynchronize(s));
  return GS_OK;
}

int hetrd_impl(gs_handle* h, bool cplx, int64_t n, const double* A, int64_t lda, int upper, double* d, double* e,
               double* Q, int64_t ldq, int flags) {
  if (!h) return GS_E_ARG;
  if (n < 0 || lda < std::max<int64_t>(n, 1) || (Q && ldq < std::max<int64_t>(n, 1)))
    return handle_fail(h, GS_E_DIM, "bad dimension");
  if (n == 0) return GS_OK;
  if (!A || !d || (n > 1 && !e)) return handle_fail(h, GS_E_ARG, "null pointer");
  if (n > heev_big_max_n()) return handle_fail(h, GS_E_UNSUPPORTED, "order above gs_heev_max_n");
  const bool big = n > heev_max_n(cplx);
  std::lock_guard<std::mutex> lk(handle_mutex(h));
  HCHK(hipSetDevice(handle_device(h)));
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  const size_t es = cplx ? 16 : 8, nn = (size_t)n * n;
  void* bigws = nullptr;
  if (big) {
    DevBuf& ws = handle_buf(h, 13);
    HCHK(ws.ensure(heev_big_ws_bytes((int)n, cplx)));
    bigws = ws.p;
  }
  HeevArgs a;
  a.count = 1;
  a.n = (int)n;
  a.upper = upper ? 1 : 0;
  a.mode = 2;
  if (dev) {
    a.A = A;
    a.lda = lda;
    a.W = d;
    a.E = e;
    a.Z = Q;
    a.ldz = ldq;
    HCHK(big ? launch_heev_big(a, cplx, bigws, s) : launch_heev(a, cplx, s));
    if (!(flags & GS_ASYNC)) HCHK(hipStreamSynchronize(s));
    return GS_OK;
  }
  const size_t bA = up256(nn * es), bd = up256((size_t)n * 8), bQ = Q ? up256(nn * es) : 0;
  DevBuf& buf = handle_buf(h, 12);
  HCHK(buf.ensure(bA + 2 * bd + bQ));
  char* p = static_cast<char*>(buf.p);
  double* dA = reinterpret_cast<double*>(p);
  double* dd = reinterpret_cast<double*>(p + bA);
  double* de = reinterpret_cast<double*>(p + bA + bd);
  double* dQ = Q ? reinterpret_cast<double*>(p + bA + 2 * bd) : nullptr;
  HCHK(copy_cols(dA, n * es, A, lda * es, n * es, n, hipMemcpyHostToDevice, s));
  a.A = dA;
  a.lda = n;
  a.W = dd;
  a.E = de;
  a.Z = dQ;
  a.ldz = n;
  HCHK(big ? launch_heev_big(a, cplx, bigws, s) : launch_heev(a, cplx, s));
  HCHK(hipMemcpyAsync(d, dd, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  if (n > 1) HCHK(hipMemcpyAsync(e, de, (size_t)(n - 1) * 8, hipMemcpyDeviceToHost, s));
  if (Q) HCHK(copy_cols(Q, ldq * es, dQ, n * es, n * es, n, hipMemcpyDeviceToHost, s));
  HCHK(hipStreamSynchronize(s));
  return GS_OK;
}

}  // namespace

extern "C" {

int gs_heev_z(gs_handle* h, int64_t count, int64_t n, const double* A, int64_t lda, int64_t strideA, int upper,
              double* W, double* V, int64_t ldv, int64_t strideV, int32_t* info, int flags) {
  return heev_impl(h, true, count, n, A, lda, strideA, upper, W, V, ldv, strideV, info, flags);
}

int gs_heev_d(gs_handle* h, int64_t count, int64_t n, const double* A, int64_t lda, int64_t strideA, int upper,
              double* W, double* V, int64_t ldv, int64_t strideV, int32_t* info, int flags) {
  return heev_impl(h, false, count, n, A, lda, strideA, upper, W, V, ldv, strideV, info, flags);
}

int gs_heev_max_n(int is_complex) {
  (void)is_complex;
  return heev_big_max_n();
}

int gs_hetrd_z(gs_handle* h, int64_t n, const double* A, int64_t lda, int upper, double* d, double* e, double* Q,
               int64_t ldq, int flags) {
  return hetrd_impl(h, true, n, A, lda, upper, d, e, Q, ldq, flags);
}

int gs_hetrd_d(gs_handle* h, int64_t n, const double* A, int64_t lda, int upper, double* d, double* e, double* Q,
               int64_t ldq, int flags) {
  return hetrd_impl(h, false, n, A, lda, upper, d, e, Q, ldq, flags);
}

int gs_steqr(gs_handle* h, int64_t n, double* d, double* e, double* Z, int64_t ldz, int64_t zrows, int zcomplex,
             int32_t* info, int flags) {
  if (!h) return GS_E_ARG;
  if (n < 0 || zrows < 0 || (Z && ldz < std::max<int64_t>(zrows, 1))) return handle_fail(h, GS_E_DIM, "bad dimension");
  if (n == 0) {
    if (info) info[0] = 0;
    return GS_OK;
  }
  if (!d || (n > 1 && !e)) return handle_fail(h, GS_E_ARG, "null pointer");
  if (n > heev_big_max_n()) return handle_fail(h, GS_E_UNSUPPORTED, "order above 4096");
  if (zrows > INT32_MAX) return handle_fail(h, GS_E_DIM, "too many rows of Z");
  std::lock_guard<std::mutex> lk(handle_mutex(h));
  HCHK(hipSetDevice(handle_device(h)));
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  const size_t es = zcomplex ? 16 : 8;
  const bool wantz = Z != nullptr && zrows > 0;
  SteqrArgs a;
  a.n = (int)n;
  a.zrows = wantz ? (int)zrows : 0;
  a.maxsweeps = 30 * (int)n;
  // above the LDS log's range: the rotations through a stream in HBM
  void* bigws = nullptr;
  if (n > steqr_max_n()) {
    DevBuf& ws = handle_buf(h, 13);
    HCHK(ws.ensure(steqr_big_ws_bytes((int)n)));
    bigws = ws.p;
  }
  auto run = [&]() { return bigws ? launch_steqr_big(a, zcomplex != 0, bigws, s) : launch_steqr(a, zcomplex != 0, s); };
  if (dev) {
    a.d = d;
    a.e = e;
    a.Z = wantz ? Z : nullptr;
    a.ldz = ldz;
    a.info = info;
    HCHK(run());
    if (!(flags & GS_ASYNC)) HCHK(hipStreamSynchronize(s));
    return GS_OK;
  }
  const size_t bd = up256((size_t)n * 8), bZ = wantz ? up256((size_t)zrows * n * es) : 0;
  DevBuf& buf = handle_buf(h, 12);
  HCHK(buf.ensure(2 * bd + bZ + 256));
  char* p = static_cast<char*>(buf.p);
  double* dd = reinterpret_cast<double*>(p);
  double* de = reinterpret_cast<double*>(p + bd);
  int32_t* dI = reinterpret_cast<int32_t*>(p + 2 * bd);
  double* dZ = wantz ? reinterpret_cast<double*>(p + 2 * bd + 256) : nullptr;
  HCHK(hipMemcpyAsync(dd, d, (size_t)n * 8, hipMemcpyHostToDevice, s));
  if (n > 1) HCHK(hipMemcpyAsync(de, e, (size_t)(n - 1) * 8, hipMemcpyHostToDevice, s));
  if (wantz) HCHK(copy_cols(dZ, zrows * es, Z, ldz * es, zrows * es, n, hipMemcpyHostToDevice, s));
  a.d = dd;
  a.e = de;
  a.Z = dZ;
  a.ldz = zrows;
  a.info = dI;
  HCHK(run());
  HCHK(hipMemcpyAsync(d, dd, (size_t)n * 8, hipMemcpyDeviceToHost, s));
  if (n > 1) HCHK(hipMemcpyAsync(e, de, (size_t)(n - 1) * 8, hipMemcpyDeviceToHost, s));
  if (wantz) HCHK(copy_cols(Z, ldz * es, dZ, zrows * es, zrows * es, n, hipMemcpyDeviceToHost, s));
  if (info) HCHK(hipMemcpyAsync(info, dI, 4, hipMemcpyDeviceToHost, s));
  HCHK(hipStreamSynchronize(s));
  return GS_OK;
}

}  // extern "C"
