// Host entry points for the eigenvector kernels (gs_trevc_*), see
// include/gschur_mi355x.h.  Shares the handle of gschur_api.cpp through the
// small accessor interface declared in handle.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/gschur_mi355x.h"
#include "handle.h"
#include "internal.h"

using namespace gs;

namespace {

int run_trevc(gs_handle* h, bool cplx, int64_t count, const int32_t* n, const int64_t* off, const double* T,
              int64_t ldt, const double* Z, int64_t ldz, double* V, int64_t ldv, int left, int normalize, int flags,
              bool single) {
  if (!h || count < 0) return GS_E_ARG;
  if (count == 0) return GS_OK;
  if (!n || !off || !T || !V) return handle_fail(h, GS_E_ARG, "null pointer");
  const int src = check_spans(h, count, n, off, nullptr);
  if (src != GS_OK) return src;
  int nmax = 0;
  for (int64_t i = 0; i < count; ++i) nmax = std::max(nmax, (int)n[i]);
  if (nmax == 0) return GS_OK;
  // orders above the one-workgroup kernels' limit run one matrix at a time
  // through the whole-GPU path, up to trevc_big_max_n()
  const int lim = cplx ? ztrevc_max_n() : dtrevc_max_n();
  if (nmax > trevc_big_max_n()) return handle_fail(h, GS_E_UNSUPPORTED, "order above the eigenvector kernel limit");
  std::lock_guard<std::mutex> lk(handle_mutex(h));
  hipError_t e = hipSetDevice(handle_device(h));
  if (e != hipSuccess) return handle_fail_hip(h, e, "hipSetDevice");
  hipStream_t s = handle_stream(h);
  const size_t es = cplx ? 16 : 8;
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  // element spans of T, Z and V (batched: ld = n for all three)
  int64_t spanT = 0, spanZ = 0, spanV = 0;
  for (int64_t i = 0; i < count; ++i) {
    spanT = std::max<int64_t>(spanT, off[i] + (single ? ldt : n[i]) * n[i]);
    spanZ = std::max<int64_t>(spanZ, off[i] + (single ? ldz : n[i]) * n[i]);
    spanV = std::max<int64_t>(spanV, off[i] + (single ? ldv : n[i]) * n[i]);
  }
  const int64_t span = spanV;
  // metadata
  std::vector<int32_t> order(count);
  for (int64_t i = 0; i < count; ++i) order[i] = (int32_t)i;
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return n[x] > n[y]; });
  DevBuf& meta = handle_buf(h, 0);
  const size_t mbytes = sizeof(int32_t) * 2 * count + sizeof(int64_t) * count + 64;
  if ((e = meta.ensure(mbytes)) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
  char* mp = static_cast<char*>(meta.p);
  int64_t* d_off = reinterpret_cast<int64_t*>(mp);
  int32_t* d_n = reinterpret_cast<int32_t*>(mp + sizeof(int64_t) * count);
  int32_t* d_order = d_n + count;
  std::vector<char> hm(mbytes);
  std::memcpy(hm.data(), off, sizeof(int64_t) * count);
  std::memcpy(hm.data() + sizeof(int64_t) * count, n, sizeof(int32_t) * count);
  std::memcpy(hm.data() + sizeof(int64_t) * count + sizeof(int32_t) * count, order.data(), sizeof(int32_t) * count);
  if ((e = hipMemcpyAsync(meta.p, hm.data(), mbytes - 64, hipMemcpyHostToDevice, s)) != hipSuccess)
    return handle_fail_hip(h, e, "hipMemcpyAsync");
  const double *dT = T, *dZ = Z;
  double* dV = V;
  if (!dev) {
    DevBuf& bT = handle_buf(h, 1);
    DevBuf& bZ = handle_buf(h, 2);
    DevBuf& bV = handle_buf(h, 3);
    if ((e = bT.ensure(es * spanT)) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
    if ((e = bV.ensure(16 * spanV)) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
    // a single strided operand occupies ld (n - 1) + n elements of the
    // caller's buffer: copy column by column, never past its last column
    if (single)
      e = copy_cols(bT.p, es * ldt, T, es * ldt, es * n[0], n[0], hipMemcpyHostToDevice, s);
    else
      e = hipMemcpyAsync(bT.p, T, es * spanT, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return handle_fail_hip(h, e, "hipMemcpyAsync");
    dT = static_cast<const double*>(bT.p);
    if (Z) {
      if ((e = bZ.ensure(es * spanZ)) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
      if (single)
        e = copy_cols(bZ.p, es * ldz, Z, es * ldz, es * n[0], n[0], hipMemcpyHostToDevice, s);
      else
        e = hipMemcpyAsync(bZ.p, Z, es * spanZ, hipMemcpyHostToDevice, s);
      if (e != hipSuccess) return handle_fail_hip(h, e, "hipMemcpyAsync");
      dZ = static_cast<const double*>(bZ.p);
    }
    dV = static_cast<double*>(bV.p);
  }
  TrevcArgs a;
  a.order = d_order;
  a.count = (int)count;
  a.n = d_n;
  a.off = d_off;
  a.T = dT;
  a.Z = dZ;
  a.V = dV;
  a.left = left;
  a.normalize = normalize;
  a.ldt = single ? ldt : 0;
  a.ldz = single ? ldz : 0;
  a.ldv = single ? ldv : 0;
  // the order array lists the matrices largest first: those above lim lead
  int64_t nbig = 0;
  while (nbig < count && n[order[nbig]] > lim) ++nbig;
  if (nbig > 0) {
    DevBuf& ws = handle_buf(h, 4);
    const size_t wsb = cplx ? ztrevc_big_ws_bytes(nmax) : dtrevc_big_ws_bytes(nmax);
    if ((e = ws.ensure(wsb)) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
    const int64_t ew = cplx ? 2 : 1;  // doubles per element of T and Z
    for (int64_t b = 0; b < nbig; ++b) {
      const int32_t m = order[b];
      TrevcArgs one = a;
      one.T = dT + ew * off[m];
      one.Z = dZ ? dZ + ew * off[m] : nullptr;
      one.V = dV + 2 * off[m];
      one.ldt = single ? ldt : n[m];
      one.ldz = single ? ldz : n[m];
      one.ldv = single ? ldv : n[m];
      e = cplx ? launch_ztrevc_big(one, n[m], ws.p, s) : launch_dtrevc_big(one, n[m], ws.p, s);
      if (e != hipSuccess) return handle_fail_hip(h, e, "kernel launch");
    }
  }
  if (nbig < count && n[order[nbig]] > 0) {
    a.order = d_order + nbig;
    a.count = (int)(count - nbig);
    e = cplx ? launch_ztrevc(a, n[order[nbig]], s) : launch_dtrevc(a, n[order[nbig]], s);
    if (e != hipSuccess) return handle_fail_hip(h, e, "kernel launch");
  }
  if (!dev) {
    if (single) {
      e = copy_cols(V, 16 * ldv, dV, 16 * ldv, 16 * n[0], n[0], hipMemcpyDeviceToHost, s);
      if (e != hipSuccess) return handle_fail_hip(h, e, "hipMemcpy2DAsync");
    } else {
      std::vector<double> hv(2 * (size_t)span);
      if ((e = hipMemcpyAsync(hv.data(), dV, 16 * span, hipMemcpyDeviceToHost, s)) != hipSuccess)
        return handle_fail_hip(h, e, "hipMemcpyAsync");
      if ((e = hipStreamSynchronize(s)) != hipSuccess) return handle_fail_hip(h, e, "hipStreamSynchronize");
      for (int64_t i = 0; i < count; ++i)
        std::memcpy(V + 2 * off[i], hv.data() + 2 * off[i], 16 * (size_t)n[i] * n[i]);
      return GS_OK;
    }
  }
  if (!(flags & GS_ASYNC) || !dev) {
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return handle_fail_hip(h, e, "hipStreamSynchronize");
  }
  return GS_OK;
}

}  // namespace

extern "C" {

int gs_trevc_z(gs_handle* h, int64_t n, const double* T, int64_t ldt, const double* Z, int64_t ldz, double* V,
               int64_t ldv, int left, int normalize, int flags) {
  if (n < 0 || ldt < std::max<int64_t>(n, 1) || ldv < std::max<int64_t>(n, 1) || (Z && ldz < std::max<int64_t>(n, 1)))
    return GS_E_DIM;
  const int32_t nn = (int32_t)n;
  const int64_t o = 0;
  return run_trevc(h, true, 1, &nn, &o, T, ldt, Z, Z ? ldz : ldt, V, ldv, left, normalize, flags, true);
}

int gs_trevc_d(gs_handle* h, int64_t n, const double* T, int64_t ldt, const double* Z, int64_t ldz, double* V,
               int64_t ldv, int left, int normalize, int flags) {
  if (n < 0 || ldt < std::max<int64_t>(n, 1) || ldv < std::max<int64_t>(n, 1) || (Z && ldz < std::max<int64_t>(n, 1)))
    return GS_E_DIM;
  const int32_t nn = (int32_t)n;
  const int64_t o = 0;
  return run_trevc(h, false, 1, &nn, &o, T, ldt, Z, Z ? ldz : ldt, V, ldv, left, normalize, flags, true);
}

int gs_trevc_batched_z(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, const double* T,
                       const double* Z, double* V, int left, int normalize, int flags) {
  return run_trevc(h, true, count, n, off, T, 0, Z, 0, V, 0, left, normalize, flags, false);
}

int gs_trevc_batched_d(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, const double* T,
                       const double* Z, double* V, int left, int normalize, int flags) {
  return run_trevc(h, false, count, n, off, T, 0, Z, 0, V, 0, left, normalize, flags, false);
}

}  // extern "C"
