// Host entry point of the eigenvalue VJP (gs_eig_vjp_batched_d): recompute
// right and left eigenvectors from (T, Z) with the eigenvector kernel, then
// assemble Abar.  All three launches run on the handle's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/gschur_mi355x.h"
#include "handle.h"
#include "internal.h"

using namespace gs;

extern "C" int gs_eig_vjp_batched_d(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off,
                                    const double* T, const double* Z, const double* gbar, double* Abar,
                                    int32_t* info, int flags) {
  if (!h || count < 0) return GS_E_ARG;
  if (count == 0) return GS_OK;
  if (!n || !off || !T || !Z || !gbar || !Abar || !info) return handle_fail(h, GS_E_ARG, "null pointer");
  int nmax = 0;
  int64_t span = 0, wtot = 0;
  const int src = check_spans(h, count, n, off, &span);
  if (src != GS_OK) return src;
  std::vector<int64_t> woff(count);
  for (int64_t i = 0; i < count; ++i) {
    nmax = std::max(nmax, (int)n[i]);
    woff[i] = wtot;
    wtot += n[i];
  }
  if (nmax > trevc_big_max_n()) return handle_fail(h, GS_E_UNSUPPORTED, "order above the VJP kernel limit");
  std::lock_guard<std::mutex> lk(handle_mutex(h));
  hipError_t e = hipSetDevice(handle_device(h));
  if (e != hipSuccess) return handle_fail_hip(h, e, "hipSetDevice");
  hipStream_t s = handle_stream(h);
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  std::vector<int32_t> order(count);
  for (int64_t i = 0; i < count; ++i) order[i] = (int32_t)i;
  std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return n[x] > n[y]; });
  DevBuf& meta = handle_buf(h, 4);
  const size_t mbytes = sizeof(int64_t) * 2 * count + sizeof(int32_t) * 2 * count;
  if ((e = meta.ensure(mbytes + 64)) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
  char* mp = static_cast<char*>(meta.p);
  int64_t* d_off = reinterpret_cast<int64_t*>(mp);
  int64_t* d_woff = d_off + count;
  int32_t* d_n = reinterpret_cast<int32_t*>(d_woff + count);
  int32_t* d_order = d_n + count;
  std::vector<char> hm(mbytes);
  std::memcpy(hm.data(), off, sizeof(int64_t) * count);
  std::memcpy(hm.data() + sizeof(int64_t) * count, woff.data(), sizeof(int64_t) * count);
  std::memcpy(hm.data() + sizeof(int64_t) * 2 * count, n, sizeof(int32_t) * count);
  std::memcpy(hm.data() + sizeof(int64_t) * 2 * count + sizeof(int32_t) * count, order.data(),
              sizeof(int32_t) * count);
  if ((e = hipMemcpyAsync(meta.p, hm.data(), mbytes, hipMemcpyHostToDevice, s)) != hipSuccess)
    return handle_fail_hip(h, e, "hipMemcpyAsync");
  DevBuf& bVR = handle_buf(h, 5);
  DevBuf& bVL = handle_buf(h, 6);
  if ((e = bVR.ensure(16 * span)) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
  if ((e = bVL.ensure(16 * span)) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
  const double *dT = T, *dZ = Z, *dG = gbar;
  double* dA = Abar;
  int32_t* dI = info;
  DevBuf& stage = handle_buf(h, 7);
  if (!dev) {
    // one staging block: T, Z, Abar (real, span each), gbar (complex), info
    const size_t bytes = 8 * 3 * (size_t)span + 16 * (size_t)wtot + 4 * (size_t)count + 64;
    if ((e = stage.ensure(bytes)) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
    double* base = static_cast<double*>(stage.p);
    double* sT = base;
    double* sZ = base + span;
    double* sA = base + 2 * span;
    double* sG = base + 3 * span;
    int32_t* sI = reinterpret_cast<int32_t*>(sG + 2 * wtot);
    if ((e = hipMemcpyAsync(sT, T, 8 * span, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(sZ, Z, 8 * span, hipMemcpyHostToDevice, s)) != hipSuccess ||
        (e = hipMemcpyAsync(sG, gbar, 16 * wtot, hipMemcpyHostToDevice, s)) != hipSuccess)
      return handle_fail_hip(h, e, "hipMemcpyAsync");
    dT = sT;
    dZ = sZ;
    dG = sG;
    dA = sA;
    dI = sI;
  }
  TrevcArgs t;
  t.order = d_order;
  t.count = (int)count;
  t.n = d_n;
  t.off = d_off;
  t.T = dT;
  t.Z = dZ;
  t.normalize = 1;
  t.ldt = t.ldz = t.ldv = 0;
  // right and left eigenvectors: matrices above the one-workgroup kernels'
  // orders (first in the descending order) one at a time on the whole-GPU
  // path, the rest in one launch per side
  int64_t nbigv = 0;
  while (nbigv < count && n[order[nbigv]] > dtrevc_max_n()) ++nbigv;
  if (nbigv > 0) {
    DevBuf& ws = handle_buf(h, 13);
    if ((e = ws.ensure(dtrevc_big_ws_bytes(nmax))) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
    for (int64_t b = 0; b < nbigv; ++b) {
      const int32_t m = order[b];
      TrevcArgs one = t;
      one.T = dT + off[m];
      one.Z = dZ + off[m];
      one.ldt = one.ldz = one.ldv = n[m];
      for (int left = 0; left < 2; ++left) {
        one.left = left;
        one.V = static_cast<double*>(left ? bVL.p : bVR.p) + 2 * off[m];
        if ((e = launch_dtrevc_big(one, n[m], ws.p, s)) != hipSuccess)
          return handle_fail_hip(h, e, "eigenvector launch");
      }
    }
  }
  if (nbigv < count) {
    t.order = d_order + nbigv;
    t.count = (int)(count - nbigv);
    const int nsmall = n[order[nbigv]];
    t.V = static_cast<double*>(bVR.p);
    t.left = 0;
    if ((e = launch_dtrevc(t, nsmall, s)) != hipSuccess) return handle_fail_hip(h, e, "eigenvector launch");
    t.V = static_cast<double*>(bVL.p);
    t.left = 1;
    if ((e = launch_dtrevc(t, nsmall, s)) != hipSuccess) return handle_fail_hip(h, e, "eigenvector launch");
  }
  // orders whose two factors fit LDS: one workgroup per matrix; larger
  // ones (first in the descending order): c_k and the planes of each
  // matrix, then Abar = P Q^T as one FP64 MFMA GEMM of n x n x 2n
  int64_t nlarge = 0;
  while (nlarge < count && vjp_lds_bytes(n[order[nlarge]]) > 160 * 1024) ++nlarge;
  if (nlarge < count) {
    VjpArgs v;
    v.order = d_order + nlarge;
    v.count = (int)(count - nlarge);
    v.n = d_n;
    v.off = d_off;
    v.woff = d_woff;
    v.VR = static_cast<const double*>(bVR.p);
    v.VL = static_cast<const double*>(bVL.p);
    v.gbar = dG;
    v.Abar = dA;
    v.info = dI;
    if ((e = launch_vjp_assemble(v, n[order[nlarge]], s)) != hipSuccess) return handle_fail_hip(h, e, "VJP launch");
  }
  if (nlarge > 0) {
    const int nl = n[order[0]];
    DevBuf& planes = handle_buf(h, 12);
    if ((e = planes.ensure(8 * 4 * (size_t)nl * nl)) != hipSuccess) return handle_fail_hip(h, e, "hipMalloc");
    double* P = static_cast<double*>(planes.p);
    double* Q = P + 2 * (size_t)nl * nl;
    const double* vr = static_cast<const double*>(bVR.p);
    const double* vl = static_cast<const double*>(bVL.p);
    for (int64_t q = 0; q < nlarge; ++q) {
      const int32_t mi = order[q];
      const int nm = n[mi];
      if ((e = launch_vjp_info_clear(dI + mi, 1, s)) != hipSuccess ||
          (e = launch_vjp_planes(nm, vr + 2 * off[mi], vl + 2 * off[mi], dG + 2 * woff[mi], P, Q, dI + mi, s)) !=
              hipSuccess ||
          (e = launch_dgemm(false, true, nm, nm, 2 * nm, 1.0, P, nm, Q, nm, 0.0, dA + off[mi], nm, s)) != hipSuccess)
        return handle_fail_hip(h, e, "VJP launch");
    }
  }
  if (!dev) {
    std::vector<double> ha((size_t)span);
    if ((e = hipMemcpyAsync(ha.data(), dA, 8 * span, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = hipMemcpyAsync(info, dI, 4 * count, hipMemcpyDeviceToHost, s)) != hipSuccess)
      return handle_fail_hip(h, e, "hipMemcpyAsync");
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return handle_fail_hip(h, e, "hipStreamSynchronize");
    for (int64_t i = 0; i < count; ++i) std::memcpy(Abar + off[i], ha.data() + off[i], 8 * (size_t)n[i] * n[i]);
    return GS_OK;
  }
  if (!(flags & GS_ASYNC))
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return handle_fail_hip(h, e, "hipStreamSynchronize");
  return GS_OK;
}
