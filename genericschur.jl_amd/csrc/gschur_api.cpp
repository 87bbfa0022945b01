// Host implementation of the C ABI in include/gschur_mi355x.h.
//
// The handle owns a HIP stream, grow-only device staging buffers and a
// cache of batch plans.  A batch plan sorts the matrices into size bins
// (largest first inside each bin, so the hardware dispatcher hands the long
// jobs out first), uploads the per-matrix metadata once, and records the
// per-bin launch shape (threads per workgroup, LDS bytes).  Repeated calls
// with the same sizes and offsets reuse the plan and, for device pointers,
// a captured hipGraph of the whole batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gschur_mi355x.h"
#include "handle.h"
#include "internal.h"

using namespace gs;

namespace gs {

struct Bin {
  int first;  // index into the order array
  int count;
  int nmax;
  int nmin;
  int nt_hq;  // threads per workgroup of the Hessenberg+Z launch
  int nt_qr;  // threads per workgroup of the QR launch
  int branch; // graph branch (stream) the bin runs on
  int zlog;   // tuned Z log of the QR launch: 1 on, 0 off, -1 the default rule
  double work;
};

struct Plan {
  bool is_complex = false;
  int64_t count = 0;
  std::vector<int32_t> n;
  std::vector<int64_t> off;
  std::vector<int64_t> woff;
  std::vector<int32_t> order;
  std::vector<Bin> bins;
  int64_t total_elems = 0;  // elements spanned by A/T/Z (max off + n^2)
  int64_t total_w = 0;
  // Z-log offsets (complex entries, count + 1 each): normal and tiny
  // capacities, and their totals
  std::vector<int64_t> loff[2];
  // device metadata
  DevBuf d_order, d_n, d_off, d_woff, d_stats, d_scl, d_loff[2], d_llen, d_qctr;
  // captured graph for device-pointer execution
  hipGraphExec_t gexec = nullptr;
  const void* g_ptrs[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int g_flags = -1;
  ~Plan() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    d_order.release();
    d_n.release();
    d_off.release();
    d_woff.release();
    d_stats.release();
    d_scl.release();
    d_loff[0].release();
    d_loff[1].release();
    d_llen.release();
    d_qctr.release();
  }
};

// Device view of a plan: the metadata arrays a batch's launches read, from
// a cached Plan's buffers or from the no-allocation scratch array.
struct PlanDev {
  bool is_complex = false;
  const std::vector<Bin>* bins = nullptr;
  const int32_t* order = nullptr;
  const int32_t* n = nullptr;
  const int64_t* off = nullptr;
  const int64_t* woff = nullptr;
  int64_t* stats = nullptr;
  double* scl = nullptr;
  const int64_t* loff[2] = {nullptr, nullptr};
  int64_t* llen = nullptr;
  int* qctr = nullptr;  // work-queue counter of the persistent QR launch, or null
};

static PlanDev plan_dev(const Plan* p) {
  PlanDev d;
  d.is_complex = p->is_complex;
  d.bins = &p->bins;
  d.order = static_cast<const int32_t*>(p->d_order.p);
  d.n = static_cast<const int32_t*>(p->d_n.p);
  d.off = static_cast<const int64_t*>(p->d_off.p);
  d.woff = static_cast<const int64_t*>(p->d_woff.p);
  d.stats = static_cast<int64_t*>(p->d_stats.p);
  d.scl = static_cast<double*>(p->d_scl.p);
  d.loff[0] = static_cast<const int64_t*>(p->d_loff[0].p);
  d.loff[1] = static_cast<const int64_t*>(p->d_loff[1].p);
  d.llen = static_cast<int64_t*>(p->d_llen.p);
  d.qctr = static_cast<int*>(p->d_qctr.p);
  return d;
}

}  // namespace gs

static int fail_hip(gs_handle* h, hipError_t e, const char* where) { return handle_fail_hip(h, e, where); }
static int fail(gs_handle* h, int code, const char* msg) { return handle_fail(h, code, msg); }

// GSCHUR_QR_BAND=1 selects the band-resident QR kernels everywhere (A/B
// measurements), as GS_QR_BAND does per call; GSCHUR_QR_NOPRIO=1 keeps the
// chasing wavefront of the windowed kernel at the default priority
// Number of leading bins whose Hessenberg launches precede every other
// launch of a batch: two for complex, one for real batches (GSCHUR_HQ_LEAD
// overrides, 0..3).
static int hq_lead(const Tuning& t, bool cplx) {
  static const int v = [] {
    const char* e = std::getenv("GSCHUR_HQ_LEAD");
    if (!e) return -1;
    const int x = std::atoi(e);
    return x < 0 ? 0 : (x > 3 ? 3 : x);
  }();
  if (v >= 0) return v;
  if (t.lead >= 0) return t.lead;
  return cplx ? 2 : 1;
}

// Complex batches can run their QR iterations as one persistent work-queue
// launch over all size bins (largest matrix first) after the bins'
// Hessenberg launches, instead of one QR launch per bin on the graph's
// branches (measured: 139.0k against 139.4k-139.8k matrices/s per bin on
// the default workload, so the default is per bin).  The tuning table or
// GSCHUR_QR_QUEUE=0/1 (A/B runs, first) choose.
static bool qr_queue(const Tuning& t, bool cplx) {
  static const int v = [] {
    const char* e = std::getenv("GSCHUR_QR_QUEUE");
    return e ? (e[0] == '1' ? 1 : 0) : -1;
  }();
  if (!cplx) return false;
  if (v >= 0) return v == 1;
  return t.queue == 1;
}
// workgroups of the persistent QR launch: eight two-wavefront workgroups
// per CU by default (the occupancy of the per-bin launches)
static int qr_queue_slots(const Tuning& t) {
  static const int cu = [] {
    int dev = 0, x = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&x, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && x > 0)
      return x;
    return 256;
  }();
  static const int env = [] {
    const char* e = std::getenv("GSCHUR_QR_QUEUE_SLOTS");
    return e ? std::max(1, std::min(16, std::atoi(e))) : 0;
  }();
  const int per = env ? env : (t.slots > 0 ? t.slots : 8);
  return cu * per;
}

static int env_qr_band() {
  static const int v = [] {
    const char* e = std::getenv("GSCHUR_QR_BAND");
    const char* p = std::getenv("GSCHUR_QR_NOPRIO");
    return ((e && e[0] == '1') ? GS_QR_BAND : 0) | ((p && p[0] == '1') ? 0x4000 : 0);
  }();
  return v;
}

#define HIPCHK(h, call)                                  \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return fail_hip((h), e_, #call); \
  } while (0)

extern "C" {

int gs_version(void) { return 10000; }

int gs_create(gs_handle** out, int device) {
  if (!out) return GS_E_ARG;
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess) return GS_E_HIP - (int)e;
  if (device < 0 || device >= ndev) return GS_E_ARG;
  e = hipSetDevice(device);
  if (e != hipSuccess) return GS_E_HIP - (int)e;
  gs_handle* h = new gs_handle();
  h->device = device;
  e = hipStreamCreateWithFlags(&h->own_stream, hipStreamNonBlocking);
  h->stream = h->own_stream;
  if (e == hipSuccess) e = hipEventCreate(&h->ev0);
  if (e == hipSuccess) e = hipEventCreate(&h->ev1);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&h->evfork, hipEventDisableTiming);
  for (int i = 0; i < kBranches && e == hipSuccess; ++i) {
    e = hipStreamCreateWithFlags(&h->branch[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->evjoin[i], hipEventDisableTiming);
  }
  for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&h->evlead[i], hipEventDisableTiming);
  // this GPU's scratch array of the no-allocation entry point: its device
  // address is looked up once per device, here, with the device current
  void* scratch_ptr = nullptr;
  if (e == hipSuccess) e = noalloc_scratch(device, &scratch_ptr);
  if (e != hipSuccess) {
    delete h;
    return GS_E_HIP - (int)e;
  }
  *out = h;
  return GS_OK;
}

int gs_destroy(gs_handle* h) {
  if (!h) return GS_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  h->plans.clear();
  h->bA.release();
  h->bT.release();
  h->bZ.release();
  h->bW.release();
  h->bInfo.release();
  h->bWs.release();
  h->bStats.release();
  h->bMeta.release();
  h->bScl.release();
  h->bLog.release();
  for (int i = 0; i < kAuxBufs; ++i) h->aux[i].release();
  for (int i = 0; i < kBranches; ++i) {
    if (h->evjoin[i]) (void)hipEventDestroy(h->evjoin[i]);
    if (h->branch[i]) (void)hipStreamDestroy(h->branch[i]);
  }
  if (h->evfork) (void)hipEventDestroy(h->evfork);
  for (int i = 0; i < 3; ++i)
    if (h->evlead[i]) (void)hipEventDestroy(h->evlead[i]);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  for (hipEvent_t ev : h->tev) (void)hipEventDestroy(ev);
  h->lwork.release();
  if (h->lhost) (void)hipHostFree(h->lhost);
  if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
  delete h;
  return GS_OK;
}

const char* gs_last_error(const gs_handle* h) { return h ? h->err.c_str() : "null handle"; }

int gs_synchronize(gs_handle* h) {
  if (!h) return GS_E_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GS_OK;
}

int gs_set_stream(gs_handle* h, void* stream) {
  if (!h) return GS_E_ARG;
  std::lock_guard<std::mutex> lk(h->mu);
  h->stream = stream ? static_cast<hipStream_t>(stream) : h->own_stream;
  return GS_OK;
}

int gs_batched_max_n(int is_complex) { return is_complex ? zschur_max_lds_n() : dschur_max_lds_n(); }

int gs_set_tuning(gs_handle* h, int is_complex, const int32_t* rows, int nrows) {
  if (!h || nrows < 0 || (nrows > 0 && !rows)) return GS_E_ARG;
  Tuning t;
  for (int i = 0; i < nrows; ++i) {
    const int32_t kind = rows[4 * i], a = rows[4 * i + 1], b = rows[4 * i + 2], v = rows[4 * i + 3];
    if (kind == 0) {
      if (a == 1 && v >= 4 && v <= 64) t.bin_w = v;
      else if (a == 2 && v >= 0 && v <= 3) t.lead = v;
      else if (a == 3 && (v == 0 || v == 1)) t.queue = v;
      else if (a == 4 && v >= 1 && v <= 16) t.slots = v;
      else return handle_fail(h, GS_E_ARG, "bad tuning entry");
    } else if (kind == 1 || kind == 2) {
      if (a < 1 || b < a) return handle_fail(h, GS_E_ARG, "bad tuning range");
      if (kind == 1 && !(v == 0 || v == 64 || v == 256 || v == 512 || v == 1024))
        return handle_fail(h, GS_E_ARG, "bad tuning workgroup size");
      if (kind == 1 && v == 64 && b > 32) return handle_fail(h, GS_E_ARG, "64 threads only up to n = 32");
      if (kind == 2 && !(v == -1 || v == 0 || v == 1)) return handle_fail(h, GS_E_ARG, "bad tuning Z-log entry");
      Tuning::Row* r = nullptr;
      for (Tuning::Row& x : t.rows)
        if (x.nlo == a && x.nhi == b) r = &x;
      if (!r) {
        t.rows.push_back(Tuning::Row{a, b, 0, -1});
        r = &t.rows.back();
      }
      if (kind == 1) r->nt_hq = v;
      else r->zlog = v;
    } else {
      return handle_fail(h, GS_E_ARG, "bad tuning kind");
    }
  }
  std::lock_guard<std::mutex> lk(h->mu);
  (void)hipSetDevice(h->device);
  // cached plans (and their graphs) were built with the old parameters
  if (h->stream) HIPCHK(h, hipStreamSynchronize(h->stream));
  h->plans.clear();
  h->last_plan = nullptr;
  h->tune[is_complex ? 1 : 0] = std::move(t);
  return GS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ plans
// Hessenberg+Q workgroup size above n = 32: 512 threads, except 256 for
// complex bins up to n = 64 (measured, each launch alone: complex n = 81..96
// 5.7 -> 4.6 ms with 512, n = 33..48 0.64 -> 0.77 ms; real n ~ U{16..128}
// in 32-wide bins 143.9k -> 146.2k matrices/s with 512 for the n = 33..64
// bin).  GSCHUR_HQ_NT overrides it above n = 32 (A/B measurements; 256, 512
// or 1024).
static int default_nt_hq(bool cplx, int nmax) {
  static const int force = [] {
    const char* e = std::getenv("GSCHUR_HQ_NT");
    const int v = e ? std::atoi(e) : 0;
    return (v == 256 || v == 512 || v == 1024) ? v : 0;
  }();
  if (nmax <= 32) return 64;
  if (force) return force;
  return (!cplx || nmax > 64) ? 512 : 256;
}

// The tuning table's row for a bin whose largest order is nmax, or null.
static const Tuning::Row* tune_row(const Tuning& t, int nmax) {
  for (const Tuning::Row& r : t.rows)
    if (nmax >= r.nlo && nmax <= r.nhi) return &r;
  return nullptr;
}

// Hessenberg + Q workgroup size of a bin (or a single matrix) of largest
// order nmax: GSCHUR_HQ_NT, then the tuning table, then the default.  The
// single-matrix path uses the same choice, so a matrix gives the same bits
// alone and in a batch.
static int bin_nt_hq(const Tuning& t, bool cplx, int nmax) {
  int nt = default_nt_hq(cplx, nmax);
  const Tuning::Row* r = tune_row(t, nmax);
  if (r && r->nt_hq && !std::getenv("GSCHUR_HQ_NT")) nt = r->nt_hq;
  return nt;
}
static int default_nt_qr(bool cplx, int nmax) {
  (void)cplx;
  (void)nmax;
  return 64;
}

// Size-bin width: 16 orders for complex batches, 32 for real ones (measured
// on n ~ U{16..128} real: 136.4k matrices/s with 16, 144.0k with 32; complex
// n ~ U{8..96}: 16 and 24 within noise, 8, 12 and 32 slower).
// GSCHUR_BIN_W overrides it (A/B measurements, 4 .. 64).
static int bin_width(const Tuning& t, bool cplx) {
  static const int w = [] {
    const char* e = std::getenv("GSCHUR_BIN_W");
    const int v = e ? std::atoi(e) : 0;
    return (v >= 4 && v <= 64) ? v : 0;
  }();
  if (w) return w;
  if (t.bin_w) return t.bin_w;
  return cplx ? 16 : 32;
}

static std::string plan_key(bool cplx, int64_t count, const int32_t* n, const int64_t* off) {
  std::string k;
  k.reserve(16 + (size_t)count * 12);
  k.push_back(cplx ? 'z' : 'd');
  k.append(reinterpret_cast<const char*>(&count), sizeof count);
  k.append(reinterpret_cast<const char*>(n), sizeof(int32_t) * (size_t)count);
  k.append(reinterpret_cast<const char*>(off), sizeof(int64_t) * (size_t)count);
  return k;
}

// Host part of a plan: size bins of width bin_width, the largest bin first
// and the largest matrix first inside a bin (order[]), each bin's launch
// shape and graph branch, and the eigenvalue offsets woff[].  Orders must
// already be validated.
static void plan_bins(const Tuning& tn, bool cplx, int64_t count, const int32_t* n, std::vector<int32_t>& order,
                      std::vector<Bin>& bins,
                      std::vector<int64_t>& woff) {
  const int nmax_lds = cplx ? zschur_max_lds_n() : dschur_max_lds_n();
  order.clear();
  bins.clear();
  woff.resize(count);
  int64_t w = 0;
  for (int64_t i = 0; i < count; ++i) {
    woff[i] = w;
    w += n[i];
  }
  const int bw = bin_width(tn, cplx);
  const int nb = (nmax_lds + bw - 1) / bw + 1;
  std::vector<std::vector<int32_t>> byb(nb);
  for (int64_t i = 0; i < count; ++i) {
    const int b = n[i] <= 0 ? 0 : (n[i] - 1) / bw;
    byb[b].push_back((int32_t)i);
  }
  for (int b = nb - 1; b >= 0; --b) {
    auto& v = byb[b];
    if (v.empty()) continue;
    std::stable_sort(v.begin(), v.end(), [&](int32_t x, int32_t y) { return n[x] > n[y]; });
    Bin bin;
    bin.first = (int)order.size();
    bin.count = (int)v.size();
    bin.nmax = std::max(1, n[v.front()]);
    bin.nmin = n[v.back()];
    bin.nt_hq = bin_nt_hq(tn, cplx, bin.nmax);
    const Tuning::Row* r = tune_row(tn, bin.nmax);
    bin.zlog = r ? r->zlog : -1;
    bin.nt_qr = default_nt_qr(cplx, bin.nmax);
    bin.work = 0;
    for (int32_t i : v) bin.work += (double)n[i] * n[i] * n[i];
    bin.branch = 0;
    bins.push_back(bin);
    order.insert(order.end(), v.begin(), v.end());
  }
  // bins to graph branches, heaviest first onto the least loaded branch
  std::vector<int> idx(bins.size());
  for (size_t i = 0; i < idx.size(); ++i) idx[i] = (int)i;
  std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return bins[x].work > bins[y].work; });
  double load[kBranches] = {0, 0, 0, 0};
  for (int i : idx) {
    int best = 0;
    for (int b = 1; b < kBranches; ++b)
      if (load[b] < load[best]) best = b;
    bins[i].branch = best;
    load[best] += bins[i].work;
  }
}

static int check_orders(gs_handle* h, bool cplx, int64_t count, const int32_t* n) {
  const int nmax_lds = cplx ? zschur_max_lds_n() : dschur_max_lds_n();
  for (int64_t i = 0; i < count; ++i)
    if (n[i] < 0 || n[i] > nmax_lds) return fail(h, GS_E_DIM, "matrix order outside the batched range");
  return GS_OK;
}

static int build_plan(gs_handle* h, bool cplx, int64_t count, const int32_t* n, const int64_t* off, Plan** out) {
  const std::string key = plan_key(cplx, count, n, off);
  auto it = h->plans.find(key);
  if (it != h->plans.end()) {
    *out = it->second.get();
    return GS_OK;
  }
  int rc = check_orders(h, cplx, count, n);
  if (rc != GS_OK) return rc;
  int64_t span = 0;
  rc = check_spans(h, count, n, off, &span);
  if (rc != GS_OK) return rc;
  std::unique_ptr<Plan> p(new Plan());
  p->is_complex = cplx;
  p->count = count;
  p->n.assign(n, n + count);
  p->off.assign(off, off + count);
  plan_bins(h->tune[cplx ? 1 : 0], cplx, count, n, p->order, p->bins, p->woff);
  p->total_w = count ? p->woff[count - 1] + n[count - 1] : 0;
  p->total_elems = span;
  {
    // Z-log regions: complex entries for complex input, doubles for real
    for (int t = 0; t < 2; ++t) {
      p->loff[t].assign(count + 1, 0);
      for (int64_t i = 0; i < count; ++i)
        p->loff[t][i + 1] =
            p->loff[t][i] + (n[i] > 1 ? (cplx ? zlog_capacity(n[i], t == 1) : dlog_capacity(n[i], t == 1)) : 0);
      HIPCHK(h, p->d_loff[t].ensure(sizeof(int64_t) * (size_t)(count + 1)));
      HIPCHK(h, hipMemcpy(p->d_loff[t].p, p->loff[t].data(), sizeof(int64_t) * (size_t)(count + 1),
                          hipMemcpyHostToDevice));
    }
    HIPCHK(h, p->d_llen.ensure(sizeof(int64_t) * std::max<int64_t>(count, 1)));
  }
  HIPCHK(h, p->d_qctr.ensure(256));
  HIPCHK(h, p->d_order.ensure(sizeof(int32_t) * std::max<int64_t>(count, 1)));
  HIPCHK(h, p->d_n.ensure(sizeof(int32_t) * std::max<int64_t>(count, 1)));
  HIPCHK(h, p->d_off.ensure(sizeof(int64_t) * std::max<int64_t>(count, 1)));
  HIPCHK(h, p->d_woff.ensure(sizeof(int64_t) * std::max<int64_t>(count, 1)));
  HIPCHK(h, p->d_stats.ensure(sizeof(int64_t) * 2 * std::max<int64_t>(count, 1)));
  HIPCHK(h, p->d_scl.ensure(sizeof(double) * 2 * std::max<int64_t>(count, 1)));
  if (count > 0) {
    HIPCHK(h, hipMemcpy(p->d_order.p, p->order.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(p->d_n.p, p->n.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(p->d_off.p, p->off.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice));
    HIPCHK(h, hipMemcpy(p->d_woff.p, p->woff.data(), sizeof(int64_t) * count, hipMemcpyHostToDevice));
  }
  Plan* raw = p.get();
  if (h->plans.size() > 64) h->plans.clear();
  h->plans[key] = std::move(p);
  *out = raw;
  return GS_OK;
}

// Bin orders whose QR launch logs its Z updates.  The smallest: measured on
// MI355X with each launch alone, a bin's QR + replay against in-place
// updates: n = 81..96 (1470 matrices) 25.1 ms vs 32.4 ms, n = 65..80 (1479)
// 16.7 ms vs 15.0 ms.  The largest real one: the replay holds all of Z in
// LDS with its record buffers (128^2 doubles + 6 KB).
constexpr int kZlogMinN = 81;
constexpr int kDlogMinN = 33;
constexpr int kDlogMaxN = 128;

// Whether the batched QR launches of this call log their Z updates
// (windowed kernels, Z wanted, not GS_ZDIRECT).
static bool uses_zlog(int kflags) { return (kflags & GS_WANTZ) && !(kflags & (GS_QR_BAND | GS_ZDIRECT)); }
// Whether bin b's QR launch logs its Z updates.
static bool bin_zlog(bool cplx, const Bin& b, int kflags) {
  if (!uses_zlog(kflags)) return false;
  if (!cplx && b.nmax > kDlogMaxN) return false;
  if (kflags & GS_ZLOG_TINY) return true;
  if (b.zlog >= 0) return b.zlog == 1;
  return b.nmax >= (cplx ? kZlogMinN : kDlogMinN);
}

// Work-queue form of a complex batch (tuning queue = 1): the bins'
// Hessenberg launches (the leading bins' first, alone, on branch 0, the
// others on branches 1 .. 3), then one persistent QR launch over the whole
// batch, largest matrix first, and the Z-log replays.  Every matrix of order
// >= zmin logs its Z updates, zmin the smallest order of the logging bins,
// so that the replays of the bins at or above zmin cover exactly the
// logging matrices.  (Measured: a second QR launch for the leading bins,
// started beside the other bins' Hessenberg launches, ran slower, 57.4
// against 56.4 ms per batch: the smaller matrices' workgroups then took
// slots before the largest matrices had started.)
static int queue_zmin(const std::vector<Bin>& bins, bool zl, bool cplx, int flags) {
  int zmin = 1 << 30;
  for (const Bin& b : bins)
    if (zl && bin_zlog(cplx, b, flags)) zmin = std::min(zmin, b.nmin);
  if ((flags & GS_ZLOG_TINY) && zmin < (1 << 30)) zmin = 0;
  return zmin;
}
// events of the QR launch in the timed pass (after the bins' six each)
static size_t queue_event(size_t nbins) { return 6 * nbins; }

static hipError_t launch_plan_queue(gs_handle* h, const PlanDev& pd, BatchArgs a, int flags, double* zlog, bool timed,
                                    const Tuning& tn) {
  const std::vector<Bin>& bins = *pd.bins;
  const int nb = (int)bins.size();
  const int lead = std::min(hq_lead(tn, true), nb);
  const int zmin = queue_zmin(bins, zlog != nullptr, true, flags);
  hipError_t e;
  auto st = [&](int i) { return timed ? h->stream : h->branch[i]; };
  if (!timed) {
    if ((e = hipEventRecord(h->evfork, h->stream)) != hipSuccess) return e;
    for (int i = 0; i < kBranches; ++i)
      if ((e = hipStreamWaitEvent(h->branch[i], h->evfork, 0)) != hipSuccess) return e;
  }
  auto hq = [&](int bi, hipStream_t s) -> hipError_t {
    const Bin& b = bins[bi];
    a.first = b.first;
    a.count = b.count;
    hipError_t e2;
    if (timed && (e2 = hipEventRecord(h->tev[6 * bi], s)) != hipSuccess) return e2;
    if ((e2 = launch_zhq(a, b.nmax, b.nt_hq, s)) != hipSuccess) return e2;
    if (timed && (e2 = hipEventRecord(h->tev[6 * bi + 1], s)) != hipSuccess) return e2;
    return hipSuccess;
  };
  for (int bi = 0; bi < lead; ++bi)
    if ((e = hq(bi, st(0))) != hipSuccess) return e;
  if (!timed && lead > 0 && (e = hipEventRecord(h->evlead[0], st(0))) != hipSuccess) return e;
  bool usedb[kBranches] = {true, false, false, false};
  for (int bi = lead; bi < nb; ++bi) {
    const int br = 1 + (bi - lead) % (kBranches - 1);
    if (!timed && !usedb[br] && lead > 0 && (e = hipStreamWaitEvent(st(br), h->evlead[0], 0)) != hipSuccess)
      return e;
    usedb[br] = true;
    if ((e = hq(bi, st(br))) != hipSuccess) return e;
  }
  // every Hessenberg launch done: the QR launch and the replays on h->stream
  if (!timed)
    for (int i = 0; i < kBranches; ++i) {
      if ((e = hipEventRecord(h->evjoin[i], h->branch[i])) != hipSuccess) return e;
      if ((e = hipStreamWaitEvent(h->stream, h->evjoin[i], 0)) != hipSuccess) return e;
    }
  int64_t tot = 0;
  for (const Bin& b : bins) tot += b.count;
  if (tot == 0) return hipSuccess;
  a.first = 0;
  a.count = (int)tot;
  a.zlog = zmin < (1 << 30) ? zlog : nullptr;
  a.zlog_min = zmin;
  a.qctr = pd.qctr;
  if ((e = launch_fill_i32(a.qctr, 1, h->stream)) != hipSuccess) return e;
  if (timed && (e = hipEventRecord(h->tev[queue_event(nb)], h->stream)) != hipSuccess) return e;
  if ((e = launch_zqrw_queue(a, qr_queue_slots(tn), h->stream)) != hipSuccess) return e;
  if (timed && (e = hipEventRecord(h->tev[queue_event(nb) + 1], h->stream)) != hipSuccess) return e;
  a.qctr = nullptr;
  a.zlog_min = 0;
  for (int bi = 0; bi < nb; ++bi) {
    const Bin& b = bins[bi];
    if (!(zmin < (1 << 30) && b.nmin >= zmin)) continue;
    a.first = b.first;
    a.count = b.count;
    a.zlog = zlog;
    if (timed && (e = hipEventRecord(h->tev[6 * bi + 4], h->stream)) != hipSuccess) return e;
    if ((e = launch_zlog(a, b.nmax, h->stream)) != hipSuccess) return e;
    if (timed && (e = hipEventRecord(h->tev[6 * bi + 5], h->stream)) != hipSuccess) return e;
  }
  return hipSuccess;
}

// With timed set, every launch runs alone, in order on the handle's stream,
// between two events (eager launches only, never under capture): events
// 6b .. 6b+5 for bin b (Hessenberg, QR, Z log), so each record is the
// launch's own duration.  zlog: the Z-log buffer when uses_zlog.
static hipError_t launch_plan(gs_handle* h, const PlanDev& pd, const double* A, double* T, double* Z, double* W,
                              int32_t* info, int flags, double* zlog, bool timed = false) {
  const PlanDev* p = &pd;
  const std::vector<Bin>& bins = *pd.bins;
  BatchArgs a;
  a.order = pd.order;
  a.n = pd.n;
  a.off = pd.off;
  a.woff = pd.woff;
  a.A = A;
  a.T = T;
  a.Z = Z;
  a.W = W;
  a.info = info;
  a.stats = pd.stats;
  a.scl = pd.scl;
  a.flags = flags;
  a.maxiter = 0;
  a.maxinner = 0;
  a.zlog = nullptr;
  a.loff = nullptr;
  a.llen = nullptr;
  const bool zl = zlog && uses_zlog(flags);
  if (zl) {
    a.loff = pd.loff[(flags & GS_ZLOG_TINY) ? 1 : 0];
    a.llen = pd.llen;
  }
  // fork: every bin runs its Hessenberg launch then its QR launch on one of
  // kBranches streams, so the LDS left free by a large bin's workgroups is
  // filled by the smaller bins' workgroups
  const Tuning& tn = h->tune[p->is_complex ? 1 : 0];
  if (pd.qctr && !(flags & GS_QR_BAND) && qr_queue(tn, p->is_complex))
    return launch_plan_queue(h, pd, a, flags, zl ? zlog : nullptr, timed, tn);
  hipError_t e = hipEventRecord(h->evfork, h->stream);
  if (e != hipSuccess) return e;
  bool used[kBranches] = {false, false, false, false};
  for (const Bin& b : bins) used[b.branch] = true;
  for (int i = 0; i < kBranches; ++i)
    if (used[i] && (e = hipStreamWaitEvent(h->branch[i], h->evfork, 0)) != hipSuccess) return e;
  // the leading (largest) bins' Hessenberg launches run first, alone: the
  // largest bin's chain (Hessenberg, QR, Z log) is the batch's longest, and
  // started later it would finish last with the GPU half idle
  const int lead = (int)std::min<size_t>((size_t)hq_lead(tn, p->is_complex), bins.size());
  auto bstream = [&](const Bin& b) { return timed ? h->stream : h->branch[b.branch]; };
  auto launch_hq = [&](size_t bi) -> hipError_t {
    const Bin& b = bins[bi];
    a.first = b.first;
    a.count = b.count;
    hipStream_t s = bstream(b);
    hipError_t e2;
    if (timed && (e2 = hipEventRecord(h->tev[6 * bi], s)) != hipSuccess) return e2;
    e2 = p->is_complex ? launch_zhq(a, b.nmax, b.nt_hq, s) : launch_dhq(a, b.nmax, b.nt_hq, s);
    if (e2 != hipSuccess) return e2;
    if (timed && (e2 = hipEventRecord(h->tev[6 * bi + 1], s)) != hipSuccess) return e2;
    return hipSuccess;
  };
  for (int bi = 0; bi < lead; ++bi) {
    if ((e = launch_hq(bi)) != hipSuccess) return e;
    if ((e = hipEventRecord(h->evlead[bi], bstream(bins[bi]))) != hipSuccess) return e;
  }
  bool waited[kBranches] = {false, false, false, false};
  for (size_t bi = 0; bi < bins.size(); ++bi) {
    const Bin& b = bins[bi];
    hipStream_t s = bstream(b);
    if ((int)bi >= lead) {
      if (!waited[b.branch]) {
        for (int j = 0; j < lead; ++j)
          if (bins[j].branch != b.branch && (e = hipStreamWaitEvent(s, h->evlead[j], 0)) != hipSuccess) return e;
        waited[b.branch] = true;
      }
      if ((e = launch_hq(bi)) != hipSuccess) return e;
    }
    a.first = b.first;
    a.count = b.count;
    // the Z log pays off where Z's rows are long (its replay holds all of Z
    // in LDS); GS_ZLOG_TINY exercises it on every bin
    a.zlog = (zl && bin_zlog(p->is_complex, b, flags)) ? zlog : nullptr;
    if (timed && (e = hipEventRecord(h->tev[6 * bi + 2], s)) != hipSuccess) return e;
    if (flags & GS_QR_BAND)
      e = p->is_complex ? launch_zqr(a, b.nmax, b.nt_qr, s) : launch_dqr(a, b.nmax, b.nt_qr, s);
    else
      e = p->is_complex ? launch_zqrw(a, s) : launch_dqrw(a, s);
    if (e != hipSuccess) return e;
    if (timed && (e = hipEventRecord(h->tev[6 * bi + 3], s)) != hipSuccess) return e;
    if (timed && (e = hipEventRecord(h->tev[6 * bi + 4], s)) != hipSuccess) return e;
    if (a.zlog && (e = p->is_complex ? launch_zlog(a, b.nmax, s) : launch_dlog(a, b.nmax, s)) != hipSuccess) return e;
    if (timed && (e = hipEventRecord(h->tev[6 * bi + 5], s)) != hipSuccess) return e;
  }
  for (int i = 0; i < kBranches; ++i) {
    if (!used[i]) continue;
    if ((e = hipEventRecord(h->evjoin[i], h->branch[i])) != hipSuccess) return e;
    if ((e = hipStreamWaitEvent(h->stream, h->evjoin[i], 0)) != hipSuccess) return e;
  }
  return hipSuccess;
}

// One eager execution of the plan with an event pair around every launch,
// then the per-launch records for gs_last_launches.
static int run_timed(gs_handle* h, Plan* p, const double* A, double* T, double* Z, double* W, int32_t* info,
                     int kflags, double* zlog) {
  const size_t need = 6 * p->bins.size() + 2;
  while (h->tev.size() < need) {
    hipEvent_t ev = nullptr;
    HIPCHK(h, hipEventCreate(&ev));
    h->tev.push_back(ev);
  }
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  hipError_t le = launch_plan(h, plan_dev(p), A, T, Z, W, info, kflags, zlog, true);
  if (le != hipSuccess) return fail_hip(h, le, "kernel launch (timed)");
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->launch_recs.clear();
  // the work-queue form: one QR record for the whole batch, and the Z-log
  // replays of the bins at or above the smallest logging order
  const Tuning& tn = h->tune[p->is_complex ? 1 : 0];
  const bool queue = !(kflags & GS_QR_BAND) && qr_queue(tn, p->is_complex);
  const int zmin = queue ? queue_zmin(p->bins, zlog && uses_zlog(kflags), p->is_complex, kflags) : (1 << 30);
  auto put = [&](int kind, int nmin, int nmax, int64_t cnt, double work, hipEvent_t e0, hipEvent_t e1) -> int {
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, e0, e1));
    const double rec[GS_LAUNCH_REC] = {(double)kind, (double)nmin, (double)nmax, (double)cnt, work,
                                       1000.0 * (double)ms};
    h->launch_recs.insert(h->launch_recs.end(), rec, rec + GS_LAUNCH_REC);
    return GS_OK;
  };
  int rc = GS_OK;
  int nmin = 1 << 30, nmax = 0;
  int64_t cnt = 0;
  double work = 0;
  for (size_t bi = 0; bi < p->bins.size() && rc == GS_OK; ++bi) {
    const Bin& b = p->bins[bi];
    nmin = std::min(nmin, b.nmin);
    nmax = std::max(nmax, b.nmax);
    cnt += b.count;
    work += b.work;
    // kinds 0 (Hessenberg + Q), 1 (QR sweeps), 4 (Z log applied)
    rc = put(0, b.nmin, b.nmax, b.count, b.work, h->tev[6 * bi], h->tev[6 * bi + 1]);
    if (rc == GS_OK && !queue) rc = put(1, b.nmin, b.nmax, b.count, b.work, h->tev[6 * bi + 2], h->tev[6 * bi + 3]);
    const bool zb = queue ? (zmin < (1 << 30) && b.nmin >= zmin) : (zlog && bin_zlog(p->is_complex, b, kflags));
    if (rc == GS_OK && zb) rc = put(4, b.nmin, b.nmax, b.count, b.work, h->tev[6 * bi + 4], h->tev[6 * bi + 5]);
  }
  if (rc == GS_OK && queue && cnt > 0) {
    const size_t qev = queue_event(p->bins.size());
    rc = put(1, nmin, nmax, cnt, work, h->tev[qev], h->tev[qev + 1]);
  }
  return rc;
}

static int run_batched(gs_handle* h, bool cplx, int64_t count, const int32_t* n, const int64_t* off,
                       const double* A, double* T, double* Z, double* W, int32_t* info, int flags) {
  if (!h || count < 0 || (count > 0 && (!n || !off || !A || !T || !W || !info))) return fail(h, GS_E_ARG, "bad argument");
  if ((flags & GS_WANTZ) && count > 0 && !Z) return fail(h, GS_E_ARG, "Z required with GS_WANTZ");
  if (flags & GS_HESSENBERG) return fail(h, GS_E_UNSUPPORTED, "GS_HESSENBERG is single-matrix only");
  std::lock_guard<std::mutex> lk(h->mu);
  HIPCHK(h, hipSetDevice(h->device));
  Plan* p = nullptr;
  int rc = build_plan(h, cplx, count, n, off, &p);
  if (rc != GS_OK) return rc;
  h->last_plan = p;
  h->last_stats_from_plan = true;
  h->launch_recs.clear();
  if (count == 0) return GS_OK;
  // GSCHUR_REFERENCE_SWEEP=1 selects the sweep written in the reference's
  // form (LDS round trips, full divisions) for A/B measurements
  static const bool ref_sweep = [] {
    const char* v = std::getenv("GSCHUR_REFERENCE_SWEEP");
    return v && v[0] == '1';
  }();
  const int kflags = (flags & (GS_WANTZ | GS_SCALE | GS_QR_BAND | GS_ZDIRECT | GS_ZLOG_TINY)) |
                     (ref_sweep ? 0x1000 : 0) | env_qr_band();
  const size_t es = cplx ? 16 : 8;
  double* zlog = nullptr;
  if (uses_zlog(kflags)) {
    const int64_t entries = p->loff[(kflags & GS_ZLOG_TINY) ? 1 : 0][count];
    HIPCHK(h, h->bLog.ensure(es * (size_t)std::max<int64_t>(entries, 1)));
    zlog = static_cast<double*>(h->bLog.p);
  }
  double* dZdummy = nullptr;
  if (flags & GS_DEVICE_PTRS) {
    if (!(flags & GS_WANTZ)) {
      HIPCHK(h, h->bZ.ensure(es * (size_t)p->total_elems));
      dZdummy = static_cast<double*>(h->bZ.p);
    }
    double* Zd = (flags & GS_WANTZ) ? Z : dZdummy;
    if (flags & GS_TIMING) return run_timed(h, p, A, T, Zd, W, info, kflags, zlog);
    const void* ptrs[6] = {A, T, Zd, W, info, zlog};
    if (!p->gexec || std::memcmp(ptrs, p->g_ptrs, sizeof ptrs) != 0 || p->g_flags != kflags) {
      if (p->gexec) {
        (void)hipGraphExecDestroy(p->gexec);
        p->gexec = nullptr;
      }
      hipGraph_t g = nullptr;
      HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
      hipError_t le = launch_plan(h, plan_dev(p), A, T, Zd, W, info, kflags, zlog);
      hipError_t ce = hipStreamEndCapture(h->stream, &g);
      if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail_hip(h, le, "kernel launch (capture)");
      }
      if (ce != hipSuccess) return fail_hip(h, ce, "hipStreamEndCapture");
      hipError_t ie = hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ie != hipSuccess) {
        p->gexec = nullptr;
        return fail_hip(h, ie, "hipGraphInstantiate");
      }
      std::memcpy(p->g_ptrs, ptrs, sizeof ptrs);
      p->g_flags = kflags;
    }
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    HIPCHK(h, hipGraphLaunch(p->gexec, h->stream));
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    if (!(flags & GS_ASYNC)) HIPCHK(h, hipStreamSynchronize(h->stream));
    return GS_OK;
  }
  // host pointers: stage through device buffers
  const size_t nel = (size_t)p->total_elems;
  HIPCHK(h, h->bA.ensure(es * nel));
  HIPCHK(h, h->bT.ensure(es * nel));
  HIPCHK(h, h->bZ.ensure(es * nel));
  HIPCHK(h, h->bW.ensure(16 * (size_t)std::max<int64_t>(p->total_w, 1)));
  HIPCHK(h, h->bInfo.ensure(sizeof(int32_t) * (size_t)count));
  HIPCHK(h, hipMemcpyAsync(h->bA.p, A, es * nel, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  hipError_t le = launch_plan(h, plan_dev(p), static_cast<double*>(h->bA.p), static_cast<double*>(h->bT.p),
                              static_cast<double*>(h->bZ.p), static_cast<double*>(h->bW.p),
                              static_cast<int32_t*>(h->bInfo.p), kflags, zlog);
  if (le != hipSuccess) return fail_hip(h, le, "kernel launch");
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  // copy back only the matrices' own elements (leave caller memory between
  // matrices untouched)
  std::vector<double> hostT(nel * (es / 8)), hostZ;
  HIPCHK(h, hipMemcpyAsync(hostT.data(), h->bT.p, es * nel, hipMemcpyDeviceToHost, h->stream));
  if (flags & GS_WANTZ) {
    hostZ.resize(nel * (es / 8));
    HIPCHK(h, hipMemcpyAsync(hostZ.data(), h->bZ.p, es * nel, hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(h, hipMemcpyAsync(W, h->bW.p, 16 * (size_t)p->total_w, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(info, h->bInfo.p, sizeof(int32_t) * count, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  const size_t ed = es / 8;
  for (int64_t i = 0; i < count; ++i) {
    const size_t o = (size_t)off[i] * ed, len = (size_t)n[i] * n[i] * ed;
    std::memcpy(T + o, hostT.data() + o, 8 * len);
    if (flags & GS_WANTZ) std::memcpy(Z + o, hostZ.data() + o, 8 * len);
  }
  return GS_OK;
}

extern "C" int gs_schur_batched_z(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, const double* A,
                                  double* T, double* Z, double* W, int32_t* info, int flags) {
  return run_batched(h, true, count, n, off, A, T, Z, W, info, flags);
}

extern "C" int gs_schur_batched_d(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, const double* A,
                                  double* T, double* Z, double* W, int32_t* info, int flags) {
  return run_batched(h, false, count, n, off, A, T, Z, W, info, flags);
}

// The batched complex path without device allocation: the per-matrix
// metadata and the scale factors live in the GPU's fixed scratch array, Z is
// updated in place during the sweeps (no Z log), and the launches run
// eagerly on the handle's stream.  Synchronous: the scratch array and the
// handle's host staging block are reused by the next call.
extern "C" int gs_schur_batched_z_noalloc(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off,
                                          const double* A, double* T, double* Z, double* W, int32_t* info,
                                          int flags) {
  if (!h || count < 0 || (count > 0 && (!n || !off || !A || !T || !Z || !W || !info)))
    return fail(h, GS_E_ARG, "bad argument");
  if (count > kNoAllocMaxCount)
    return fail(h, GS_E_ARG, "gs_schur_batched_z_noalloc: more than 16384 matrices in one call");
  std::lock_guard<std::mutex> lk(h->mu);
  HIPCHK(h, hipSetDevice(h->device));
  int rc = check_orders(h, true, count, n);
  if (rc != GS_OK) return rc;
  rc = check_spans(h, count, n, off, nullptr);
  if (rc != GS_OK) return rc;
  h->last_plan = nullptr;
  h->last_stats_from_plan = false;
  h->launch_recs.clear();
  if (count == 0) return GS_OK;
  void* base = nullptr;
  HIPCHK(h, noalloc_scratch(h->device, &base));
  std::vector<int32_t> order;
  std::vector<Bin> bins;
  std::vector<int64_t> woff;
  plan_bins(h->tune[1], true, count, n, order, bins, woff);
  // scratch layout: order, n, off, woff (uploaded), stats, scl (written by
  // the kernels before they are read), each segment 256-byte aligned
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  const size_t c = (size_t)count;
  const size_t o_n = al(4 * c), o_off = o_n + al(4 * c), o_woff = o_off + al(8 * c), o_stats = o_woff + al(8 * c);
  const size_t o_scl = o_stats + al(16 * c), end = o_scl + al(16 * c);
  if (end > kNoAllocScratchBytes) return fail(h, GS_E_ARG, "gs_schur_batched_z_noalloc: scratch too small");
  std::vector<unsigned char>& hs = h->noalloc_host;
  hs.resize(o_stats);
  std::memcpy(hs.data(), order.data(), 4 * c);
  std::memcpy(hs.data() + o_n, n, 4 * c);
  std::memcpy(hs.data() + o_off, off, 8 * c);
  std::memcpy(hs.data() + o_woff, woff.data(), 8 * c);
  unsigned char* b = static_cast<unsigned char*>(base);
  HIPCHK(h, hipMemcpyAsync(b, hs.data(), o_stats, hipMemcpyHostToDevice, h->stream));
  PlanDev pd;
  pd.is_complex = true;
  pd.bins = &bins;
  pd.order = reinterpret_cast<const int32_t*>(b);
  pd.n = reinterpret_cast<const int32_t*>(b + o_n);
  pd.off = reinterpret_cast<const int64_t*>(b + o_off);
  pd.woff = reinterpret_cast<const int64_t*>(b + o_woff);
  pd.stats = reinterpret_cast<int64_t*>(b + o_stats);
  pd.scl = reinterpret_cast<double*>(b + o_scl);
  const int kflags = GS_WANTZ | (flags & GS_SCALE) | GS_ZDIRECT;
  const hipError_t le = launch_plan(h, pd, A, T, Z, W, info, kflags, nullptr);
  if (le != hipSuccess) return fail_hip(h, le, "kernel launch");
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GS_OK;
}

// ------------------------------------------------------------ single matrix
// The large real path (dlarge.hip) with its workspace; optional per-phase
// and per-GEMM event records for gs_last_launches.
static hipError_t run_large(gs_handle* h, int64_t n, double* dA, int64_t lda, double* dT, int64_t ldt, double* dZ,
                            int64_t ldz, double* dW, int32_t* dInfo, int kflags, int64_t maxiter, bool timing) {
  LargeWorkSizes sz;
  const size_t nd = large_work_doubles(n, &sz);
  hipError_t e = h->lwork.ensure(8 * nd);
  if (e != hipSuccess) return e;
  if (!h->lhost) {
    e = hipHostMalloc(&h->lhost, 1024, hipHostMallocDefault);
    if (e != hipSuccess) {
      h->lhost = nullptr;
      return e;
    }
  }
  double* base = static_cast<double*>(h->lwork.p);
  LargeWork w;
  // segment sizes are multiples of 32 doubles and sum to nd
  double* const end = base + nd;
  auto carve = [&](int64_t cnt) {
    double* q = base;
    base += cnt;
    return q;
  };
  w.Y = carve(sz.Y);
  w.V = carve(sz.V);
  w.Ytmp = carve(sz.Ytmp);
  w.Tall = carve(sz.Tall);
  w.Wb = carve(sz.Wb);
  w.W2 = carve(sz.W2);
  w.tau = carve(sz.tau);
  w.tv = carve(sz.tv);
  w.part = carve(sz.part);
  w.scl = carve(sz.scl);
  w.lg = carve(sz.lg);
  w.st = carve(sz.st);
  w.gk = carve(sz.gk);
  if (base != end) return hipErrorInvalidValue;
  w.host_state = h->lhost;
  LargeArgs a;
  a.n = n;
  a.A = dA;
  a.lda = lda;
  a.T = dT;
  a.ldt = ldt;
  a.Z = dZ;
  a.ldz = ldz;
  a.W = dW;
  a.info = dInfo;
  a.stats = static_cast<int64_t*>(h->bStats.p);
  a.flags = kflags;
  a.maxiter = maxiter;
  LargeTimer tm;
  tm.pool = &h->tev;
  e = large_schur_d(a, w, h->stream, timing ? &tm : nullptr);
  if (e != hipSuccess) {
    h->large_where = w.where;
    return e;
  }
  if (!timing) return e;
  if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return e;
  for (const LargeTimer::Rec& r : tm.recs) {
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, h->tev[r.e0], h->tev[r.e1])) != hipSuccess) return e;
    const double rec[GS_LAUNCH_REC] = {(double)r.kind, (double)r.a, (double)r.b, (double)r.c, r.flops,
                                       1000.0 * (double)ms};
    h->launch_recs.insert(h->launch_recs.end(), rec, rec + GS_LAUNCH_REC);
  }
  for (int ph = 0; ph < 5; ++ph) {
    float ms = 0.f;
    if ((e = hipEventElapsedTime(&ms, h->tev[tm.phase_ev[ph]], h->tev[tm.phase_ev[ph + 1]])) != hipSuccess) return e;
    const double rec[GS_LAUNCH_REC] = {3.0, (double)ph, 0.0, 0.0, 0.0, 1000.0 * (double)ms};
    h->launch_recs.insert(h->launch_recs.end(), rec, rec + GS_LAUNCH_REC);
  }
  return hipSuccess;
}

// Threads of the windowed QR kernel for one matrix above the LDS-resident
// orders.
static int win_nt(int n) {
  static const int force = [] {
    const char* e = std::getenv("GSCHUR_WIN_NT");
    const int v = e ? std::atoi(e) : 0;
    return (v == 128 || v == 256 || v == 512 || v == 1024) ? v : 0;
  }();
  if (force) return force;
  return n <= 128 ? 256 : (n <= 256 ? 512 : 1024);
}

static int run_single(gs_handle* h, bool cplx, int64_t n, double* A, int64_t lda, double* T, int64_t ldt, double* Z,
                      int64_t ldz, double* W, int flags, int64_t maxiter, int64_t maxinner, int32_t* info) {
  if (!h || !info) return fail(h, GS_E_ARG, "bad argument");
  if (n < 0 || lda < std::max<int64_t>(n, 1) || ldt < std::max<int64_t>(n, 1) ||
      ((flags & GS_WANTZ) && ldz < std::max<int64_t>(n, 1))) {
    *info = GS_INFO_DIM;
    return fail(h, GS_E_DIM, "DimensionMismatch: leading dimension smaller than n");
  }
  if (n > 0 && (!A || !T || !W || ((flags & GS_WANTZ) && !Z))) return fail(h, GS_E_ARG, "null pointer");
  std::lock_guard<std::mutex> lk(h->mu);
  HIPCHK(h, hipSetDevice(h->device));
  h->last_stats_from_plan = false;
  if (n == 0) {
    *info = 0;
    return GS_OK;
  }
  const size_t es = cplx ? 16 : 8;
  const size_t ed = es / 8;
  const bool dev = (flags & GS_DEVICE_PTRS) != 0;
  const int kflags = (flags & (GS_WANTZ | GS_SCALE | GS_HESSENBERG | GS_CHECK_SUBDIAG | GS_QR_BAND)) | env_qr_band();
  // device operands, packed with ld = n when staged
  double *dA, *dT, *dZ, *dW;
  int32_t* dInfo;
  int64_t dlda = lda, dldt = ldt, dldz = ldz;
  const size_t nn = (size_t)n * n;
  HIPCHK(h, h->bInfo.ensure(sizeof(int32_t)));
  HIPCHK(h, h->bStats.ensure(2 * sizeof(int64_t)));
  HIPCHK(h, h->bScl.ensure(2 * sizeof(double)));
  HIPCHK(h, h->bWs.ensure(std::max(es * 5 * (size_t)n + 64, 8 * zhq_global_ws_doubles(n))));
  if (dev) {
    dA = A;
    dT = T;
    dW = W;
    dInfo = static_cast<int32_t*>(h->bInfo.p);
    if (flags & GS_WANTZ) {
      dZ = Z;
    } else {
      HIPCHK(h, h->bZ.ensure(es * nn));
      dZ = static_cast<double*>(h->bZ.p);
      dldz = n;
    }
  } else {
    HIPCHK(h, h->bA.ensure(es * nn));
    HIPCHK(h, h->bT.ensure(es * nn));
    HIPCHK(h, h->bZ.ensure(es * nn));
    HIPCHK(h, h->bW.ensure(16 * (size_t)n));
    dA = static_cast<double*>(h->bA.p);
    dT = static_cast<double*>(h->bT.p);
    dZ = static_cast<double*>(h->bZ.p);
    dW = static_cast<double*>(h->bW.p);
    dInfo = static_cast<int32_t*>(h->bInfo.p);
    dlda = dldt = dldz = n;
    HIPCHK(h, copy_cols(dA, es * n, A, es * lda, es * n, n, hipMemcpyHostToDevice, h->stream));
    if ((flags & GS_HESSENBERG) && (flags & GS_WANTZ))
      HIPCHK(h, copy_cols(dZ, es * n, Z, es * ldz, es * n, n, hipMemcpyHostToDevice, h->stream));
  }
  const int nlds = cplx ? zschur_max_lds_n() : dschur_max_lds_n();
  const bool use_lds = !(flags & GS_PLAIN) && n <= nlds && dlda == n && dldt == n && dldz == n;
  const bool use_large = !cplx && !use_lds && !(flags & (GS_PLAIN | GS_HESSENBERG)) && n > nlds;
  // complex orders above the LDS-resident range: Hessenberg + Q in global
  // memory, then the windowed QR kernel with H and Z in HBM (ld = n)
  const bool use_win = cplx && !use_lds && !(flags & GS_PLAIN) && n > nlds && n <= INT32_MAX && dlda == n &&
                       dldt == n && dldz == n;
  h->launch_recs.clear();
  HIPCHK(h, hipEventRecord(h->ev0, h->stream));
  hipError_t le;
  if (use_large) {
    h->large_where.clear();
    le = run_large(h, n, dA, dlda, dT, dldt, dZ, dldz, dW, dInfo, kflags, maxiter, (flags & GS_TIMING) != 0);
    if (le != hipSuccess)
      return fail_hip(h, le, h->large_where.empty() ? "large path" : ("large path: " + h->large_where).c_str());
  } else if (use_lds) {
    HIPCHK(h, h->bMeta.ensure(64));
    // metadata for a batch of one: order = 0, n, off = 0, woff = 0
    struct Meta {
      int32_t order, n;
      int64_t off, woff;
    } meta = {0, (int32_t)n, 0, 0};
    HIPCHK(h, hipMemcpyAsync(h->bMeta.p, &meta, sizeof meta, hipMemcpyHostToDevice, h->stream));
    char* mp = static_cast<char*>(h->bMeta.p);
    BatchArgs a;
    a.order = reinterpret_cast<const int32_t*>(mp + offsetof(Meta, order));
    a.n = reinterpret_cast<const int32_t*>(mp + offsetof(Meta, n));
    a.off = reinterpret_cast<const int64_t*>(mp + offsetof(Meta, off));
    a.woff = reinterpret_cast<const int64_t*>(mp + offsetof(Meta, woff));
    a.first = 0;
    a.count = 1;
    a.A = dA;
    a.T = dT;
    a.Z = dZ;
    a.W = dW;
    a.info = dInfo;
    a.stats = static_cast<int64_t*>(h->bStats.p);
    a.scl = static_cast<double*>(h->bScl.p);
    a.flags = kflags;
    a.maxiter = maxiter;
    a.maxinner = maxinner;
    a.zlog = nullptr;  // one matrix: Z updated in place
    a.loff = nullptr;
    a.llen = nullptr;
    const int nthq = bin_nt_hq(h->tune[cplx ? 1 : 0], cplx, (int)n), ntqr = default_nt_qr(cplx, (int)n);
    le = cplx ? launch_zhq(a, (int)n, nthq, h->stream) : launch_dhq(a, (int)n, nthq, h->stream);
    if (le == hipSuccess) {
      if (kflags & GS_QR_BAND)
        le = cplx ? launch_zqr(a, (int)n, ntqr, h->stream) : launch_dqr(a, (int)n, ntqr, h->stream);
      else
        le = cplx ? launch_zqrw(a, h->stream) : launch_dqrw(a, h->stream);
    }
  } else if (use_win) {
    HIPCHK(h, h->bMeta.ensure(128));
    // Z log from order 768 up (unless GS_ZDIRECT; GS_ZLOG_TINY forces it for
    // tests): the chain wavefronts then update H only, and the log is
    // replayed onto Z's rows over the whole GPU after the sweeps (measured:
    // order 1024 2.84 -> 2.50 s; orders 256 and 512 slower, 0.216 -> 0.242
    // and 0.434 -> 0.471 s, the replay's row chains outweighing the chains
    // saved)
    const bool zlg = (kflags & GS_WANTZ) && !(flags & GS_ZDIRECT) && (n >= 768 || (flags & GS_ZLOG_TINY));
    const int64_t lcap = zlg ? zlog_capacity((int)n, (flags & GS_ZLOG_TINY) != 0) : 0;
    if (zlg) HIPCHK(h, h->bLog.ensure(16 * (size_t)lcap));
    struct Meta {
      int32_t order, n;
      int64_t off, woff;
      int64_t loff[2], llen;
    } meta = {0, (int32_t)n, 0, 0, {0, lcap}, 0};
    HIPCHK(h, hipMemcpyAsync(h->bMeta.p, &meta, sizeof meta, hipMemcpyHostToDevice, h->stream));
    SingleArgs g;
    g.n = n;
    g.A = dA;
    g.lda = dlda;
    g.T = dT;
    g.ldt = dldt;
    g.Z = dZ;
    g.ldz = dldz;
    g.W = dW;
    g.ws = static_cast<double*>(h->bWs.p);
    g.info = dInfo;
    g.stats = static_cast<int64_t*>(h->bStats.p);
    g.flags = kflags;
    g.maxiter = maxiter;
    g.maxinner = maxinner;
    le = launch_zhq_global(g, static_cast<double*>(h->bScl.p), h->stream);
    if (le == hipSuccess) {
      char* mp = static_cast<char*>(h->bMeta.p);
      BatchArgs a;
      a.order = reinterpret_cast<const int32_t*>(mp + offsetof(Meta, order));
      a.n = reinterpret_cast<const int32_t*>(mp + offsetof(Meta, n));
      a.off = reinterpret_cast<const int64_t*>(mp + offsetof(Meta, off));
      a.woff = reinterpret_cast<const int64_t*>(mp + offsetof(Meta, woff));
      a.first = 0;
      a.count = 1;
      a.A = dA;
      a.T = dT;
      a.Z = dZ;
      a.W = dW;
      a.info = dInfo;
      a.stats = static_cast<int64_t*>(h->bStats.p);
      a.scl = static_cast<double*>(h->bScl.p);
      a.flags = kflags;
      a.maxiter = maxiter;
      a.maxinner = maxinner;
      a.zlog = zlg ? static_cast<double*>(h->bLog.p) : nullptr;
      a.loff = zlg ? reinterpret_cast<const int64_t*>(mp + offsetof(Meta, loff)) : nullptr;
      a.llen = zlg ? reinterpret_cast<int64_t*>(mp + offsetof(Meta, llen)) : nullptr;
      a.zlog_min = 0;
      // chain wavefronts in proportion to the lines a chunk updates (about
      // 2n rows and columns): GSCHUR_WIN_NT overrides (128 .. 1024)
      le = launch_zqrw_single(a, win_nt((int)n), h->stream);
      if (le == hipSuccess && zlg)
        le = launch_zlog_rows((int)n, dZ, n, static_cast<const double*>(h->bLog.p),
                              reinterpret_cast<const int64_t*>(mp + offsetof(Meta, llen)), h->stream);
    }
  } else {
    SingleArgs a;
    a.n = n;
    a.A = dA;
    a.lda = dlda;
    a.T = dT;
    a.ldt = dldt;
    a.Z = dZ;
    a.ldz = dldz;
    a.W = dW;
    a.ws = static_cast<double*>(h->bWs.p);
    a.info = dInfo;
    a.stats = static_cast<int64_t*>(h->bStats.p);
    a.flags = kflags;
    a.maxiter = maxiter;
    a.maxinner = maxinner;
    const int nt = n <= 64 ? 64 : (n <= 512 ? 256 : 1024);
    le = cplx ? launch_zschur_global(a, nt, h->stream) : launch_dschur_global(a, nt, h->stream);
  }
  if (le != hipSuccess) return fail_hip(h, le, "kernel launch");
  HIPCHK(h, hipEventRecord(h->ev1, h->stream));
  HIPCHK(h, hipMemcpyAsync(info, dInfo, sizeof(int32_t), hipMemcpyDeviceToHost, h->stream));
  if (!dev) {
    HIPCHK(h, copy_cols(T, es * ldt, dT, es * n, es * n, n, hipMemcpyDeviceToHost, h->stream));
    if (flags & GS_WANTZ)
      HIPCHK(h, copy_cols(Z, es * ldz, dZ, es * n, es * n, n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(W, dW, 16 * (size_t)n, hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(h, hipStreamSynchronize(h->stream));
  (void)ed;
  return GS_OK;
}

extern "C" int gs_schur_z(gs_handle* h, int64_t n, double* A, int64_t lda, double* T, int64_t ldt, double* Z,
                          int64_t ldz, double* W, int flags, int64_t maxiter, int64_t maxinner, int32_t* info) {
  return run_single(h, true, n, A, lda, T, ldt, Z, ldz, W, flags, maxiter, maxinner, info);
}

extern "C" int gs_schur_d(gs_handle* h, int64_t n, double* A, int64_t lda, double* T, int64_t ldt, double* Z,
                          int64_t ldz, double* W, int flags, int64_t maxiter, int64_t maxinner, int32_t* info) {
  return run_single(h, false, n, A, lda, T, ldt, Z, ldz, W, flags, maxiter, maxinner, info);
}

extern "C" int gs_dgemm(gs_handle* h, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha,
                        const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                        int64_t ldc) {
  if (!h || m < 0 || n < 0 || k < 0 || m > INT32_MAX || n > INT32_MAX || k > INT32_MAX)
    return fail(h, GS_E_ARG, "bad argument");
  if (m == 0 || n == 0) return GS_OK;
  if (!A || !B || !C || ldc < m || lda < (ta ? k : m) || ldb < (tb ? n : k))
    return fail(h, GS_E_ARG, "bad argument");
  std::lock_guard<std::mutex> lk(h->mu);
  HIPCHK(h, hipSetDevice(h->device));
  const hipError_t e = launch_dgemm(ta != 0, tb != 0, (int)m, (int)n, (int)k, alpha, A, lda, B, ldb, beta, C, ldc,
                                    h->stream);
  if (e != hipSuccess) return fail_hip(h, e, "gs_dgemm");
  return GS_OK;
}

extern "C" int gs_last_launches(const gs_handle* hc, double* out, int max_records) {
  if (!hc) return 0;
  const int avail = (int)(hc->launch_recs.size() / GS_LAUNCH_REC);
  if (out && max_records > 0)
    std::copy(hc->launch_recs.begin(),
              hc->launch_recs.begin() + (size_t)std::min(avail, max_records) * GS_LAUNCH_REC, out);
  return avail;
}

extern "C" int gs_last_stats(const gs_handle* hc, int64_t* out, int len) {
  gs_handle* h = const_cast<gs_handle*>(hc);
  if (!h || !out || len <= 0) return 0;
  int64_t sweeps = 0, steps = 0;
  if (h->last_stats_from_plan && h->last_plan && h->last_plan->count > 0) {
    std::vector<int64_t> st(2 * (size_t)h->last_plan->count);
    if (hipMemcpy(st.data(), h->last_plan->d_stats.p, sizeof(int64_t) * st.size(), hipMemcpyDeviceToHost) !=
        hipSuccess)
      return 0;
    for (size_t i = 0; i < st.size(); i += 2) {
      sweeps += st[i];
      steps += st[i + 1];
    }
  } else if (h->bStats.p) {
    int64_t st[2] = {0, 0};
    if (hipMemcpy(st, h->bStats.p, sizeof st, hipMemcpyDeviceToHost) != hipSuccess) return 0;
    sweeps = st[0];
    steps = st[1];
  }
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, h->ev0, h->ev1) != hipSuccess) ms = 0.f;
  const int64_t vals[3] = {sweeps, steps, (int64_t)(ms * 1000.0f)};
  int k = 0;
  for (; k < len && k < 3; ++k) out[k] = vals[k];
  return k;
}
