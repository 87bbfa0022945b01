// The library handle (struct gs_handle) shared by the host translation
// units, and small helpers around it.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gschur_mi355x.h"

namespace gs {

// grow-only device buffer
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&p, want);
    if (e == hipSuccess) cap = want;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct Plan;
constexpr int kBranches = 4;
constexpr int kAuxBufs = 14;

// Launch parameters of batched calls set by gs_set_tuning (a table written
// by the autotuner); zero or -1 fields keep the built-in defaults.
struct Tuning {
  int bin_w = 0;   // size-bin width
  int lead = -1;   // leading bins whose Hessenberg launches run first, alone
  int queue = -1;  // complex: one work-queue QR launch over all bins (1) or one per bin (0)
  int slots = 0;   // workgroups per CU of the work-queue launch
  struct Row {
    int nlo, nhi;  // bins whose largest order lies in [nlo, nhi]
    int nt_hq;     // Hessenberg + Q workgroup size (0: default)
    int zlog;      // Z log of the QR launch: 1 on, 0 off, -1 default
  };
  std::vector<Row> rows;
};

}  // namespace gs

struct gs_handle {
  int device = 0;
  hipStream_t stream = nullptr;      // stream in use
  hipStream_t own_stream = nullptr;  // stream created with the handle
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipStream_t branch[gs::kBranches] = {};
  hipEvent_t evfork = nullptr, evjoin[gs::kBranches] = {};
  hipEvent_t evlead[3] = {};  // after the leading bins' Hessenberg launches
  std::string err;
  gs::DevBuf bA, bT, bZ, bW, bInfo, bWs, bStats, bMeta, bScl;
  gs::DevBuf bLog;  // Z log of the batched complex QR launches
  gs::DevBuf aux[gs::kAuxBufs];  // scratch of the eigenvector / VJP entry points
  std::map<std::string, std::unique_ptr<gs::Plan>> plans;
  gs::Tuning tune[2];  // [0] real, [1] complex batches
  gs::Plan* last_plan = nullptr;
  bool last_stats_from_plan = false;
  // GS_TIMING: events around each launch (two per launch) and the records
  // read back by gs_last_launches
  std::vector<hipEvent_t> tev;
  std::vector<double> launch_recs;
  // workspace of the large-matrix real path and its pinned state record
  gs::DevBuf lwork;
  void* lhost = nullptr;
  std::string large_where;
  // host staging block of the no-allocation batched entry point
  std::vector<unsigned char> noalloc_host;
  std::mutex mu;
};

namespace gs {

inline int handle_fail_hip(gs_handle* h, hipError_t e, const char* where) {
  if (h) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "%s: %s", where, hipGetErrorString(e));
    h->err = buf;
  }
  return GS_E_HIP - (int)e;
}
inline int handle_fail(gs_handle* h, int code, const char* msg) {
  if (h) h->err = msg;
  return code;
}
// Batched operand layout: matrix i occupies elements [off[i], off[i] + n[i]^2)
// of its arrays (ld = n[i]).  Every n[i] must be >= 0 (GS_E_DIM otherwise),
// every off[i] >= 0, and the nonempty spans of different matrices must be
// disjoint (GS_E_ARG otherwise), so that no kernel writes outside the
// caller's arrays or into another matrix.  *span receives max(off[i] + n[i]^2).
inline int check_spans(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, int64_t* span) {
  std::vector<std::pair<int64_t, int64_t>> iv;
  iv.reserve((size_t)count);
  int64_t s = 0;
  for (int64_t i = 0; i < count; ++i) {
    if (n[i] < 0) return handle_fail(h, GS_E_DIM, "negative matrix order");
    if (off[i] < 0) return handle_fail(h, GS_E_ARG, "negative matrix offset");
    const int64_t len = (int64_t)n[i] * n[i];
    if (off[i] > INT64_MAX - len) return handle_fail(h, GS_E_ARG, "matrix offset out of range");
    if (len == 0) continue;
    iv.emplace_back(off[i], off[i] + len);
    s = std::max(s, off[i] + len);
  }
  std::sort(iv.begin(), iv.end());
  for (size_t k = 1; k < iv.size(); ++k)
    if (iv[k].first < iv[k - 1].second) return handle_fail(h, GS_E_ARG, "overlapping matrix spans");
  if (span) *span = s;
  return GS_OK;
}

// Host <-> device copy of a column-major block (cols columns of width
// bytes, pitches in bytes): one linear copy when both sides are packed,
// the pitched copy otherwise (which moves pageable memory column by column).
inline hipError_t copy_cols(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width, size_t cols,
                            hipMemcpyKind kind, hipStream_t s) {
  if (dpitch == width && spitch == width) return hipMemcpyAsync(dst, src, width * cols, kind, s);
  return hipMemcpy2DAsync(dst, dpitch, src, spitch, width, cols, kind, s);
}

inline std::mutex& handle_mutex(gs_handle* h) { return h->mu; }
inline int handle_device(gs_handle* h) { return h->device; }
inline hipStream_t handle_stream(gs_handle* h) { return h->stream; }
inline DevBuf& handle_buf(gs_handle* h, int i) { return h->aux[i]; }

}  // namespace gs
