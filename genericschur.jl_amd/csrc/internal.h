// Internal declarations shared by the host API and the kernel translation
// units of libgschur_mi355x.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace gs {

// Arguments of the batched Schur kernels.  Matrix m has order n[m], lives at
// element offset off[m] of A/T/Z (column-major, ld = n[m]) and owns the
// eigenvalue slots woff[m] .. woff[m] + n[m] - 1 of W.  Launch item b
// processes matrix order[first + b].
struct BatchArgs {
  const int32_t* order;
  int first;
  int count;
  const int32_t* n;
  const int64_t* off;
  const int64_t* woff;
  const double* A;
  double* T;
  double* Z;
  double* W;
  int32_t* info;
  int64_t* stats;  // 2 per matrix (sweeps, bulge steps) or null
  double* scl;     // 2 per matrix: (cscale, anrm) of _scale!, or (0, 0)
  int flags;
  int64_t maxiter;   // <= 0: 100 n
  int64_t maxinner;  // <= 0: 30 n
  // Z log of the complex windowed QR kernel (null: Z updated in place):
  // matrix m's records occupy zlog[2 loff[m] .. 2 loff[m+1]) (complex
  // entries as double pairs); the kernel stores the length used in llen[m]
  double* zlog;
  const int64_t* loff;
  int64_t* llen;
  // with zlog set, only matrices of order >= zlog_min log their Z updates
  int zlog_min = 0;
  // work queue of the persistent QR launch: matrices order[first + q] are
  // taken in turn, q from an atomic counter (zero before the launch)
  int* qctr = nullptr;
};

// Arguments of the single-matrix global-memory kernels (any n).
struct SingleArgs {
  int64_t n;
  double* A;  // input (device), ld = lda
  int64_t lda;
  double* T;
  int64_t ldt;
  double* Z;
  int64_t ldz;
  double* W;
  double* ws;  // device scratch: >= 5 n complex / 5 n real elements
  int32_t* info;
  int64_t* stats;
  int flags;
  int64_t maxiter;
  int64_t maxinner;
};

// Batched kernels are two launches per size bin: the Hessenberg reduction
// and Q formation with the whole matrix in LDS ("hq"), then the QR sweeps
// with only H's band in LDS ("qr").
size_t zhq_lds_bytes(int n, int nt);
size_t zqr_lds_bytes(int n, int nt);
size_t dhq_lds_bytes(int n, int nt);
size_t dqr_lds_bytes(int n, int nt);
int zschur_max_lds_n();
int dschur_max_lds_n();

hipError_t launch_zhq(const BatchArgs& a, int nmax, int nt, hipStream_t s);
hipError_t launch_zqr(const BatchArgs& a, int nmax, int nt, hipStream_t s);
// windowed QR launch: one wavefront per matrix, H and Z in HBM
hipError_t launch_zqrw(const BatchArgs& a, hipStream_t s);
// The persistent form of launch_zqrw: `slots` workgroups take the matrices
// order[first .. first + count) from the work queue a.qctr, in order.
hipError_t launch_zqrw_queue(const BatchArgs& a, int slots, hipStream_t s);
// Zeroes n ints (a kernel, so that it can sit in a captured graph).
hipError_t launch_fill_i32(int* p, int n, hipStream_t s);
hipError_t launch_zlog(const BatchArgs& a, int nmax, hipStream_t s);
// Z-log capacity of one matrix of order n, in complex entries (tiny: one
// sweep's record, the least the kernel accepts)
inline int64_t zlog_capacity(int n, bool tiny) {
  const int64_t one_sweep = 3 + 2 * (int64_t)n;
  return tiny ? one_sweep : 6 * (int64_t)n * n + 16 * (int64_t)n + 64;
}
hipError_t launch_dqrw(const BatchArgs& a, hipStream_t s);
hipError_t launch_dlog(const BatchArgs& a, int nmax, hipStream_t s);
// Real Z-log capacity of one matrix of order n, in doubles (tiny: one sweep
// and one rotation, the least the kernel accepts)
inline int64_t dlog_capacity(int n, bool tiny) {
  const int64_t one = 2 + 3 * (int64_t)n + 3;
  return tiny ? one : 5 * (int64_t)n * n + 16 * (int64_t)n + 64;
}
hipError_t launch_dhq(const BatchArgs& a, int nmax, int nt, hipStream_t s);
hipError_t launch_dqr(const BatchArgs& a, int nmax, int nt, hipStream_t s);
hipError_t launch_zschur_global(const SingleArgs& a, int nt, hipStream_t s);
// one matrix above the LDS-resident orders: Hessenberg + Q in global memory
// (H in a.T, Q in a.Z, ld = n), then launch_zqrw on a batch of one
hipError_t launch_zhq_global(const SingleArgs& a, double* scl, hipStream_t s);
// doubles of a.ws launch_zhq_global needs (tau, two vectors, gemv partials)
size_t zhq_global_ws_doubles(int64_t n);
// the windowed QR kernel with nt threads (128 .. 1024): one chasing
// wavefront, nt / 64 - 1 chain wavefronts
hipError_t launch_zqrw_single(const BatchArgs& a, int nt, hipStream_t s);
// the Z log of one matrix (llen entries at lg) replayed onto Z's rows in
// HBM (ld ld) over the whole GPU
hipError_t launch_zlog_rows(int n, double* Z, int64_t ld, const double* lg, const int64_t* llen, hipStream_t s);
hipError_t launch_dschur_global(const SingleArgs& a, int nt, hipStream_t s);

// Eigenvector kernels (one workgroup per matrix).  T, Z at element offset
// off[m] (complex elements for the complex kernel), V (complex) at complex
// offset off[m].  ld* > 0 gives leading dimensions for a single matrix;
// 0 means ld = n.
struct TrevcArgs {
  const int32_t* order;
  int count;
  const int32_t* n;
  const int64_t* off;
  const double* T;
  const double* Z;  // may be null
  double* V;
  int left;
  int normalize;
  int64_t ldt, ldz, ldv;
};
typedef TrevcArgs TrevcArgsD;

size_t ztrevc_lds_bytes(int n);
size_t dtrevc_lds_bytes(int n);
// largest order the eigenvector kernels take (T in global memory above the
// LDS-resident orders 128 complex / 192 real)
int ztrevc_max_n();
int dtrevc_max_n();
hipError_t launch_ztrevc(const TrevcArgs& a, int nmax, hipStream_t s);
hipError_t launch_dtrevc(const TrevcArgsD& a, int nmax, hipStream_t s);
// One matrix of order z/dtrevc_max_n() < n <= trevc_big_max_n() over the
// whole GPU (solves one per workgroup with the vector in LDS, the
// back-transform as FP64 MFMA GEMMs).  a.T, a.Z, a.V point at the matrix,
// a.ld* its leading dimensions (0: n); order/count/n/off are not read.
// ws: z/dtrevc_big_ws_bytes(n) bytes of device memory.
int trevc_big_max_n();
size_t dtrevc_big_ws_bytes(int n);
size_t ztrevc_big_ws_bytes(int n);
hipError_t launch_dtrevc_big(const TrevcArgsD& a, int n, void* ws, hipStream_t s);
hipError_t launch_ztrevc_big(const TrevcArgs& a, int n, void* ws, hipStream_t s);

// Eigenvalue VJP assembly: VR, VL (complex) at complex offset off[m], gbar
// (complex) at woff[m], Abar (real) at off[m].
struct VjpArgs {
  const int32_t* order;
  int count;
  const int32_t* n;
  const int64_t* off;
  const int64_t* woff;
  const double* VR;
  const double* VL;
  const double* gbar;
  double* Abar;
  int32_t* info;
};
size_t vjp_lds_bytes(int n);
hipError_t launch_vjp_assemble(const VjpArgs& a, int nmax, hipStream_t s);
// one matrix above the LDS-resident VJP orders: c_k and the planes
// P = [Re(Y C) | Im(Y C)], Q = [Re X | Im X] (n x 2n, ld n) of the GEMM
// Abar = P Q^T (info set to 1 when some y_k^H x_k vanished); and a fill
// kernel that clears info[0 .. count)
hipError_t launch_vjp_planes(int n, const double* VR, const double* VL, const double* g, double* P, double* Q,
                             int32_t* info, hipStream_t s);
hipError_t launch_vjp_info_clear(int32_t* info, int count, hipStream_t s);


// eigen! pieces (eigen.hip): balancing of one matrix (out = ilo, ihi,
// trivial, status), the balancer's back-transform of complex columns,
// triangularize, _enormalize!, reciprocal eigenvalue condition numbers
hipError_t launch_balance(bool cplx, double* A, int64_t lda, int n, int scale, int permute, int32_t* out, int32_t* sp,
                          double* D, hipStream_t s);
hipError_t launch_balback(int n, int m, const int32_t* iit, const int32_t* sp, const double* D, double* V,
                          int64_t ldv, int inverse, hipStream_t s);
hipError_t launch_triangularize(int n, const double* T, int64_t ldt, const double* Z, int64_t ldz, double* Tc,
                                int64_t ldtc, double* Zc, int64_t ldzc, double* W, hipStream_t s);
hipError_t launch_enormalize(int n, int m, double* V, int64_t ldv, hipStream_t s);
// gordschur! on a complex Schur form (T, Z: ld ldt, ldz; select: n ints)
// and eigvalscond / subspacesep (work: 3 nsub (n - nsub) complex entries,
// out: s, sep), one wavefront each
// Hermitian / real symmetric eigenproblem (heev.hip).  mode 0: eigenvalues
// W (and eigenvectors Z when non-null; E is workspace of count x n); mode 2:
// the tridiagonal matrix (W = d, E = e) and Q (Z when non-null).  A: lower (upper = 0) or upper triangle
// of each Hermitian matrix (lda, stride sA elements).
struct HeevArgs {
  int count = 0, n = 0, upper = 0, mode = 0, sort = 0, maxsweeps = 0;
  const void* A = nullptr;
  int64_t lda = 0, sA = 0;
  double* W = nullptr;
  int64_t sW = 0;
  double* E = nullptr;
  int64_t sE = 0;
  void* Z = nullptr;
  int64_t ldz = 0, sZ = 0;
  int* info = nullptr;
  int* sweeps = nullptr;
  // mode 0 workspace: per matrix a rotation record stream (logcap records
  // of 16 bytes; null: eigenvalues only), its length, the QL/QR state
  void* log = nullptr;
  int logcap = 0;
  int* logpos = nullptr;
  void* state = nullptr;
};
struct SteqrArgs {
  int n = 0, zrows = 0, maxsweeps = 0;
  double* d = nullptr;
  double* e = nullptr;
  void* Z = nullptr;
  int64_t ldz = 0;
  int* info = nullptr;
  int* sweeps = nullptr;
};
// LDS shared by the reduction's vectors (tau, two reflector buffers, w) and the rotation log
constexpr size_t heev_union_bytes(int n, size_t es) {
  return ((4 * (size_t)n * es > 56 * (size_t)n ? 4 * (size_t)n * es : 56 * (size_t)n) + 15) / 16 * 16;
}
constexpr int heev_max_n(bool cplx) { return cplx ? 96 : 128; }
constexpr int steqr_max_n() { return 2048; }
size_t steqr_lds_bytes(int n);
size_t heev_log_records(int n);   // records per matrix of the mode 0 stream
size_t heev_state_bytes();        // bytes per matrix of the QL/QR state
hipError_t launch_heev(const HeevArgs& a, bool cplx, hipStream_t s);
// Orders heev_max_n < n <= heev_big_max_n, one matrix at a time over the
// whole GPU (the working matrix, Q and the eigenvectors in HBM); modes 0
// and 2 as launch_heev (E is output only in mode 2).  ws holds
// heev_big_ws_bytes(n, cplx) bytes of device memory.
constexpr int heev_big_max_n() { return 4096; }
size_t heev_big_ws_bytes(int n, bool cplx);
hipError_t launch_heev_big(const HeevArgs& a, bool cplx, void* ws, hipStream_t s);
// gs_steqr for steqr_max_n < n <= heev_big_max_n: the recurrence's
// rotations through a record stream in HBM (ws: steqr_big_ws_bytes(n)),
// replayed onto Z over the whole GPU; the host waits once per stream round
size_t steqr_big_ws_bytes(int n);
hipError_t launch_steqr_big(const SteqrArgs& a, bool zcplx, void* ws, hipStream_t s);
hipError_t launch_steqr(const SteqrArgs& a, bool zcplx, hipStream_t s);

hipError_t launch_ordschur_d(int n, double* T, int64_t ldt, double* Z, int64_t ldz, bool wantz, const int32_t* select,
                             double* W, int32_t* info, hipStream_t s);
hipError_t launch_ordschur(int n, double* T, int64_t ldt, double* Z, int64_t ldz, bool wantz, const int32_t* select,
                           double* W, hipStream_t s);
hipError_t launch_eigcond(int n, const double* T, int64_t ldt, int nsub, double* work, double* out, hipStream_t s);
hipError_t launch_eig_rcond(int n, int m, const double* VL, int64_t ldl, const double* VR, int64_t ldr, double* rc,
                            hipStream_t s);

// No-allocation batched entry point (gs_schur_batched_z_noalloc): the size
// of its fixed device scratch array (one per GPU, noalloc.hip) and the most
// matrices one call takes (56 bytes of metadata each, 256-byte aligned
// segments).
constexpr size_t kNoAllocScratchBytes = size_t(1) << 20;
constexpr int64_t kNoAllocMaxCount = 16384;
// Device address of this GPU's scratch array (fetched with
// hipGetSymbolAddress once per device, then cached).
hipError_t noalloc_scratch(int device, void** ptr);

// FP64 MFMA GEMM: C(m x n) = beta C + alpha op(A) op(B), column-major.
hipError_t launch_dgemm(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int64_t lda,
                        const double* B, int64_t ldb, double beta, double* C, int64_t ldc, hipStream_t s);
// the same product split over K into partial products in ws (for m <= 32
// and long k), then summed into C; ws holds dgemm_splitk_doubles(m, n, k)
size_t dgemm_splitk_doubles(int m, int n, int k);
hipError_t launch_dgemm_splitk(bool ta, bool tb, int m, int n, int k, double alpha, const double* A, int64_t lda,
                               const double* B, int64_t ldb, double beta, double* C, int64_t ldc, double* ws,
                               hipStream_t s);

// ---------------------------------------------------------- large real path
// One real matrix of any order on the whole GPU: blocked Hessenberg with
// compact-WY GEMM updates, blocked Q formation, chunked double-shift QR.
struct LargeArgs {
  int64_t n;
  const double* A;  // device, ld = lda (not modified)
  int64_t lda;
  double* T;  // working H, then T
  int64_t ldt;
  double* Z;  // n x n (required: Q accumulation target even without wantZ)
  int64_t ldz;
  double* W;  // n complex eigenvalues
  int32_t* info;   // device
  int64_t* stats;  // device, 2 values, or null
  int flags;
  int64_t maxiter;
};

struct LargeWorkSizes {
  int64_t Y, V, Ytmp, Tall, Wb, W2, tau, tv, part, scl, lg, st, gk;
};

// device workspace carved from one allocation, plus a pinned host record
struct LargeWork {
  double *Y, *V, *Ytmp, *Tall, *Wb, *W2, *tau, *tv, *part, *scl, *lg;
  void* st;
  double* gk;  // split-K partial products
  void* host_state;
  const char* where = "";  // the call that failed, for the error message
};

// events around phases and GEMM launches (only when timing is requested)
struct LargeTimer {
  std::vector<hipEvent_t>* pool = nullptr;
  size_t used = 0;
  struct Rec {
    int kind;  // 2: GEMM (a, b, c = m, n, k), 3: phase (a = phase index)
    int a, b, c;
    double flops;
    size_t e0, e1;
  };
  std::vector<Rec> recs;
  size_t phase_ev[6] = {0, 0, 0, 0, 0, 0};
  hipError_t ev(hipStream_t s, size_t& idx) {
    if (used == pool->size()) {
      hipEvent_t e = nullptr;
      hipError_t r = hipEventCreate(&e);
      if (r != hipSuccess) return r;
      pool->push_back(e);
    }
    idx = used++;
    return hipEventRecord((*pool)[idx], s);
  }
  hipError_t mark(hipStream_t s, int ph) { return ev(s, phase_ev[ph]); }
};

inline hipError_t tm_gemm(LargeTimer* tm, hipStream_t s, bool ta, bool tb, int m, int n, int k, double alpha,
                          const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C,
                          int64_t ldc, double* splitk_ws = nullptr) {
  if (m <= 0 || n <= 0) return hipSuccess;
  size_t e0 = 0, e1 = 0;
  hipError_t r;
  if (tm && (r = tm->ev(s, e0)) != hipSuccess) return r;
  r = splitk_ws ? launch_dgemm_splitk(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, splitk_ws, s)
                : launch_dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, s);
  if (r != hipSuccess) return r;
  if (tm) {
    if ((r = tm->ev(s, e1)) != hipSuccess) return r;
    tm->recs.push_back({2, m, n, k, 2.0 * m * n * (double)k, e0, e1});
  }
  return hipSuccess;
}

size_t large_work_doubles(int64_t n, LargeWorkSizes* sz);
size_t large_state_bytes();
hipError_t large_schur_d(const LargeArgs& a, LargeWork& w, hipStream_t s, LargeTimer* tm);

}  // namespace gs
