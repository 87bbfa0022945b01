// Internal declarations shared by the host API and the kernel translation
// units of libgschur_mi355x.  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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
  double* ws;  // device scratch: >= 4 n complex / 4 n real elements
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
hipError_t launch_dhq(const BatchArgs& a, int nmax, int nt, hipStream_t s);
hipError_t launch_dqr(const BatchArgs& a, int nmax, int nt, hipStream_t s);
hipError_t launch_zschur_global(const SingleArgs& a, int nt, hipStream_t s);
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
hipError_t launch_ztrevc(const TrevcArgs& a, int nmax, hipStream_t s);
hipError_t launch_dtrevc(const TrevcArgsD& a, int nmax, hipStream_t s);

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

}  // namespace gs
