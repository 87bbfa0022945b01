/*
 * gschur_mi355x.h — C ABI of the MI355X (gfx950) Schur library.
 *
 * The library computes the Schur decomposition A = Z T Z^H of square
 * Float64 / ComplexF64 matrices the way GenericSchur.jl's `gschur!` does —
 * Householder Hessenberg reduction, Francis QR sweeps with deflation and
 * accumulation of Z (single shift for complex input, double shift to the
 * standard-form quasi-triangular T for real input) — plus the eigenvector
 * back-substitution behind `eigvecs`/`geigvecs` and an eigenvalue
 * vector-Jacobian product.  All work runs in hand-written HIP kernels.
 *
 * Conventions
 *   - Matrices are column-major; complex data are interleaved (re, im)
 *     doubles, i.e. Julia ComplexF64 / C `double _Complex`.  `double*`
 *     arguments of complex routines point at the first real part.
 *   - Eigenvalues are always complex (interleaved pairs), in the diagonal
 *     order of T, as `Schur.values` in the reference.
 *   - Without GS_DEVICE_PTRS every data pointer is host memory and the call
 *     stages it through device buffers; with it every data pointer is device
 *     memory owned by the caller.  Size/offset arrays are always host memory.
 *   - Calls are synchronous unless GS_ASYNC is set (device pointers only).
 *   - One handle serialises its calls on its own HIP stream; distinct
 *     handles are independent and may be used from different threads.
 *
 * Return value of every call: GS_OK or a negative status (GS_E_*).
 * Per-matrix `info`: 0 = converged, > 0 = iteration limit reached
 * (UnconvergedException), -1 = bad dimension (DimensionMismatch),
 * -2 = non-real subdiagonal in Hessenberg input (ArgumentError).
 */
#ifndef GSCHUR_MI355X_H
#define GSCHUR_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gs_handle gs_handle;

/* flags */
#define GS_WANTZ 0x01        /* accumulate Schur vectors (wantZ = true)        */
#define GS_SCALE 0x02        /* over/underflow scaling of A (scale = true)     */
#define GS_HESSENBERG 0x04   /* A is upper Hessenberg; with GS_WANTZ, Z holds  */
                             /* the transformation to update (gschur!(H, Z))   */
#define GS_CHECK_SUBDIAG 0x08 /* with GS_HESSENBERG, complex: reject non-real  */
                              /* subdiagonals (checksd = true)                 */
#define GS_DEVICE_PTRS 0x10  /* data pointers are device memory                */
#define GS_ASYNC 0x20        /* do not synchronise before returning            */
#define GS_PLAIN 0x40        /* single matrix: force the unblocked global-     */
                             /* memory kernel (the plain GPU form)             */
#define GS_QR_BAND 0x100     /* LDS-resident QR phase: keep H's band in LDS    */
                             /* (up to 80 KB per matrix) instead of the        */
                             /* default windowed kernel (H in HBM, a 32 x 32   */
                             /* chase window in LDS); same arithmetic           */
#define GS_ZDIRECT 0x200     /* complex (batched, and single from order 768):  */
                             /* apply each chunk's reflectors to Z in HBM      */
                             /* during the QR sweeps instead of logging them   */
                             /* and applying the log to Z after the sweeps     */
                             /* (batched: Z in LDS; single: Z's rows over the  */
                             /* GPU); same arithmetic and order (comparisons)  */
#define GS_ZLOG_TINY 0x400   /* complex, testing: a Z log of one sweep per     */
                             /* matrix, so that the QR kernel applies the log  */
                             /* to Z itself before almost every sweep (single: */
                             /* the log at any order above the LDS range)      */
#define GS_SORT 0x800       /* Hermitian eigensolver: eigenvalues ascending, */
                             /* eigenvectors alike (sorteig! with eigsortby)  */
#define GS_TIMING 0x80       /* record launch durations for gs_last_launches:  */
                             /* batched with GS_DEVICE_PTRS (launched eagerly, */
                             /* no graph, events around every kernel on its    */
                             /* stream) and the large real single-matrix path  */
                             /* (events around its phases and GEMMs)           */

/* status codes */
#define GS_OK 0
#define GS_E_ARG (-3)
#define GS_E_DIM (-1)
#define GS_E_UNSUPPORTED (-4)
#define GS_E_HIP (-100) /* -100 - hipError_t */

/* info codes */
#define GS_INFO_DIM (-1)
#define GS_INFO_NONREAL_SUBDIAG (-2)

int gs_version(void);

/* Context: owns the device, stream, staging buffers, size-bin tables and
 * captured graphs.  No counterpart in the reference (stateless Julia). */
int gs_create(gs_handle** h, int device);
int gs_destroy(gs_handle* h);
const char* gs_last_error(const gs_handle* h);
int gs_synchronize(gs_handle* h);
/* Run this handle's work on the caller's HIP stream (a hipStream_t passed as
 * void*), e.g. torch.cuda.current_stream().cuda_stream; NULL restores the
 * handle's own stream.  No counterpart in the reference. */
int gs_set_stream(gs_handle* h, void* stream);

/* Single matrix, ComplexF64.
 * Replaces gschur!(A::StridedMatrix{Complex{T}}; wantZ, scale, maxiter,
 * maxinner) — GenericSchur.jl:350-372, with the QR driver of :194-335 and the
 * sweep of :374-504; with GS_HESSENBERG, gschur!(H::Hessenberg, Z) of
 * :194-335.  A is n x n (lda >= n) and is overwritten (destroyed), as in the
 * reference.  T (ldt), Z (ldz, referenced with GS_WANTZ) receive the factors,
 * W (2n doubles) the eigenvalues diag(T).  maxiter <= 0 selects 100 n,
 * maxinner <= 0 selects 30 n.
 * Orders up to gs_batched_max_n(1) run the LDS-resident kernels; larger ones
 * (unless GS_PLAIN) a global-memory Hessenberg reduction and Q formation
 * followed by the windowed QR kernel (H and Z in HBM, the chase in an LDS
 * window, more chain wavefronts for larger n).  GS_PLAIN runs the unblocked
 * single-workgroup kernel. */
int gs_schur_z(gs_handle* h, int64_t n, double* A, int64_t lda, double* T, int64_t ldt, double* Z, int64_t ldz,
               double* W, int flags, int64_t maxiter, int64_t maxinner, int32_t* info);

/* Single matrix, Float64.
 * Replaces gschur!(A::StridedMatrix{<:AbstractFloat}; wantZ, scale, maxiter)
 * — GenericSchur.jl:805-835 with the QR driver :513-699, the double-shift
 * sweep :837-952 and the 2x2 standardisation _gs2x2! :716-803.  T is the
 * standard-form quasi-triangular factor; W (2n doubles) holds the
 * eigenvalues by position with the +imag member of a pair first.
 * maxinner is ignored (the reference's real driver has none).
 * Orders above gs_batched_max_n(0) take the large-matrix path unless
 * GS_PLAIN or GS_HESSENBERG is set: blocked Hessenberg reduction and Q
 * formation with FP64 MFMA GEMM updates, then the same double-shift QR
 * iteration with each sweep chased in LDS-resident chunks whose reflectors
 * are applied to the rest of H and to Z by the whole GPU (same per-element
 * update order as the unblocked sweep).  GS_PLAIN runs the unblocked
 * single-workgroup kernel instead. */
int gs_schur_d(gs_handle* h, int64_t n, double* A, int64_t lda, double* T, int64_t ldt, double* Z, int64_t ldz,
               double* W, int flags, int64_t maxiter, int64_t maxinner, int32_t* info);

/* Batched: `count` matrices of orders n[i], packed column-major (ld = n[i])
 * at element offsets off[i] in A, T and Z (complex elements for _z).  W holds
 * the eigenvalues of matrix i at complex offset sum_{j<i} n[j].  info[i] per
 * matrix.  A is read only.  One call is one graph launch of the size-binned
 * LDS-resident kernels; n[i] must not exceed gs_batched_max_n(). */
int gs_schur_batched_z(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, const double* A,
                       double* T, double* Z, double* W, int32_t* info, int flags);
int gs_schur_batched_d(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, const double* A,
                       double* T, double* Z, double* W, int32_t* info, int flags);
int gs_batched_max_n(int is_complex);

/* Launch parameters of this handle's batched calls for one element type
 * (no counterpart in the reference), as written by the autotuner
 * (genericschur/autotune.py).  rows holds nrows entries of four int32
 * {kind, a, b, value}:
 *   kind 0: a = 1 size-bin width (4..64), 2 leading Hessenberg launches
 *           (0..3), 3 one work-queue QR launch over all bins (complex; 0/1),
 *           4 workgroups per CU of that launch (1..16);
 *   kind 1: Hessenberg + Q workgroup size (64 up to n = 32, 256, 512 or
 *           1024; 0 = default) of bins whose largest order lies in [a, b];
 *   kind 2: Z log of those bins' QR launches (1 on, 0 off, -1 default).
 * nrows = 0 restores the defaults.  None of these changes the results of a
 * matrix between its batched and single-matrix calls on one handle (the
 * single-matrix LDS path takes the same Hessenberg workgroup size); a
 * workgroup size does change the rounding of the Hessenberg reduction.
 * Cached plans and graphs are dropped. */
int gs_set_tuning(gs_handle* h, int is_complex, const int32_t* rows, int nrows);

/* Batched ComplexF64 Schur without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and info codes as
 * gs_schur_batched_z, with these differences:
 *   - every data pointer is device memory (GS_DEVICE_PTRS is implied) and
 *     Z is always written (GS_WANTZ is implied; GS_SCALE is honoured, other
 *     flags are ignored);
 *   - count <= 16384 (GS_E_ARG above): the per-matrix metadata lives in a
 *     fixed 1 MiB device array that each GPU holds until the process ends;
 *   - Z is updated in place during the sweeps (the GS_ZDIRECT form; same
 *     results bit for bit as the Z-log form), the launches run eagerly on
 *     the handle's stream, and the call returns when they are done;
 *   - the scratch array is shared by every call on one GPU: calls that
 *     overlap on one GPU (two handles, two threads) race.
 * It calls no hipMalloc, hipFree or other allocating routine. */
int gs_schur_batched_z_noalloc(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, const double* A,
                               double* T, double* Z, double* W, int32_t* info, int flags);

/* Eigenvectors of a Schur factorization (T, Z), columns in the order of the
 * eigenvalues, normalised to unit 2-norm with the largest component real
 * (geigvecs(S; left) — vectors.jl:12-20 with _enormalize! util.jl:572-592;
 * complex right/left: vectors.jl:45-131 / :372-460 with _usolve!/_cusolve!
 * util.jl:128-461; real right/left: vectors.jl:133-362 / :462-698 with
 * _xsolve/_safe_lu_solve! util.jl:603-751).  Z may be NULL (vectors of T).
 * V is n x n complex (ldv).  normalize = 0 returns the max-abs1-scaled
 * vectors before _enormalize!.  Orders up to 4096 (GS_E_UNSUPPORTED above):
 * one workgroup per matrix up to 1024 (T in LDS up to 128 complex / 192
 * real, read from global memory above); above 1024 one matrix at a time
 * over the whole GPU, with n x n workspaces (real 40 n^2 bytes, complex
 * 64 n^2) held by the handle. */
int gs_trevc_z(gs_handle* h, int64_t n, const double* T, int64_t ldt, const double* Z, int64_t ldz, double* V,
               int64_t ldv, int left, int normalize, int flags);
int gs_trevc_d(gs_handle* h, int64_t n, const double* T, int64_t ldt, const double* Z, int64_t ldz, double* V,
               int64_t ldv, int left, int normalize, int flags);
/* Batched eigenvectors: T, Z at off[i] (complex elements for _z, real for
 * _d); V (complex) at complex offset off[i]. */
int gs_trevc_batched_z(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, const double* T,
                       const double* Z, double* V, int left, int normalize, int flags);
int gs_trevc_batched_d(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, const double* T,
                       const double* Z, double* V, int left, int normalize, int flags);

/* Eigenvalue vector-Jacobian product (no counterpart in the reference; the
 * backward of an eigenvalue rrule).  For the real Schur factorizations
 * (T, Z) of real matrices A_i (as returned by gs_schur_batched_d) and
 * cotangents gbar_k = dL/dRe(lambda_k) + i dL/dIm(lambda_k), one complex
 * value per eigenvalue slot at the eigenvalue offsets of gs_schur_batched_d,
 * writes Abar_i = Re( sum_k gbar_k / conj(y_k^H x_k) y_k x_k^H )
 * = Re( X^{-H} diag(gbar) X^H ), with the right (x_k) and left (y_k)
 * eigenvectors recomputed from T and Z on the device.  Abar is real at
 * off[i].  info[i] = 0, or 1 when some y_k^H x_k vanished (defective).
 * Orders up to 4096: up to 70 the assembly runs with both eigenvector
 * factors in LDS, above as the FP64 MFMA GEMM Abar = P Q^T with
 * P = [Re(Y C) | Im(Y C)], Q = [Re X | Im X], C = diag(c_k); above 1024
 * the eigenvectors are recomputed one matrix at a time over the GPU. */
int gs_eig_vjp_batched_d(gs_handle* h, int64_t count, const int32_t* n, const int64_t* off, const double* T,
                         const double* Z, const double* gbar, double* Abar, int32_t* info, int flags);

/* ---- eigen! around the Schur factorization (pirates.jl:63-160) ---- */

/* Balancing: balance!(A; scale, permute) — balance.jl:33-199 (xGEBAL's
 * permutation search and power-of-2 scaling, p = 1, algo = :pr).  A (n x n,
 * lda) is balanced in place.  out4 receives {ilo, ihi, trivial, status}
 * (1-based ilo/ihi as Balancer.ilo/ihi; status 0, -3 when a NaN was met
 * while scaling — the reference's error("NaN encountered while balancing")
 * — or -4 when the scaling sweeps did not settle); sp (n int32) the
 * permutation record, 1-based (Balancer.prow = sp[0 .. ilo-2],
 * Balancer.pcol = sp[ihi .. n-1]); D (n doubles) the scale factors.
 * With GS_DEVICE_PTRS every array is device memory. */
int gs_balance_z(gs_handle* h, int64_t n, double* A, int64_t lda, int scale, int permute, int32_t* out4, int32_t* sp,
                 double* D, int flags);
int gs_balance_d(gs_handle* h, int64_t n, double* A, int64_t lda, int scale, int permute, int32_t* out4, int32_t* sp,
                 double* D, int flags);
/* The balancer's back-transform of the m complex columns of V (n x m, ldv):
 * lmul!(B, V) for right eigenvectors (inverse = 0) or ldiv!(B, V) for left
 * eigenvectors (inverse = 1) — balance.jl:201-261.  bal3 = {ilo, ihi,
 * trivial}, sp and D as returned by gs_balance_*. */
int gs_balance_apply(gs_handle* h, int64_t n, int64_t m, const int32_t* bal3, const int32_t* sp, const double* D,
                     double* V, int64_t ldv, int inverse, int flags);
/* triangularize(S) — triang.jl:9-43: the standard-form real Schur
 * factorization (T, Z) to the complex triangular one (Tc, Zc), W = diag(Tc)
 * (n complex).  Z may be NULL (then Zc is not referenced). */
int gs_triangularize_d(gs_handle* h, int64_t n, const double* T, int64_t ldt, const double* Z, int64_t ldz,
                       double* Tc, int64_t ldtc, double* Zc, int64_t ldzc, double* W, int flags);
/* _enormalize!(V) — util.jl:572-592: each of the m complex columns of V
 * scaled to unit 2-norm with its first largest entry real and positive. */
int gs_enormalize(gs_handle* h, int64_t n, int64_t m, double* V, int64_t ldv, int flags);
/* Reciprocal eigenvalue condition numbers of eigen!(A; jce = true) —
 * pirates.jl:112-120: rconde[j] = |vl_j^H vr_j| / ||vl_j|| / ||vr_j|| over
 * the m complex columns of VL and VR. */
int gs_eig_rcond(gs_handle* h, int64_t n, int64_t m, const double* VL, int64_t ldvl, const double* VR, int64_t ldvr,
                 double* rconde, int flags);

/* ---- reordering and conditioning of a complex Schur form ---- */

/* gordschur!(T, Z, select) — ordschur.jl:23-73 (complex): the eigenvalues
 * with select[k] != 0 (n int32, host memory) are moved to the front of the
 * triangular T (n x n complex, ldt), in their order, by adjacent swaps with
 * Givens rotations (_trexchange!, LinearAlgebra.givensAlgorithm), Z (ldz)
 * updated alike (NULL: no Z), T made exactly upper triangular, W (n complex)
 * = diag(T).  T and Z are overwritten. */
int gs_ordschur_z(gs_handle* h, int64_t n, double* T, int64_t ldt, double* Z, int64_t ldz, const int32_t* select,
                  double* W, int flags);
/* gordschur!(T, Z, select) for a real Schur form — rordschur.jl:3-45, with
 * _trexchange! (:54-165, LAPACK xTREXC), _swap1or2 (:168-297, xLAEXC) and
 * _syl1or2 (:300-466, xLASY2): the blocks (1 x 1, or standard 2 x 2 pairs,
 * selected when either member is) with select[k] != 0 (n int32, host
 * memory) moved to the front of the quasi-triangular T (n x n, ldt) by
 * adjacent block swaps, Z (ldz; NULL: none) updated alike, the swapped
 * 2 x 2 blocks restandardized (_gs2x2!); W (n complex) the eigenvalues.
 * info[0] (host memory) = 0, or k when the swap moving the block at k was
 * refused as too ill-conditioned (IllConditionException(k)); T and Z then
 * hold the swaps done before it and W is not written.  Bit for bit the
 * reference's operation order (its BLAS calls in the reference BLAS
 * order). */
int gs_ordschur_d(gs_handle* h, int64_t n, double* T, int64_t ldt, double* Z, int64_t ldz, const int32_t* select,
                  double* W, int32_t* info, int flags);
/* eigvalscond(S, nsub) and subspacesep(S, nsub) — ordschur.jl:85-131 with
 * trsylvester!/adjtrsylvester! (sylvester.jl:24-139) and norm1est!
 * (norm1est.jl:21-100): out2[0] = the reciprocal condition number of the
 * nsub leading eigenvalues of the triangular T, out2[1] = the estimated
 * separation of its leading invariant subspace; 1 <= nsub <= n - 1
 * (GS_E_ARG otherwise; the reference's estimator is undefined there).
 * out2 is host memory. */
int gs_eigcond_z(gs_handle* h, int64_t n, const double* T, int64_t ldt, int64_t nsub, double* out2, int flags);

/* ---- Hermitian / real symmetric eigenproblem (QR iteration) ---- */

/* geigen!(A::Hermitian, QRIteration()) / geigvals!(A) — GenericSchur.jl:
 * 124-146: _hehessenberg! (hessenberg.jl:27-95, LAPACK xHETD2 with
 * rank2update!), _materializeQ (hessenberg.jl:133-232), gschur!(H, V)
 * (symtridiag.jl:14-41) with the QL/QR iteration _gschur!(SymTridiagonal,
 * QRIteration(), V) (symtridiag.jl:77-427, LAPACK xSTEQR), and, with
 * GS_SORT, sorteig!(λ, V, eigsortby) (a stable ascending sort).
 * `count` matrices of order n (A: lda, strideA elements between matrices;
 * only the lower (upper = 0) or upper (upper = 1) triangle is read; A is not
 * modified).  W receives count x n eigenvalues (stride n doubles); with
 * GS_WANTZ, V (ldv, strideV) the eigenvectors (complex for _z).  info[k]
 * (may be NULL): 0, or the number of unconverged off-diagonal entries when
 * the iteration stopped after 30 n sweeps (the reference's loop has no
 * limit).  n <= gs_heev_max_n() = 4096 (GS_E_UNSUPPORTED above): one
 * workgroup per matrix with the matrix in LDS up to 96 (complex) / 128
 * (real); above, one matrix at a time over the whole GPU with the working
 * matrices in device memory held by the handle.  Host pointers are staged
 * through device buffers; with GS_DEVICE_PTRS, A, W, V and info are device
 * memory. */
int gs_heev_z(gs_handle* h, int64_t count, int64_t n, const double* A, int64_t lda, int64_t strideA, int upper,
              double* W, double* V, int64_t ldv, int64_t strideV, int32_t* info, int flags);
int gs_heev_d(gs_handle* h, int64_t count, int64_t n, const double* A, int64_t lda, int64_t strideA, int upper,
              double* W, double* V, int64_t ldv, int64_t strideV, int32_t* info, int flags);
int gs_heev_max_n(int is_complex);

/* _hehessenberg!(Hermitian(A, uplo)) + _materializeQ(H) — hessenberg.jl:
 * 19-95, 133-232: the real symmetric tridiagonal matrix (d: n, e: n - 1) and
 * Q (n x n, ldq; NULL: not formed) with A = Q T Q^H, from the lower (upper =
 * 0) or upper triangle of A (not modified).  n <= gs_heev_max_n(). */
int gs_hetrd_z(gs_handle* h, int64_t n, const double* A, int64_t lda, int upper, double* d, double* e, double* Q,
               int64_t ldq, int flags);
int gs_hetrd_d(gs_handle* h, int64_t n, const double* A, int64_t lda, int upper, double* d, double* e, double* Q,
               int64_t ldq, int flags);

/* _gschur!(SymTridiagonal(d, e), QRIteration(), Z) — symtridiag.jl:77-427
 * (the step of gschur!(H::Hessenberg{SymTridiagonal}, Z), symtridiag.jl:
 * 14-41, and of geigen!(::SymTridiagonal), :43-49): d (n) is overwritten by
 * the eigenvalues (unsorted), e (n - 1) destroyed, and the zrows x n matrix Z
 * (ldz; complex when zcomplex; NULL: eigenvalues only) multiplied by the
 * rotations, bit for bit as the reference's operation order gives them.
 * info[0] as for gs_heev_z.  n <= 4096: above 2048 the rotations pass
 * through a record stream in device memory held by the handle, replayed
 * onto Z's rows over the whole GPU (the host waits once per stream round;
 * the results are the same bit for bit). */
int gs_steqr(gs_handle* h, int64_t n, double* d, double* e, double* Z, int64_t ldz, int64_t zrows, int zcomplex,
             int32_t* info, int flags);

/* Statistics of the last Schur call on this handle: out[0] = QR sweeps,
 * out[1] = bulge steps, out[2] = kernel time in microseconds (events on the
 * handle's stream, 0 when unavailable).  Returns the number written. */
int gs_last_stats(const gs_handle* h, int64_t* out, int len);

/* Per-launch records of the last Schur call made with GS_TIMING
 * (measurement support; no counterpart in the reference).  Each record is
 * GS_LAUNCH_REC doubles {kind, a, b, c, d, microseconds} (HIP events on the
 * launch's stream):
 *   kind 0 / 1  batched Hessenberg + Q kernel / QR kernel of one size bin:
 *               a, b = smallest and largest order, c = matrices, d = sum n^3
 *   kind 2      GEMM of the large real path: a, b, c = m, n, k, d = 2 m n k
 *   kind 3      phase of the large real path: a = 0 copy + scale,
 *               1 Hessenberg, 2 Q formation, 3 QR iterations, 4 finish
 *   kind 4      batched complex: the Z log applied after a bin's QR kernel
 *               (fields as kind 0 / 1)
 * Writes at most max_records records and returns the number available. */
#define GS_LAUNCH_REC 6
int gs_last_launches(const gs_handle* h, double* out, int max_records);

/* C(m x n) = beta C + alpha op(A) op(B) on the handle's stream with the
 * FP64 MFMA GEMM kernel of the large real path (column-major device
 * pointers; op = transpose when ta / tb is nonzero).  Support routine with
 * no counterpart in the reference: it exposes the GEMM for measurement. */
int gs_dgemm(gs_handle* h, int ta, int tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
             int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc);

#ifdef __cplusplus
}
#endif

#endif /* GSCHUR_MI355X_H */
