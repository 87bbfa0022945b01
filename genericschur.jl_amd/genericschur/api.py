"""GenericSchur-style API over the MI355X library.

Mirrors the entry points of GenericSchur.jl that the Schur hot path serves:

==========================  ====================================================
this module                 reference
==========================  ====================================================
``gschur_(A, ...)``         ``gschur!(A; wantZ, scale, maxiter, maxinner)``
                            (GenericSchur.jl:350-372 complex, :805-835 real)
``gschur(A, ...)``          ``gschur(A)`` (GenericSchur.jl:343)
``schur_(A, ...)``          ``LinearAlgebra.schur!(A)`` pirate (pirates.jl:8-10)
``gschur_hessenberg_(H,Z)`` ``gschur!(H::Hessenberg, Z)`` (GenericSchur.jl:194,513)
``eigvals_(A, sortby)``     ``LinearAlgebra.eigvals!(A)`` pirate (pirates.jl:17-27)
``geigvecs(S, left)``       ``geigvecs(S; left)`` (vectors.jl:12-20)
``eigvecs(S, left)``        ``eigvecs(S::Schur; left)`` pirate (pirates.jl:38-48)
``eigen_(A, ...)``          ``LinearAlgebra.eigen!(A; permute, scale, sortby,
                            jvl, jvr, jce)`` pirate (pirates.jl:63-160)
``balance_(A, ...)``        ``balance!(A; scale, permute)`` (balance.jl:33-199)
``lmul_(B, V)``, ``ldiv_``  ``lmul!/ldiv!(B::Balancer, V)`` (balance.jl:201-261)
``triangularize(S)``        ``triangularize(S::Schur{<:Real})`` (triang.jl:9-43)
``enormalize_(V)``          ``_enormalize!(V)`` (util.jl:572-592)
``ordschur_(S, select)``    ``gordschur!`` / ``ordschur!`` (ordschur.jl:23-73
                            complex, rordschur.jl:3-45 real)
``eigvalscond``,            ``eigvalscond(S, nsub)``, ``subspacesep(S, nsub)``
``subspacesep``             (ordschur.jl:85-131)
``geigen_hermitian(A)``     ``geigen!(A::Hermitian, QRIteration())``
                            (GenericSchur.jl:124-135), ``geigvals_hermitian``
                            (:137-146), ``heev_batched`` (a batch of them)
``hessenberg_hermitian``    ``_hessenberg!(Hermitian(A))`` + ``_materializeQ``
                            (hessenberg.jl:19-232)
``steqr_(d, e, Z)``         ``_gschur!(SymTridiagonal, QRIteration(), Z)``
                            (symtridiag.jl:77-427); ``gschur_symtridiag``,
                            ``geigen_symtridiag``, ``geigvals_symtridiag``
                            (symtridiag.jl:14-55)
``Schur``                   ``LinearAlgebra.Schur`` (fields T, Z, values)
``UnconvergedException``    ``GenericSchur.UnconvergedException``
``IllConditionException``   ``GenericSchur.IllConditionException``
``DimensionMismatch``       ``Base.DimensionMismatch``
``ArgumentError``           ``Base.ArgumentError``
==========================  ====================================================

Matrices are numpy arrays indexed ``A[i, j]`` as in Julia; Float64 input
takes the real double-shift path, ComplexF64 input the complex single-shift
path.  Batched entry points take lists of matrices (host) or packed device
tensors.  Every computation runs in the HIP library; nothing here computes a
decomposition on the CPU.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import (GS_CHECK_SUBDIAG, GS_DEVICE_PTRS, GS_HESSENBERG, GS_PLAIN, GS_SCALE, GS_WANTZ, GSError)


class UnconvergedException(Exception):
    """Iteration limit reached (GenericSchur.jl:39-45)."""

    def __init__(self, msg):
        super().__init__("Convergence failure; " + msg)
        self.msg = msg


class DimensionMismatch(ValueError):
    pass


class ArgumentError(ValueError):
    pass


class Schur:
    """Schur factorization A = Z T Z^H (LinearAlgebra.Schur: T, Z, values)."""

    __slots__ = ("T", "Z", "values")

    def __init__(self, T, Z, values):
        self.T = T
        self.Z = Z
        self.values = values

    @property
    def vectors(self):
        return self.Z

    @property
    def Schur(self):
        return self.T

    def __iter__(self):
        return iter((self.T, self.Z, self.values))

    def __repr__(self):
        return f"Schur(T={self.T.shape} {self.T.dtype}, Z={self.Z.shape}, values={self.values.shape})"


def _checksquare(A):
    if A.ndim != 2 or A.shape[0] != A.shape[1]:
        raise DimensionMismatch(f"matrix is not square: dimensions are {A.shape}")
    return A.shape[0]


def _kind(A):
    if A.dtype == np.complex128:
        return True
    if A.dtype == np.float64:
        return False
    raise TypeError(f"element type {A.dtype} is not served by this library (Float64 / ComplexF64 only)")


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _raise_info(info, maxiter, n):
    if info == 0:
        return
    if info > 0:
        mi = 100 * n if maxiter is None else maxiter
        raise UnconvergedException(f"iteration limit {mi} reached")
    if info == -1:
        raise DimensionMismatch("bad dimension")
    if info == -2:
        raise ArgumentError("algorithm assumes real subdiagonal")
    raise GSError(f"info = {info}")


def _single(A, wantZ, scale, maxiter, maxinner, hess=False, Z0=None, checksd=False, plain=False, device=0,
            timing=False, extra=0):
    cplx = _kind(A)
    n = A.shape[0]
    dt = A.dtype
    a = np.array(A, dtype=dt, order="F", copy=True)
    T = np.zeros((n, n), dtype=dt, order="F")
    Z = np.zeros((n, n), dtype=dt, order="F") if Z0 is None else np.array(Z0, dtype=dt, order="F", copy=True)
    W = np.zeros(n, dtype=np.complex128)
    info = ctypes.c_int32(0)
    flags = (GS_WANTZ if wantZ else 0) | (GS_SCALE if scale else 0) | extra
    if hess:
        flags |= GS_HESSENBERG
        if checksd:
            flags |= GS_CHECK_SUBDIAG
    if plain:
        flags |= GS_PLAIN
    if timing:
        flags |= _lib.GS_TIMING
    h = _lib.handle(device)
    f = _lib.lib().gs_schur_z if cplx else _lib.lib().gs_schur_d
    ld = max(n, 1)
    rc = f(h, n, _ptr(a), ld, _ptr(T), ld, _ptr(Z), ld, _ptr(W), flags,
           0 if maxiter is None else int(maxiter), 0 if maxinner is None else int(maxinner), ctypes.byref(info))
    if rc == _lib.GS_E_DIM:
        raise DimensionMismatch("bad dimension")
    _lib.check(rc, h, "gs_schur")
    _raise_info(info.value, maxiter, n)
    Zo = Z if wantZ else np.zeros((0, 0), dtype=dt)
    return Schur(T, Zo, W)


def gschur_(A, wantZ=True, scale=True, maxiter=None, maxinner=None, plain=False, device=0, timing=False,
            zdirect=False, zlog_tiny=False):
    """gschur!(A) -> Schur (GenericSchur.jl:350-372 / :805-835).

    Real matrices above the LDS-resident orders take the blocked large-matrix
    path unless plain=True (the unblocked single-workgroup kernel).  With
    timing=True the large path records per-phase and per-GEMM durations
    (genericschur._lib.last_launches()).  Complex matrices above the
    LDS-resident orders log Z's updates and apply the log after the sweeps;
    zdirect=True updates Z during the sweeps instead, zlog_tiny=True makes
    the log hold one sweep (comparisons; the results are the same)."""
    if _is_torch(A):
        raise TypeError("use gschur_device / gschur_batched_device for device tensors")
    A = np.asarray(A)
    _checksquare(A)
    extra = (_lib.GS_ZDIRECT if zdirect else 0) | (_lib.GS_ZLOG_TINY if zlog_tiny else 0)
    return _single(A, wantZ, scale, maxiter, maxinner, plain=plain, device=device, timing=timing, extra=extra)


def gschur_device(n, A, T, Z, W, complex_input, wantZ=True, scale=True, plain=False, timing=False, maxiter=None,
                  device=0):
    """gschur! on one device-resident column-major matrix (torch CUDA tensors
    or raw device addresses, ld = n): T, Z (n x n) and W (n complex) are
    written on the device.  Returns info (0, or > 0 when unconverged)."""
    def addr(x):
        return ctypes.c_void_p(x.data_ptr() if _is_torch(x) else int(x))

    flags = (GS_WANTZ if wantZ else 0) | (GS_SCALE if scale else 0) | GS_DEVICE_PTRS
    if plain:
        flags |= GS_PLAIN
    if timing:
        flags |= _lib.GS_TIMING
    info = ctypes.c_int32(0)
    h = _lib.handle(device)
    f = _lib.lib().gs_schur_z if complex_input else _lib.lib().gs_schur_d
    ld = max(int(n), 1)
    rc = f(h, int(n), addr(A), ld, addr(T), ld, addr(Z) if Z is not None else None, ld, addr(W), flags,
           0 if maxiter is None else int(maxiter), 0, ctypes.byref(info))
    _lib.check(rc, h, "gs_schur (device)")
    return info.value


def dgemm_device(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, device=0):
    """C = beta C + alpha op(A) op(B) with the library's FP64 MFMA GEMM
    (device pointers, column-major)."""
    def addr(x):
        return ctypes.c_void_p(x.data_ptr() if _is_torch(x) else int(x))

    h = _lib.handle(device)
    rc = _lib.lib().gs_dgemm(h, int(bool(ta)), int(bool(tb)), m, n, k, float(alpha), addr(A), lda, addr(B), ldb,
                             float(beta), addr(C), ldc)
    _lib.check(rc, h, "gs_dgemm")


def gschur(A, **kw):
    """gschur(A) = gschur!(copy(A)) (GenericSchur.jl:343)."""
    return gschur_(np.array(A, copy=True), **kw)


schur_ = gschur_


def gschur_hessenberg_(H, Z=None, maxiter=None, maxinner=None, checksd=True, device=0):
    """gschur!(H::Hessenberg, Z) (GenericSchur.jl:194-335 / :513-699).

    H is upper Hessenberg (entries below the subdiagonal are ignored); Z, if
    given, is the n x n transformation to update.  Returns Schur.
    """
    H = np.asarray(H)
    n = _checksquare(H)
    if Z is not None:
        Z = np.asarray(Z)
        if Z.ndim != 2 or Z.shape[1] != n or Z.shape[0] != n:
            raise DimensionMismatch("second dimension of Z must match H")
        if Z.dtype != H.dtype:
            Z = Z.astype(H.dtype)
    cplx = _kind(H)
    return _single(H, Z is not None, False, maxiter, maxinner, hess=True, Z0=Z, checksd=(checksd and cplx),
                   device=device)


def eigvals_(A, sortby="eigsortby", **kw):
    """eigvals!(A) (pirates.jl:17-27): gschur!(A; wantZ=false), sorted by (re, im)."""
    S = gschur_(A, wantZ=False, **kw)
    v = S.values
    if sortby is None:
        return v
    if sortby == "eigsortby":
        return v[np.lexsort((v.imag, v.real))]
    return np.array(sorted(v, key=sortby))


def geigvecs(S, left=False, normalize=True, device=0):
    """geigvecs(S; left) (vectors.jl:12-20): eigenvectors as columns, ordered
    like S.values, unit 2-norm with the largest component real."""
    T = np.asarray(S.T)
    n = _checksquare(T)
    cplx = _kind(T)
    t = np.array(T, dtype=T.dtype, order="F", copy=True)
    Z = np.asarray(S.Z)
    z = None if Z.size == 0 else np.array(Z, dtype=T.dtype, order="F", copy=True)
    V = np.zeros((n, n), dtype=np.complex128, order="F")
    h = _lib.handle(device)
    f = _lib.lib().gs_trevc_z if cplx else _lib.lib().gs_trevc_d
    ld = max(n, 1)
    rc = f(h, n, _ptr(t), ld, None if z is None else _ptr(z), ld, _ptr(V), ld, int(bool(left)), int(bool(normalize)), 0)
    _lib.check(rc, h, "gs_trevc")
    return V


def eigvecs(S, left=False, device=0):
    """eigvecs(S::Schur; left) (pirates.jl:38-48)."""
    return geigvecs(S, left=left, device=device)


# ------------------------------------------------------------------ batched
def pack(mats, dtype):
    """Pack square matrices column-major, back to back.  Returns (n, off, flat)."""
    ns = np.array([m.shape[0] for m in mats], dtype=np.int32)
    off = np.zeros(len(mats), dtype=np.int64)
    if len(mats) > 1:
        off[1:] = np.cumsum(ns[:-1].astype(np.int64) ** 2)
    total = int(off[-1] + int(ns[-1]) ** 2) if len(mats) else 0
    flat = np.zeros(total, dtype=dtype)
    for m, o, k in zip(mats, off, ns):
        flat[o:o + k * k] = np.asarray(m, dtype=dtype).reshape(-1, order="F")
    return ns, off, flat


def unpack(ns, off, flat):
    return [flat[o:o + k * k].reshape((k, k), order="F") for k, o in zip(ns, off)]


def eig_offsets(ns):
    w = np.zeros(len(ns), dtype=np.int64)
    if len(ns) > 1:
        w[1:] = np.cumsum(np.asarray(ns[:-1], dtype=np.int64))
    return w


def gschur_batched(mats, wantZ=True, scale=True, device=0, raise_on_info=True, qr_band=False, zdirect=False,
                   zlog_tiny=False):
    """Batched gschur! of host matrices (all Float64 or all ComplexF64) in one
    graph launch.  Returns a list of Schur (and the info array).  qr_band
    selects the band-resident complex QR kernel instead of the windowed one
    (same arithmetic; for comparisons).  Complex, windowed: zdirect updates Z
    during the sweeps instead of through the Z log; zlog_tiny gives each
    matrix a log of one sweep, so the QR kernel applies it itself (tests)."""
    mats = [np.asarray(m) for m in mats]
    if not mats:
        return [], np.zeros(0, dtype=np.int32)
    for m in mats:
        _checksquare(m)
    cplx = _kind(mats[0])
    if any(_kind(m) != cplx for m in mats):
        raise TypeError("mixed element types in one batch")
    dt = np.complex128 if cplx else np.float64
    ns, off, flat = pack(mats, dt)
    T = np.zeros_like(flat)
    Z = np.zeros_like(flat)
    W = np.zeros(int(ns.sum()), dtype=np.complex128)
    info = np.zeros(len(mats), dtype=np.int32)
    flags = (GS_WANTZ if wantZ else 0) | (GS_SCALE if scale else 0) | (_lib.GS_QR_BAND if qr_band else 0)
    flags |= (_lib.GS_ZDIRECT if zdirect else 0) | (_lib.GS_ZLOG_TINY if zlog_tiny else 0)
    h = _lib.handle(device)
    f = _lib.lib().gs_schur_batched_z if cplx else _lib.lib().gs_schur_batched_d
    rc = f(h, len(mats), _ptr(ns), _ptr(off), _ptr(flat), _ptr(T), _ptr(Z), _ptr(W), _ptr(info), flags)
    if rc == _lib.GS_E_DIM:
        raise DimensionMismatch("matrix order outside the batched range")
    _lib.check(rc, h, "gs_schur_batched")
    woff = eig_offsets(ns)
    out = []
    for i, (k, o) in enumerate(zip(ns, off)):
        if raise_on_info:
            _raise_info(int(info[i]), None, int(k))
        Ti = T[o:o + k * k].reshape((k, k), order="F").copy()
        Zi = Z[o:o + k * k].reshape((k, k), order="F").copy() if wantZ else np.zeros((0, 0), dtype=dt)
        out.append(Schur(Ti, Zi, W[woff[i]:woff[i] + k].copy()))
    return out, info


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def gschur_batched_device(ns, off, A, T, Z, W, info, complex_input, wantZ=True, scale=True, asynchronous=False,
                          device=0, timing=False, zdirect=False):
    """Batched gschur! on device-resident packed data (torch CUDA tensors or
    raw device addresses).  ns/off are host int32/int64 arrays.  With timing,
    the batch runs eagerly with events around each launch (read them with
    genericschur._lib.last_launches())."""
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    off = np.ascontiguousarray(off, dtype=np.int64)

    def addr(x):
        if x is None:
            return None
        return ctypes.c_void_p(x.data_ptr() if _is_torch(x) else int(x))

    flags = (GS_WANTZ if wantZ else 0) | (GS_SCALE if scale else 0) | GS_DEVICE_PTRS
    if asynchronous:
        flags |= _lib.GS_ASYNC
    if timing:
        flags |= _lib.GS_TIMING
    if zdirect:
        flags |= _lib.GS_ZDIRECT
    h = _lib.handle(device)
    f = _lib.lib().gs_schur_batched_z if complex_input else _lib.lib().gs_schur_batched_d
    rc = f(h, len(ns), _ptr(ns), _ptr(off), addr(A), addr(T), addr(Z), addr(W), addr(info), flags)
    _lib.check(rc, h, "gs_schur_batched")


def set_stream(stream, device=0):
    """Run the library's work on `stream` (a hipStream_t address, e.g.
    torch.cuda.current_stream().cuda_stream); None restores its own stream."""
    h = _lib.handle(device)
    _lib.check(_lib.lib().gs_set_stream(h, ctypes.c_void_p(stream) if stream else None), h, "gs_set_stream")


def batched_max_n(complex_input):
    """Largest order the LDS-resident batched kernels take (gs_batched_max_n)."""
    return int(_lib.lib().gs_batched_max_n(1 if complex_input else 0))


def synchronize(device=0):
    h = _lib.handle(device)
    _lib.check(_lib.lib().gs_synchronize(h), h, "gs_synchronize")


# ------------------------------------------------------------------ eigen!
class Balancer:
    """Balancer (balance.jl:3-10): ilo, ihi (1-based), prow, pcol (1-based
    permutation records), D (scale factors), trivial."""

    __slots__ = ("ilo", "ihi", "sp", "D", "trivial")

    def __init__(self, ilo, ihi, sp, D, trivial):
        self.ilo, self.ihi, self.sp, self.D, self.trivial = ilo, ihi, sp, D, trivial

    @property
    def prow(self):
        return self.sp[: self.ilo - 1]

    @property
    def pcol(self):
        return self.sp[self.ihi:]

    def __repr__(self):
        return f"Balancer(ilo={self.ilo}, ihi={self.ihi}, trivial={self.trivial})"


class Eigen:
    """LinearAlgebra.Eigen: values, vectors, and (eigen!'s jvl / jce options)
    vectorsl, rconde."""

    __slots__ = ("values", "vectors", "vectorsl", "rconde", "rcondv")

    def __init__(self, values, vectors, vectorsl=None, rconde=None, rcondv=None):
        self.values, self.vectors, self.vectorsl, self.rconde, self.rcondv = values, vectors, vectorsl, rconde, rcondv

    def __iter__(self):
        return iter((self.values, self.vectors))


def balance_(A, scale=True, permute=True, device=0):
    """balance!(A; scale, permute) -> (Abal, Balancer) (balance.jl:33-199).
    A (numpy, Float64 or ComplexF64) is balanced on the GPU; the input array
    is not modified (a balanced copy is returned)."""
    A = np.asarray(A)
    n = _checksquare(A)
    cplx = _kind(A)
    a = np.array(A, dtype=A.dtype, order="F", copy=True)
    out = np.zeros(4, dtype=np.int32)
    sp = np.zeros(max(n, 1), dtype=np.int32)
    D = np.ones(max(n, 1))
    h = _lib.handle(device)
    f = _lib.lib().gs_balance_z if cplx else _lib.lib().gs_balance_d
    rc = f(h, n, _ptr(a), max(n, 1), int(bool(scale)), int(bool(permute)), _ptr(out), _ptr(sp), _ptr(D), 0)
    _lib.check(rc, h, "gs_balance")
    if out[3] == -3:
        raise RuntimeError("NaN encountered while balancing")
    if out[3] != 0:
        raise GSError(f"balancing did not settle (status {out[3]})")
    return a, Balancer(int(out[0]), int(out[1]), sp[:n].copy(), D[:n].copy(), bool(out[2]))


def _bal_apply(B, V, inverse, device):
    V = np.asarray(V)
    if V.ndim != 2:
        raise DimensionMismatch("eigenvector matrix expected")
    n, m = V.shape
    v = np.array(V, dtype=np.complex128, order="F", copy=True)
    b3 = np.array([B.ilo, B.ihi, int(B.trivial)], dtype=np.int32)
    sp = np.ascontiguousarray(B.sp, dtype=np.int32)
    D = np.ascontiguousarray(B.D, dtype=np.float64)
    if len(sp) != n or len(D) != n:
        raise DimensionMismatch("balancer and vectors differ in size")
    h = _lib.handle(device)
    rc = _lib.lib().gs_balance_apply(h, n, m, _ptr(b3), _ptr(sp), _ptr(D), _ptr(v), max(n, 1), int(inverse), 0)
    _lib.check(rc, h, "gs_balance_apply")
    return v


def lmul_(B, V, device=0):
    """lmul!(B::Balancer, V): back-transform of right eigenvectors
    (balance.jl:201-230).  Returns the transformed (complex) copy."""
    return _bal_apply(B, V, False, device)


def ldiv_(B, V, device=0):
    """ldiv!(B::Balancer, V): back-transform of left eigenvectors
    (balance.jl:233-261).  Returns the transformed (complex) copy."""
    return _bal_apply(B, V, True, device)


def triangularize(S, device=0):
    """triangularize(S::Schur{<:Real}) -> complex Schur (triang.jl:9-43)."""
    T = np.asarray(S.T)
    n = _checksquare(T)
    if T.dtype != np.float64:
        raise TypeError("triangularize takes a real Schur factorization")
    t = np.array(T, order="F", copy=True)
    Z = np.asarray(S.Z)
    z = None if Z.size == 0 else np.array(Z, dtype=np.float64, order="F", copy=True)
    Tc = np.zeros((n, n), dtype=np.complex128, order="F")
    Zc = np.zeros((n, n), dtype=np.complex128, order="F") if z is not None else np.zeros((0, 0), dtype=np.complex128)
    W = np.zeros(n, dtype=np.complex128)
    h = _lib.handle(device)
    ld = max(n, 1)
    rc = _lib.lib().gs_triangularize_d(h, n, _ptr(t), ld, None if z is None else _ptr(z), ld, _ptr(Tc), ld,
                                       None if z is None else _ptr(Zc), ld, _ptr(W), 0)
    _lib.check(rc, h, "gs_triangularize_d")
    return Schur(Tc, Zc, W)


# ------------------------------------------- reordering and conditioning
class IllConditionException(Exception):
    """A block swap of the real reordering was refused (rordschur.jl:32;
    GenericSchur.IllConditionException): k is the 1-based position."""

    def __init__(self, k):
        super().__init__(f"reordering refused at block {k}")
        self.k = k


def _ordschur_real(S, sel, n, device):
    t = np.array(S.T, dtype=np.float64, order="F", copy=True)
    Z = np.asarray(S.Z)
    z = None if Z.size == 0 else np.array(Z, dtype=np.float64, order="F", copy=True)
    W = np.zeros(n, dtype=np.complex128)
    info = np.zeros(1, dtype=np.int32)
    h = _lib.handle(device)
    ld = max(n, 1)
    rc = _lib.lib().gs_ordschur_d(h, n, _ptr(t), ld, None if z is None else _ptr(z), ld, _ptr(sel), _ptr(W),
                                  _ptr(info), 0)
    _lib.check(rc, h, "gs_ordschur_d")
    if info[0]:
        raise IllConditionException(int(info[0]))
    return Schur(t, z if z is not None else np.zeros((0, 0)), W)


def ordschur_(S, select, device=0):
    """gordschur!(S, select): the selected eigenvalues first, in order.
    Complex Schur factorization (ordschur.jl:23-73): a new Schur with T
    exactly upper triangular, Z updated, values = diag(T).  Real (rordschur.jl:
    3-45): blocks move whole (a 2 x 2 pair is selected when either member
    is), the moved 2 x 2 blocks are restandardized; raises
    IllConditionException(k) when a swap is refused."""
    T = np.asarray(S.T)
    n = _checksquare(T)
    sel = np.ascontiguousarray(np.asarray(select, dtype=bool).astype(np.int32))
    if sel.shape != (n,):
        raise ValueError("select must have length n")
    if T.dtype == np.float64:
        return _ordschur_real(S, sel, n, device)
    if T.dtype != np.complex128:
        raise TypeError("ordschur_ takes a Float64 or ComplexF64 Schur factorization")
    t = np.array(T, order="F", copy=True)
    Z = np.asarray(S.Z)
    z = None if Z.size == 0 else np.array(Z, dtype=np.complex128, order="F", copy=True)
    W = np.zeros(n, dtype=np.complex128)
    h = _lib.handle(device)
    ld = max(n, 1)
    rc = _lib.lib().gs_ordschur_z(h, n, _ptr(t), ld, None if z is None else _ptr(z), ld, _ptr(sel), _ptr(W), 0)
    _lib.check(rc, h, "gs_ordschur_z")
    return Schur(t, z if z is not None else np.zeros((0, 0), complex), W)


def eigcond(S, nsub, device=0):
    """(eigvalscond(S, nsub), subspacesep(S, nsub)) of a Schur factorization
    (ordschur.jl:85-131), 1 <= nsub <= n - 1.  The reference's solvers
    (trsylvester!, adjtrsylvester!, norm1est!) are generic in the element
    type and read T as triangular: a Float64 T goes through the same kernel
    as ComplexF64 with zero imaginary parts, whose complex products, Smith
    divisions, moduli and sign steps then reduce exactly to the real
    operations."""
    T = np.asarray(S.T)
    n = _checksquare(T)
    if T.dtype not in (np.complex128, np.float64):
        raise TypeError("eigcond takes a Float64 or ComplexF64 Schur factorization")
    t = np.array(T, dtype=np.complex128, order="F", copy=True)
    out = np.zeros(2)
    h = _lib.handle(device)
    rc = _lib.lib().gs_eigcond_z(h, n, _ptr(t), max(n, 1), int(nsub), _ptr(out), 0)
    _lib.check(rc, h, "gs_eigcond_z")
    return float(out[0]), float(out[1])


def eigvalscond(S, nsub, device=0):
    """eigvalscond(S, nsub) (ordschur.jl:85-98)."""
    return eigcond(S, nsub, device)[0]


def subspacesep(S, nsub, device=0):
    """subspacesep(S, nsub) (ordschur.jl:111-131)."""
    return eigcond(S, nsub, device)[1]


def enormalize_(V, device=0):
    """_enormalize!(V) (util.jl:572-592): unit 2-norm columns, the first
    largest entry of each real and positive.  Returns the normalised copy."""
    V = np.asarray(V)
    n, m = V.shape
    v = np.array(V, dtype=np.complex128, order="F", copy=True)
    h = _lib.handle(device)
    rc = _lib.lib().gs_enormalize(h, n, m, _ptr(v), max(n, 1), 0)
    _lib.check(rc, h, "gs_enormalize")
    return v


def _eig_rcond(VL, VR, device):
    n, m = VR.shape
    vl = np.array(VL, dtype=np.complex128, order="F", copy=True)
    vr = np.array(VR, dtype=np.complex128, order="F", copy=True)
    rc = np.zeros(m)
    h = _lib.handle(device)
    st = _lib.lib().gs_eig_rcond(h, n, m, _ptr(vl), max(n, 1), _ptr(vr), max(n, 1), _ptr(rc), 0)
    _lib.check(st, h, "gs_eig_rcond")
    return rc


def eigen_(A, permute=True, scale=True, sortby="eigsortby", jvl=False, jvr=True, jce=False, device=0):
    """eigen!(A; permute, scale, sortby, jvl, jvr, jce) (pirates.jl:63-160).

    Balancing (balance!), gschur!, triangularize for real input, the complex
    right (and left) eigenvector solves of geigvecs without normalisation,
    the balancer's back-transform, _enormalize!, and sorteig! by
    eigsortby = (real, imag).  With jce the reciprocal eigenvalue condition
    numbers of the balanced form are returned in rconde.  Every numerical
    step runs on the GPU; only the final sort permutation is host logic."""
    A = np.asarray(A)
    n = _checksquare(A)
    cplx = _kind(A)
    if permute or scale:
        Ab, B = balance_(A, scale=scale, permute=permute, device=device)
    else:
        Ab, B = np.array(A, copy=True), None
    S = gschur_(Ab, device=device)
    if not cplx:
        S = triangularize(S, device=device)
    vals = S.values
    do_vr = jvr or jce
    do_vl = jvl or jce
    v = geigvecs(S, left=False, normalize=False, device=device) if do_vr else np.zeros((0, 0), dtype=np.complex128)
    vl = geigvecs(S, left=True, normalize=False, device=device) if do_vl else np.zeros((0, 0), dtype=np.complex128)
    rconde = _eig_rcond(vl, v, device) if jce else None
    if jvr:
        if B is not None:
            v = lmul_(B, v, device=device)
        v = enormalize_(v, device=device)
    else:
        v = np.zeros((0, 0), dtype=np.complex128)
    if jvl:
        if B is not None:
            vl = ldiv_(B, vl, device=device)
        vl = enormalize_(vl, device=device)
    else:
        vl = None
    if sortby is not None:
        if sortby == "eigsortby":
            p = np.lexsort((vals.imag, vals.real))
        else:
            p = np.array(sorted(range(n), key=lambda i: sortby(vals[i])), dtype=np.int64)
        vals = vals[p]
        if v.size:
            v = v[:, p]
        if vl is not None:
            vl = vl[:, p]
        if rconde is not None:
            rconde = rconde[p]
    return Eigen(vals, v, vl, rconde)


# ------------------------------------------------- Hermitian / real symmetric


def _uplo(uplo):
    if uplo not in ("L", "U"):
        raise ArgumentError("uplo must be 'L' or 'U'")
    return int(uplo == "U")


def heev_max_n(complex_input):
    """Largest order of the LDS-resident Hermitian eigensolver."""
    return int(_lib.lib().gs_heev_max_n(int(bool(complex_input))))


def hessenberg_hermitian(A, uplo="L", wantQ=True, device=0):
    """_hessenberg!(Hermitian(A, uplo)) and _materializeQ (hessenberg.jl:19-95,
    133-232): the real symmetric tridiagonal matrix (d, e) and the unitary Q
    with A = Q SymTridiagonal(d, e) Q^H, from the uplo triangle of A."""
    A = np.asarray(A)
    n = _checksquare(A)
    cplx = np.iscomplexobj(A)
    dt = np.complex128 if cplx else np.float64
    a = np.array(A, dtype=dt, order="F", copy=True)
    d = np.zeros(max(n, 1))
    e = np.zeros(max(n - 1, 1))
    q = np.zeros((max(n, 1), max(n, 1)), dtype=dt, order="F") if wantQ else None
    h = _lib.handle(device)
    f = _lib.lib().gs_hetrd_z if cplx else _lib.lib().gs_hetrd_d
    rc = f(h, n, _ptr(a), max(n, 1), _uplo(uplo), _ptr(d), _ptr(e), None if q is None else _ptr(q), max(n, 1), 0)
    _lib.check(rc, h, "gs_hetrd")
    return d[:n], e[:max(n - 1, 0)], (q[:n, :n] if wantQ else None)


def _sort_host(w, V, sortby):
    """sorteig!(λ, V, sortby) for a callable sortby (stable, as Julia's)."""
    keys = [sortby(x) for x in w]
    p = sorted(range(len(w)), key=lambda i: keys[i])
    w = w[p]
    return w, (V[:, p] if V is not None else None)


_MRRR_WARNED = [False]


def _check_alg(alg):
    """The Hermitian / tridiagonal algorithm argument (GenericSchur.jl:
    118-123, symtridiag.jl:14-29): "QRIteration" runs; "RobustRepresentations"
    falls back to QR iteration with a one-time warning, as the reference does
    when no MRRR implementation is registered (symtridiag.jl:21-28);
    "DivideAndConquer" (symtri_dc.jl) is not provided by this library."""
    if alg in (None, "QRIteration"):
        return
    if alg == "RobustRepresentations":
        if not _MRRR_WARNED[0]:
            import warnings
            warnings.warn("GenericSchur does not currently provide alg=RobustRepresentations; "
                          "falling back to QRIteration")
            _MRRR_WARNED[0] = True
        return
    if alg == "DivideAndConquer":
        raise NotImplementedError("alg=DivideAndConquer (symtri_dc.jl) is not provided by this library; "
                                  "use QRIteration")
    raise ArgumentError(f"unknown algorithm {alg!r}")


def heev_batched(A, uplo="L", wantV=True, sortby="eigsortby", device=0, raise_on_info=True, alg="QRIteration"):
    """geigen! (wantV) / geigvals! with QRIteration (GenericSchur.jl:124-146)
    over a stack A (count, n, n) of Hermitian (complex) or real symmetric
    matrices, one GPU launch: returns (W (count, n), V (count, n, n) or None,
    info).  sortby: "eigsortby" (ascending, on the GPU), None, or a callable
    (applied on the host afterwards)."""
    _check_alg(alg)
    A = np.asarray(A)
    if A.ndim != 3 or A.shape[1] != A.shape[2]:
        raise DimensionMismatch("A must be (count, n, n)")
    count, n = A.shape[0], A.shape[1]
    cplx = np.iscomplexobj(A)
    dt = np.complex128 if cplx else np.float64
    a = np.ascontiguousarray(np.transpose(A, (0, 2, 1)), dtype=dt)  # column-major per matrix
    W = np.zeros((count, n))
    V = np.zeros((count, n, n), dtype=dt) if wantV else None
    info = np.zeros(max(count, 1), dtype=np.int32)
    flags = (_lib.GS_WANTZ if wantV else 0) | (_lib.GS_SORT if sortby == "eigsortby" else 0)
    h = _lib.handle(device)
    f = _lib.lib().gs_heev_z if cplx else _lib.lib().gs_heev_d
    rc = f(h, count, n, _ptr(a), max(n, 1), n * n, _uplo(uplo), _ptr(W), None if V is None else _ptr(V), max(n, 1),
           n * n, _ptr(info), flags)
    _lib.check(rc, h, "gs_heev")
    if raise_on_info and count and info[:count].any():
        raise UnconvergedException(f"{int(info[:count].max())} eigenvalues failed to converge in {30 * n} sweeps")
    V = np.transpose(V, (0, 2, 1)) if wantV else None
    if callable(sortby):
        for k in range(count):
            w2, v2 = _sort_host(W[k], V[k] if wantV else None, sortby)
            W[k] = w2
            if wantV:
                V[k] = v2
    return W, V, info[:count]


def geigen_hermitian(A, uplo="L", sortby="eigsortby", device=0, alg="QRIteration"):
    """geigen!(Hermitian(A, uplo), QRIteration(); sortby) (GenericSchur.jl:
    124-135): Eigen(values, vectors)."""
    A = np.asarray(A)
    _checksquare(A)
    W, V, _ = heev_batched(A[None], uplo=uplo, wantV=True, sortby=sortby, device=device, alg=alg)
    return Eigen(W[0], V[0])


def geigvals_hermitian(A, uplo="L", sortby="eigsortby", device=0, alg="QRIteration"):
    """geigvals!(Hermitian(A, uplo), QRIteration(); sortby) (GenericSchur.jl:
    137-146)."""
    A = np.asarray(A)
    _checksquare(A)
    W, _, _ = heev_batched(A[None], uplo=uplo, wantV=False, sortby=sortby, device=device, alg=alg)
    return W[0]


def steqr_(d, e, Z=None, device=0):
    """_gschur!(SymTridiagonal(d, e), QRIteration(), Z) (symtridiag.jl:77-172):
    returns (eigenvalues unsorted, Z times the rotations or None, e after the
    iteration, info).  Z (zrows x n) may be real or complex."""
    d = np.array(d, dtype=np.float64, copy=True).reshape(-1)
    n = d.shape[0]
    e1 = np.zeros(max(n, 1))
    ev = np.asarray(e, dtype=np.float64).reshape(-1)
    if ev.shape[0] != max(n - 1, 0):
        raise DimensionMismatch("e must have length n - 1")
    e1[:ev.shape[0]] = ev
    z = None
    zc = 0
    if Z is not None:
        Z = np.asarray(Z)
        if Z.ndim != 2 or Z.shape[1] != n:
            raise DimensionMismatch("second dimension of Z must match tridiagonal A")
        zc = int(np.iscomplexobj(Z))
        z = np.array(Z, dtype=np.complex128 if zc else np.float64, order="F", copy=True)
    info = np.zeros(1, dtype=np.int32)
    h = _lib.handle(device)
    zr = z.shape[0] if z is not None else 0
    rc = _lib.lib().gs_steqr(h, n, _ptr(d) if n else None, _ptr(e1), None if z is None else _ptr(z), max(zr, 1), zr,
                             zc, _ptr(info), 0)
    _lib.check(rc, h, "gs_steqr")
    return d, (z if z is not None else None), e1[:max(n - 1, 0)], int(info[0])


def gschur_symtridiag(d, e, Z=None, device=0, alg="QRIteration"):
    """gschur!(H::Hessenberg{<:Any, <:SymTridiagonal}, Z) (symtridiag.jl:14-41)
    on the tridiagonal (d, e) with the transformation Z: Schur(diagm(values),
    Z updated, values) (T and Z empty when Z is None)."""
    _check_alg(alg)
    w, z, _, info = steqr_(d, e, Z, device=device)
    if info:
        raise UnconvergedException(f"{info} eigenvalues failed to converge")
    if z is None:
        return Schur(np.zeros((0, 0)), np.zeros((0, 0)), w)
    T = np.zeros((w.shape[0], w.shape[0]), dtype=z.dtype)
    T[np.diag_indices(w.shape[0])] = w
    return Schur(T, z, w)


def geigen_symtridiag(d, e, device=0, alg="QRIteration"):
    """geigen!(A::SymTridiagonal, QRIteration()) (symtridiag.jl:43-49):
    Eigen(values ascending, vectors)."""
    _check_alg(alg)
    n = np.asarray(d).reshape(-1).shape[0]
    w, V, _, info = steqr_(d, e, np.eye(n), device=device)
    if info:
        raise UnconvergedException(f"{info} eigenvalues failed to converge")
    w2, V2 = _sort_host(w, V, lambda x: x)
    return Eigen(w2, V2)


def geigvals_symtridiag(d, e, device=0, alg="QRIteration"):
    """geigvals!(A::SymTridiagonal, QRIteration()) (symtridiag.jl:51-55):
    the eigenvalues in the order the iteration leaves them."""
    _check_alg(alg)
    w, _, _, info = steqr_(d, e, None, device=device)
    if info:
        raise UnconvergedException(f"{info} eigenvalues failed to converge")
    return w
