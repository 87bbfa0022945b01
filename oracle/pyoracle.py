"""TEST INFRASTRUCTURE ONLY: ctypes bindings of the CPU oracle.

The oracle is a C++ restatement of GenericSchur.jl's Schur path (see
``gschur_oracle.hpp``).  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only to check
or time against it.  Matrices are numpy arrays indexed ``A[i, j]`` like the
reference; they are handed to C in column-major (Fortran) order.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liboracle_gschur.so")
_lib = None

c_long = ctypes.c_long
c_int = ctypes.c_int
c_dp = ctypes.POINTER(ctypes.c_double)
c_lp = ctypes.POINTER(ctypes.c_long)
c_ip = ctypes.POINTER(ctypes.c_int)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            raise RuntimeError("oracle library not built: run `make -C oracle`")
        _lib = ctypes.CDLL(_LIBPATH)
        L = _lib
        L.orc_schur_z.argtypes = [c_long, c_dp, c_long, c_dp, c_long, c_dp, c_long, c_dp, c_int, c_int,
                                  c_long, c_long, c_lp]
        L.orc_schur_d.argtypes = [c_long, c_dp, c_long, c_dp, c_long, c_dp, c_long, c_dp, c_int, c_int,
                                  c_long, c_lp]
        L.orc_schur_hess_z.argtypes = [c_long, c_dp, c_long, c_dp, c_long, c_long, c_long, c_int, c_dp]
        L.orc_schur_hess_d.argtypes = [c_long, c_dp, c_long, c_dp, c_long, c_long, c_dp]
        L.orc_hessenberg_z.argtypes = [c_long, c_dp, c_long, c_dp, c_dp, c_long]
        L.orc_hessenberg_d.argtypes = [c_long, c_dp, c_long, c_dp, c_dp, c_long]
        L.orc_hessenberg_z.restype = None
        L.orc_hessenberg_d.restype = None
        L.orc_reflector_z.argtypes = [c_long, c_dp, c_dp]
        L.orc_reflector_d.argtypes = [c_long, c_dp, c_dp]
        L.orc_reflector_z.restype = None
        L.orc_reflector_d.restype = None
        L.orc_gs2x2.argtypes = [c_dp, c_dp, c_dp]
        L.orc_gs2x2.restype = None
        L.orc_eigvecs_z.argtypes = [c_long, c_dp, c_long, c_dp, c_long, c_dp, c_long, c_int, c_int]
        L.orc_eigvecs_d.argtypes = [c_long, c_dp, c_long, c_dp, c_long, c_dp, c_long, c_int, c_int]
        L.orc_eigvecs_z.restype = None
        L.orc_eigvecs_d.restype = None
        L.orc_enormalize.argtypes = [c_long, c_dp, c_long]
        L.orc_enormalize.restype = None
        L.orc_triangularize.argtypes = [c_long, c_dp, c_long, c_dp, c_long, c_dp, c_dp]
        L.orc_triangularize.restype = None
        L.orc_schur_batched_z.argtypes = [c_long, c_ip, c_lp, c_dp, c_dp, c_dp, c_dp, c_ip, c_int, c_int]
        L.orc_schur_batched_d.argtypes = [c_long, c_ip, c_lp, c_dp, c_dp, c_dp, c_dp, c_ip, c_int, c_int]
        L.orc_balance_z.argtypes = [c_long, c_dp, c_long, c_int, c_int, c_lp, c_lp, c_dp]
        L.orc_balance_d.argtypes = [c_long, c_dp, c_long, c_int, c_int, c_lp, c_lp, c_dp]
        L.orc_balance_apply.argtypes = [c_long, c_long, c_lp, c_lp, c_dp, c_dp, c_long, c_int]
        L.orc_ordschur_z.argtypes = [c_long, c_dp, c_long, c_dp, c_long, c_int, c_ip, c_dp]
        L.orc_eigvalscond_z.argtypes = [c_long, c_dp, c_long, c_long]
        L.orc_eigvalscond_z.restype = ctypes.c_double
        L.orc_subspacesep_z.argtypes = [c_long, c_dp, c_long, c_long]
        L.orc_subspacesep_z.restype = ctypes.c_double
        L.orc_trsylvester_z.argtypes = [c_long, c_dp, c_long, c_long, c_dp, c_long, c_dp, c_long, c_int, c_ip]
        L.orc_trsylvester_z.restype = ctypes.c_double
        L.orc_givens_z.argtypes = [c_dp, c_dp, c_dp]
        L.orc_ordschur_z.restype = None
        L.orc_givens_z.restype = None
        L.orc_balance_apply.restype = None
        L.orc_eig_hp_z.argtypes = [c_long, c_dp, c_dp, c_int]
        L.orc_eig_hp_d.argtypes = [c_long, c_dp, c_dp, c_int]
        L.orc_hetrd_z.argtypes = [c_long, c_dp, c_long, c_int, c_dp, c_dp, c_dp, c_dp, c_long]
        L.orc_hetrd_d.argtypes = [c_long, c_dp, c_long, c_int, c_dp, c_dp, c_dp, c_dp, c_long]
        L.orc_hetrd_z.restype = L.orc_hetrd_d.restype = None
        L.orc_steqr.argtypes = [c_long, c_dp, c_dp, c_dp, c_long, c_long, c_int, c_long, c_lp]
        L.orc_steqr.restype = c_long
        L.orc_heev_z.argtypes = [c_long, c_dp, c_long, c_int, c_dp, c_dp, c_long, c_int, c_lp]
        L.orc_heev_d.argtypes = [c_long, c_dp, c_long, c_int, c_dp, c_dp, c_long, c_int, c_lp]
        L.orc_heev_z.restype = L.orc_heev_d.restype = c_long
        L.orc_heev_batched.argtypes = [c_long, c_long, c_int, c_dp, c_dp, c_dp, c_int, c_int]
        L.orc_heev_batched.restype = None
        L.orc_givens_d.argtypes = [ctypes.c_double, ctypes.c_double, c_dp]
        L.orc_givens_d.restype = None
        L.orc_sym2x2.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, c_dp]
        L.orc_sym2x2.restype = None
        L.orc_ordschur_d.argtypes = [c_long, c_dp, c_long, c_dp, c_long, c_ip, c_dp]
        L.orc_ordschur_d.restype = c_long
        L.orc_syl1or2.argtypes = [c_int, c_int, c_dp, c_dp, c_dp, c_dp]
        L.orc_syl1or2.restype = ctypes.c_double
        L.orc_jhypot.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_jhypot.restype = ctypes.c_double
        for f in ("orc_qd_sqrt",):
            getattr(L, f).argtypes = [c_dp, c_dp]
            getattr(L, f).restype = None
        for f in ("orc_qd_div", "orc_qd_mul", "orc_qd_add"):
            getattr(L, f).argtypes = [c_dp, c_dp, c_dp]
            getattr(L, f).restype = None
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags.f_contiguous or a.ndim == 1
    return a.ctypes.data_as(c_dp)


def _fort(a, dtype):
    return np.array(a, dtype=dtype, order="F", copy=True)


UNCONVERGED = 1
DIM = -1
NONREAL_SUB = -2


def gschur(A, wantz=True, scale=True, maxiter=None, maxinner=None):
    """gschur!(A) (GenericSchur.jl:350-372 complex, :805-835 real).

    Returns (T, Z, values, info, stats) with stats = (iterations, sweeps,
    bulge-step bound).
    """
    A = np.asarray(A)
    n = A.shape[0]
    if A.ndim != 2 or A.shape[1] != n:
        raise ValueError("DimensionMismatch: matrix is not square")
    cplx = np.iscomplexobj(A)
    dt = np.complex128 if cplx else np.float64
    a = _fort(A, dt)
    T = np.zeros((n, n), dtype=dt, order="F")
    Z = np.zeros((n, n), dtype=dt, order="F")
    w = np.zeros(n, dtype=np.complex128)
    stats = np.zeros(3, dtype=np.int64)
    mi = 100 * n if maxiter is None else maxiter
    if cplx:
        mn = 30 * n if maxinner is None else maxinner
        info = lib().orc_schur_z(n, _p(a), n, _p(T), n, _p(Z), n, w.ctypes.data_as(c_dp), int(wantz),
                                 int(scale), mi, mn, stats.ctypes.data_as(c_lp))
    else:
        info = lib().orc_schur_d(n, _p(a), n, _p(T), n, _p(Z), n, w.ctypes.data_as(c_dp), int(wantz),
                                 int(scale), mi, stats.ctypes.data_as(c_lp))
    return T, (Z if wantz else np.zeros((0, 0), dtype=dt)), w, info, tuple(int(s) for s in stats)


def schur_hess(H, Z=None, maxiter=None, maxinner=None, checksd=True):
    """gschur!(H::Hessenberg, Z) (GenericSchur.jl:194-335, :513-699)."""
    H = np.asarray(H)
    n = H.shape[0]
    cplx = np.iscomplexobj(H)
    dt = np.complex128 if cplx else np.float64
    h = _fort(H, dt)
    z = None if Z is None else _fort(Z, dt)
    w = np.zeros(n, dtype=np.complex128)
    mi = 100 * n if maxiter is None else maxiter
    if cplx:
        mn = 30 * n if maxinner is None else maxinner
        info = lib().orc_schur_hess_z(n, _p(h), n, _p(z), n, mi, mn, int(checksd), w.ctypes.data_as(c_dp))
    else:
        info = lib().orc_schur_hess_d(n, _p(h), n, _p(z), n, mi, w.ctypes.data_as(c_dp))
    return h, z, w, info


def hessenberg(A):
    """_hessenberg! + _materializeQ (hessenberg.jl:3-17,150-166).

    Returns (factors, tau, Q, H) with H = triu(factors, -1).
    """
    A = np.asarray(A)
    n = A.shape[0]
    cplx = np.iscomplexobj(A)
    dt = np.complex128 if cplx else np.float64
    a = _fort(A, dt)
    tau = np.zeros(max(n - 1, 1), dtype=dt)
    Q = np.zeros((n, n), dtype=dt, order="F")
    f = lib().orc_hessenberg_z if cplx else lib().orc_hessenberg_d
    f(n, _p(a), n, tau.ctypes.data_as(c_dp), _p(Q), n)
    return a, tau[: max(n - 1, 0)], Q, np.triu(a, -1)


def reflector(x):
    """_reflector! (householder.jl:12-102). Returns (x_out, tau)."""
    x = np.array(x)
    cplx = np.iscomplexobj(x)
    dt = np.complex128 if cplx else np.float64
    xx = np.array(x, dtype=dt)
    tau = np.zeros(2)
    if cplx:
        lib().orc_reflector_z(len(xx), xx.ctypes.data_as(c_dp), tau.ctypes.data_as(c_dp))
        return xx, complex(tau[0], tau[1])
    lib().orc_reflector_d(len(xx), xx.ctypes.data_as(c_dp), tau.ctypes.data_as(c_dp))
    return xx, float(tau[0])


def gs2x2(a, b, c, d):
    """_gs2x2! (GenericSchur.jl:716-803).  Returns ((a,b,c,d), (cs,sn), w1, w2)."""
    abcd = np.array([a, b, c, d], dtype=np.float64)
    cssn = np.zeros(2)
    w = np.zeros(4)
    lib().orc_gs2x2(abcd.ctypes.data_as(c_dp), cssn.ctypes.data_as(c_dp), w.ctypes.data_as(c_dp))
    return tuple(abcd), tuple(cssn), complex(w[0], w[1]), complex(w[2], w[3])


def eigvecs(T, Z=None, left=False, normalize=True):
    """geigvecs(S; left) (vectors.jl:12-20); normalize=False skips _enormalize!."""
    T = np.asarray(T)
    n = T.shape[0]
    cplx = np.iscomplexobj(T)
    dt = np.complex128 if cplx else np.float64
    t = _fort(T, dt)
    z = None if Z is None else _fort(Z, dt)
    V = np.zeros((n, n), dtype=np.complex128, order="F")
    f = lib().orc_eigvecs_z if cplx else lib().orc_eigvecs_d
    f(n, _p(t), n, _p(z), n, _p(V), n, int(left), int(normalize))
    return V


def enormalize(V):
    """_enormalize!(V) (util.jl:572-592); V is square, as in the reference."""
    assert V.ndim == 2 and V.shape[0] == V.shape[1], "enormalize takes a square matrix"
    v = _fort(V, np.complex128)
    lib().orc_enormalize(v.shape[0], _p(v), v.shape[0])
    return v


def triangularize(T, Z):
    """triangularize (triang.jl:9-43)."""
    n = T.shape[0]
    t = _fort(T, np.float64)
    z = _fort(Z, np.float64)
    tc = np.zeros((n, n), dtype=np.complex128, order="F")
    zc = np.zeros((n, n), dtype=np.complex128, order="F")
    lib().orc_triangularize(n, _p(t), n, _p(z), n, _p(tc), _p(zc))
    return tc, zc


def schur_batched(ns, offsets, Aflat, complex_input, nthreads=0, wantz=True):
    """Batched gschur! over packed column-major matrices (OpenMP over matrices)."""
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    count = len(ns)
    dt = np.complex128 if complex_input else np.float64
    A = np.ascontiguousarray(Aflat, dtype=dt)
    T = np.zeros_like(A)
    Z = np.zeros_like(A)
    W = np.zeros(int(ns.sum()), dtype=np.complex128)
    info = np.zeros(count, dtype=np.int32)
    f = lib().orc_schur_batched_z if complex_input else lib().orc_schur_batched_d
    f(count, ns.ctypes.data_as(c_ip), offsets.ctypes.data_as(c_lp), A.ctypes.data_as(c_dp),
      T.ctypes.data_as(c_dp), Z.ctypes.data_as(c_dp), W.ctypes.data_as(c_dp), info.ctypes.data_as(c_ip),
      int(nthreads), int(wantz))
    return T, Z, W, info


def eig_hp(A, prec=2):
    """Eigenvalues of the double matrix A by the restated algorithm run in
    quad-double (prec=2, ~212 bits) or binary128 (prec=1).

    Returns (values as complex128 rounded, values as (re_hi, re_lo, im_hi,
    im_lo) array, info).
    """
    A = np.asarray(A)
    n = A.shape[0]
    cplx = np.iscomplexobj(A)
    a = _fort(A, np.complex128 if cplx else np.float64)
    w = np.zeros((n, 4))
    f = lib().orc_eig_hp_z if cplx else lib().orc_eig_hp_d
    info = f(n, _p(a), w.ctypes.data_as(c_dp), int(prec))
    vals = (w[:, 0] + w[:, 1]) + 1j * (w[:, 2] + w[:, 3])
    return vals, w, info


def qd_op(op, a, b=None):
    a = np.asarray(a, dtype=np.float64)
    r = np.zeros(4)
    if op == "sqrt":
        lib().orc_qd_sqrt(a.ctypes.data_as(c_dp), r.ctypes.data_as(c_dp))
    else:
        b = np.asarray(b, dtype=np.float64)
        getattr(lib(), "orc_qd_" + op)(a.ctypes.data_as(c_dp), b.ctypes.data_as(c_dp), r.ctypes.data_as(c_dp))
    return r


def balance(A, scale=True, permute=True):
    """balance!(A; scale, permute) (balance.jl:33-199).

    Returns (Abal, (ilo, ihi, trivial), sp, d, info); ilo/ihi are 1-based,
    sp is the permutation record (1-based), d the scale factors."""
    A = np.asarray(A)
    n = A.shape[0]
    cplx = np.iscomplexobj(A)
    a = _fort(A, np.complex128 if cplx else np.float64)
    iit = np.zeros(3, dtype=np.int64)
    sp = np.zeros(max(n, 1), dtype=np.int64)
    d = np.ones(max(n, 1))
    f = lib().orc_balance_z if cplx else lib().orc_balance_d
    info = f(n, _p(a), max(n, 1), int(scale), int(permute), iit.ctypes.data_as(c_lp), sp.ctypes.data_as(c_lp),
             d.ctypes.data_as(c_dp))
    return a, (int(iit[0]), int(iit[1]), bool(iit[2])), sp[:n], d[:n], info


def balance_apply(V, bal, sp, d, inverse=False):
    """lmul!(B, V) (inverse=False) / ldiv!(B, V) (inverse=True) (balance.jl:201-261)."""
    v = _fort(V, np.complex128)
    n, m = v.shape
    iit = np.array([bal[0], bal[1], int(bal[2])], dtype=np.int64)
    spa = np.ascontiguousarray(sp, dtype=np.int64)
    da = np.ascontiguousarray(d, dtype=np.float64)
    lib().orc_balance_apply(n, m, iit.ctypes.data_as(c_lp), spa.ctypes.data_as(c_lp), da.ctypes.data_as(c_dp),
                            _p(v), max(n, 1), int(inverse))
    return v


def eigen(A, permute=True, scale=True, sort=True, jvl=False, jce=False):
    """eigen!(A; permute, scale, sortby=eigsortby, jvl, jce) (pirates.jl:63-160),
    composed from the restated pieces: balance!, gschur!, triangularize
    (real input), the complex right / left eigenvector solves without
    normalisation, lmul!/ldiv!(B, .), _enormalize!, sorteig!.

    Returns (values, vectors, vectors_left or None, rconde or None)."""
    A = np.asarray(A)
    n = A.shape[0]
    if permute or scale:
        Ab, bal, sp, d, info = balance(A, scale=scale, permute=permute)
        assert info == 0
    else:
        Ab = np.array(A)
    T, Z, w, info, _ = gschur(Ab)
    assert info == 0
    if not np.iscomplexobj(A):
        T, Z = triangularize(T, Z)
        w = np.diag(T).copy()
    v = eigvecs(T, Z, left=False, normalize=False)
    vl = eigvecs(T, Z, left=True, normalize=False) if (jvl or jce) else None
    rc = None
    if jce:
        rc = np.array([abs(np.vdot(vl[:, j], v[:, j])) / np.linalg.norm(vl[:, j]) / np.linalg.norm(v[:, j])
                       for j in range(n)])
    if permute or scale:
        v = balance_apply(v, bal, sp, d, inverse=False)
        if vl is not None:
            vl = balance_apply(vl, bal, sp, d, inverse=True)
    v = enormalize(v)
    if vl is not None:
        vl = enormalize(vl)
    if sort:
        p = np.lexsort((w.imag, w.real))
        w, v = w[p], v[:, p]
        if vl is not None:
            vl = vl[:, p]
        if rc is not None:
            rc = rc[p]
    return w, v, (vl if jvl else None), rc


def ordschur(T, Z, select):
    """gordschur!(T, Z, select) for a complex Schur form (ordschur.jl:23-43):
    returns (T2, Z2, w)."""
    n = T.shape[0]
    t = _fort(T, np.complex128)
    z = _fort(Z, np.complex128) if Z is not None else np.zeros((1, 1), dtype=np.complex128, order="F")
    sel = np.ascontiguousarray(np.asarray(select, dtype=bool).astype(np.int32))
    w = np.zeros(max(n, 1), dtype=np.complex128)
    lib().orc_ordschur_z(n, _p(t), max(n, 1), _p(z), max(z.shape[0], 1), int(Z is not None),
                         sel.ctypes.data_as(c_ip), _p(w))
    return t, (z if Z is not None else None), w[:n]


def eigvalscond(T, nsub):
    """eigvalscond(S, nsub) (ordschur.jl:85-98) of the complex triangular T."""
    t = _fort(T, np.complex128)
    return float(lib().orc_eigvalscond_z(t.shape[0], _p(t), t.shape[0], int(nsub)))


def subspacesep(T, nsub):
    """subspacesep(S, nsub) (ordschur.jl:111-131) of the complex triangular T."""
    t = _fort(T, np.complex128)
    return float(lib().orc_subspacesep_z(t.shape[0], _p(t), t.shape[0], int(nsub)))


def trsylvester(A, B, C, adjoint=False):
    """trsylvester!(A, B, C) / adjtrsylvester!(A, B, C) (sylvester.jl:24-139),
    complex: returns (X, scale, perturbed)."""
    a = _fort(A, np.complex128)
    b = _fort(B, np.complex128)
    c = _fort(C, np.complex128)
    m, n = c.shape
    pert = np.zeros(1, dtype=np.int32)
    sc = lib().orc_trsylvester_z(m, _p(a), max(m, 1), n, _p(b), max(n, 1), _p(c), max(m, 1), int(adjoint),
                                 pert.ctypes.data_as(c_ip))
    return c, float(sc), bool(pert[0])


def givens(f, g):
    """LinearAlgebra.givensAlgorithm(f, g), complex: (c, s, r)."""
    fa = np.array([complex(f).real, complex(f).imag])
    ga = np.array([complex(g).real, complex(g).imag])
    out = np.zeros(5)
    lib().orc_givens_z(_p(fa), _p(ga), _p(out))
    return out[0], complex(out[1], out[2]), complex(out[3], out[4])


# ---------------------------------------------------------------- Hermitian


def hetrd(A, uplo="L", wantq=True):
    """_hehessenberg!(Hermitian(A, uplo)) (hessenberg.jl:27-95) and
    _materializeQ (hessenberg.jl:133-232): returns (d, e, Q or None); only
    the uplo triangle of A is read."""
    A = np.asarray(A)
    cplx = np.iscomplexobj(A)
    dt = np.complex128 if cplx else np.float64
    n = A.shape[0]
    a = _fort(A, dt)
    d = np.zeros(max(n, 1))
    e = np.zeros(max(n - 1, 1))
    tau = np.zeros(max(n - 1, 1), dtype=dt)
    q = np.zeros((max(n, 1), max(n, 1)), dtype=dt, order="F") if wantq else None
    f = lib().orc_hetrd_z if cplx else lib().orc_hetrd_d
    f(n, _p(a), max(n, 1), int(uplo == "U"), _p(d), _p(e), _p(tau), _p(q), max(n, 1))
    return d[:n], e[:max(n - 1, 0)], (q[:n, :n] if wantq else None)


def steqr(d, e, Z=None, maxsweeps=-1):
    """_gschur!(SymTridiagonal(d, e), QRIteration(), Z) (symtridiag.jl:77-172):
    returns (eigenvalues unsorted, Z updated or None, info, sweeps)."""
    d = np.array(d, dtype=np.float64, copy=True)
    n = d.shape[0]
    e = np.array(e, dtype=np.float64, copy=True).reshape(-1)
    e2 = np.zeros(max(n, 1))
    e2[:e.shape[0]] = e
    z = None
    zc = 0
    if Z is not None:
        zc = int(np.iscomplexobj(Z))
        z = _fort(Z, np.complex128 if zc else np.float64)
    sw = c_long(0)
    info = lib().orc_steqr(n, _p(d) if n else None, _p(e2), _p(z), max(z.shape[0], 1) if z is not None else 1,
                           z.shape[0] if z is not None else 0, zc, int(maxsweeps), ctypes.byref(sw))
    return d, z, int(info), int(sw.value)


def heev(A, uplo="L", wantv=True, sort=True):
    """geigen! / geigvals! with QRIteration (GenericSchur.jl:124-146) of a
    Hermitian or real symmetric matrix: returns (w, V or None, info)."""
    A = np.asarray(A)
    cplx = np.iscomplexobj(A)
    dt = np.complex128 if cplx else np.float64
    n = A.shape[0]
    a = _fort(A, dt)
    w = np.zeros(max(n, 1))
    v = np.zeros((max(n, 1), max(n, 1)), dtype=dt, order="F") if wantv else None
    f = lib().orc_heev_z if cplx else lib().orc_heev_d
    sw = c_long(0)
    info = f(n, _p(a), max(n, 1), int(uplo == "U"), _p(w), _p(v), max(n, 1), int(sort), ctypes.byref(sw))
    return w[:n], (v[:n, :n] if wantv else None), int(info)


def heev_batched(A, sort=True, wantv=True, nthreads=0):
    """heev over a (count, n, n) stack (lower triangles), OpenMP over
    matrices: returns (W (count, n), V (count, n, n) or None)."""
    A = np.asarray(A)
    cplx = np.iscomplexobj(A)
    dt = np.complex128 if cplx else np.float64
    count, n = A.shape[0], A.shape[1]
    a = np.ascontiguousarray(np.transpose(A, (0, 2, 1)).astype(dt))
    w = np.zeros((count, n))
    v = np.zeros((count, n, n), dtype=dt) if wantv else None
    lib().orc_heev_batched(count, n, int(cplx), a.ctypes.data_as(c_dp), w.ctypes.data_as(c_dp),
                           v.ctypes.data_as(c_dp) if wantv else None, int(sort), int(nthreads))
    return w, (np.transpose(v, (0, 2, 1)) if wantv else None)


def givens_re(f, g):
    """LinearAlgebra.givensAlgorithm(f, g) for real f, g: (c, s, r)."""
    out = np.zeros(3)
    lib().orc_givens_d(float(f), float(g), _p(out))
    return float(out[0]), float(out[1]), float(out[2])


def sym2x2(a, b, c):
    """_dgz_sym2x2(a, b, c, true) (symtridiag.jl:375-427): (rt1, rt2, cs, sn)."""
    out = np.zeros(4)
    lib().orc_sym2x2(float(a), float(b), float(c), _p(out))
    return tuple(float(x) for x in out)


def jhypot(x, y):
    """Julia's hypot(x, y) for Float64."""
    return float(lib().orc_jhypot(float(x), float(y)))


def ordschur_re(T, Z, select):
    """gordschur!(T, Z, select) for a real Schur form (rordschur.jl:3-45):
    returns (T2, Z2, w, info) with info = k of IllConditionException(k) or 0."""
    n = T.shape[0]
    t = _fort(T, np.float64)
    z = _fort(Z, np.float64) if Z is not None else None
    sel = np.ascontiguousarray(np.asarray(select, dtype=bool).astype(np.int32))
    w = np.zeros(max(n, 1), dtype=np.complex128)
    info = lib().orc_ordschur_d(n, _p(t), max(n, 1), _p(z), max(n, 1), sel.ctypes.data_as(c_ip), _p(w))
    return t, z, w[:n], int(info)


def syl1or2(Tl, Tr, B):
    """_syl1or2(Tl, Tr, B) (rordschur.jl:300-466): (X, scale) with
    Tl X - X Tr = scale B."""
    n1, n2 = Tl.shape[0], Tr.shape[0]
    tl = np.zeros((2, 2), order="F"); tl[:n1, :n1] = Tl
    tr = np.zeros((2, 2), order="F"); tr[:n2, :n2] = Tr
    b = np.zeros((2, 2), order="F"); b[:n1, :n2] = B
    x = np.zeros((2, 2), order="F")
    sc = lib().orc_syl1or2(n1, n2, _p(tl), _p(tr), _p(b), _p(x))
    return x[:n1, :n2].copy(), float(sc)
