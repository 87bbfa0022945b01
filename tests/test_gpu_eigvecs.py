"""GPU parity of the eigenvector kernels (geigvecs / eigvecs) against the oracle.

The oracle's own Schur factors (T, Z) are handed to both the HIP kernel and
the oracle's restatement of vectors.jl, so the comparison isolates the
back-substitution.  Both outputs are normalised to unit 2-norm with the
largest component real and positive (_enormalize!), which fixes the phase:
the vectors are compared entry by entry, and the reference's vectest
residuals (test/testfuncs.jl:5-54) are checked at tolerance 20.
"""
import numpy as np
import pytest

import checks
from genericschur import gen

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


class S:
    def __init__(self, T, Z, values):
        self.T, self.Z, self.values = T, Z, values


def _compare(oracle, gs, A, T, Z, w, tol=1e-9, use_z=True):
    n = A.shape[0]
    print(f"case n={n} dtype={A.dtype} z={use_z}", flush=True)
    Zs = Z if use_z else np.zeros((0, 0), dtype=Z.dtype)
    for left in (False, True):
        Vg = gs.geigvecs(S(T, Zs, w), left=left)
        Vo = oracle.eigvecs(T, Z if use_z else None, left=left)
        err = np.abs(Vg - Vo).max()
        assert err <= tol, (left, err)
    if use_z:
        VR = gs.geigvecs(S(T, Z, w), left=False)
        VL = gs.geigvecs(S(T, Z, w), left=True)
        checks.vectest(A, w, VR, VL, 20)


@pytest.mark.parametrize("n", [1, 2, 5, 32, 64, 96, 128, 129, 256, 400, 700])
def test_complex_vectors(gs, oracle, n):
    """Above n = 128 the kernel reads T from global memory (any order up to
    1024; geigvecs, vectors.jl:12-20, takes any order)."""
    A = gen.ginibre_complex(n, gen.stream(21, 200, n))
    T, Z, w, info, _ = oracle.gschur(A)
    assert info == 0
    _compare(oracle, gs, A, T, Z, w, tol=1e-9 if n <= 128 else 1e-8)
    if n <= 32:
        _compare(oracle, gs, A, T, Z, w, use_z=False)


@pytest.mark.parametrize("n", [1, 2, 3, 5, 32, 100, 141, 193, 256, 512, 800])
def test_real_vectors(gs, oracle, n):
    """Above n = 192 the kernel reads T from global memory."""
    A = gen.ginibre_real(n, gen.stream(22, 200, n))
    T, Z, w, info, _ = oracle.gschur(A)
    assert info == 0
    _compare(oracle, gs, A, T, Z, w, tol=1e-9 if n <= 192 else 1e-8)
    if n <= 32:
        _compare(oracle, gs, A, T, Z, w, use_z=False)


def test_real_vectors_mixed_kinds(gs, oracle):
    mats, kinds = gen.workload3(seed=9, count=12)
    for A in mats:
        T, Z, w, info, _ = oracle.gschur(A)
        assert info == 0
        _compare(oracle, gs, A, T, Z, w, tol=1e-7)


@pytest.mark.parametrize("dmode", [0, 1, 2])
def test_complex_triangular_repeated(gs, oracle, dmode):
    """Repeated / near-repeated eigenvalues (test/complex.jl:327-364) run the
    perturbed, carefully scaled solves."""
    n = 32
    r = np.random.default_rng(41 + dmode)
    d = r.random(n) + 1j * r.random(n)
    if dmode == 0:
        d[1] = d[n - 2]
    elif dmode == 1:
        d[1] = d[n - 2] * np.nextafter(1.0, 2.0)
    else:
        d[1] = d[n - 2] * np.nextafter(1.0, 0.0)
    A = np.triu(r.random((n, n)) + 1j * r.random((n, n)), 1) + np.diag(d)
    I = np.eye(n, dtype=complex)
    for left in (False, True):
        Vg = gs.geigvecs(S(A, I, np.diag(A)), left=left)
        Vo = oracle.eigvecs(A, I, left=left)
        checks.vectest(A, np.diag(A), gs.geigvecs(S(A, I, np.diag(A))), gs.geigvecs(S(A, I, np.diag(A)), left=True),
                       20)
        assert np.abs(np.abs(Vg) - np.abs(Vo)).max() <= 1e-6


def test_gpu_schur_then_vectors(gs, oracle):
    """The full drop-in chain: gschur on the GPU, then eigvecs on the GPU."""
    for cplx in (True, False):
        A = gen.ginibre_complex(40, gen.stream(5, 5, 5)) if cplx else gen.ginibre_real(40, gen.stream(5, 5, 5))
        Sg = gs.gschur(A)
        VR = gs.eigvecs(Sg)
        VL = gs.eigvecs(Sg, left=True)
        checks.vectest(A, Sg.values, VR, VL, 20)
        assert np.all(np.abs(np.linalg.norm(VR, axis=0) - 1) < 40 * EPS)


def test_vectors_nonfinite_no_crash(gs, oracle):
    """test/complex.jl:393-406: non-finite T must not throw."""
    A = gen.ginibre_complex(5, gen.stream(1, 1, 1))
    Sg = gs.gschur(A)
    for bad in (np.inf, np.nan):
        T2 = Sg.T.copy()
        T2[1, 2] = bad
        gs.geigvecs(S(T2, Sg.Z, Sg.values), left=False)
        gs.geigvecs(S(T2, Sg.Z, Sg.values), left=True)


def test_batched_vectors(gs, oracle):
    import ctypes
    from genericschur import _lib
    for cplx in (True, False):
        mats = gen.workload2(seed=8, count=10) if cplx else gen.workload3(seed=8, count=10)[0]
        dt = np.complex128 if cplx else np.float64
        Ts, Zs = [], []
        for A in mats:
            T, Z, w, info, _ = oracle.gschur(A)
            Ts.append(T)
            Zs.append(Z)
        ns, off, tflat = gs.pack(Ts, dt)
        _, _, zflat = gs.pack(Zs, dt)
        V = np.zeros(len(tflat), dtype=np.complex128)
        h = _lib.handle()
        f = _lib.lib().gs_trevc_batched_z if cplx else _lib.lib().gs_trevc_batched_d
        p = lambda x: ctypes.c_void_p(x.ctypes.data)
        for left in (0, 1):
            rc = f(h, len(ns), p(ns), p(off), p(tflat), p(zflat), p(V), left, 1, 0)
            _lib.check(rc, h, "gs_trevc_batched")
            for k, (T, Z) in enumerate(zip(Ts, Zs)):
                m, o = int(ns[k]), int(off[k])
                Vg = V[o:o + m * m].reshape((m, m), order="F")
                Vo = oracle.eigvecs(T, Z, left=bool(left))
                assert np.abs(Vg - Vo).max() <= 1e-7


def _schur_lapack(n, cplx, seed):
    """A random matrix and its Schur factors from LAPACK (scipy): the inputs
    of the orders above 1024, where the oracle's own Schur factorization
    would dominate the test's time.  The comparison isolates the
    back-substitution either way (T and Z go to both sides)."""
    import scipy.linalg as sl
    r = np.random.default_rng(seed)
    A = r.standard_normal((n, n))
    if cplx:
        A = A + 1j * r.standard_normal((n, n))
        T, Z = sl.schur(A, output="complex")
        return A, T, Z, np.diag(T).copy()
    T, Z = sl.schur(A)
    w = np.diag(T).astype(np.complex128)
    j = 0
    while j < n - 1:
        if T[j + 1, j] != 0.0:
            wi = np.sqrt(abs(T[j + 1, j])) * np.sqrt(abs(T[j, j + 1]))
            w[j] = T[j, j] + 1j * wi
            w[j + 1] = T[j, j] - 1j * wi
            j += 2
        else:
            j += 1
    return A, T, Z, w


@pytest.mark.parametrize("n,cplx", [(1100, False), (2100, False), (1100, True)])
def test_vectors_large_orders(gs, oracle, n, cplx):
    """Orders above 1024 run one matrix at a time over the whole GPU (one
    solve per workgroup with the vector in LDS, the back-transform as FP64
    GEMMs); n = 2100 takes the 4096-entry vectors.  Entrywise against the
    oracle and the reference's vectest residuals."""
    A, T, Z, w = _schur_lapack(n, cplx, 300 + n)
    _compare(oracle, gs, A, T, Z, w, tol=1e-7)


def test_vectors_large_complex_4096_entry(gs, oracle):
    """Complex n = 2100: the 4096-entry vectors of the complex solves."""
    A, T, Z, w = _schur_lapack(2100, True, 2400)
    _compare(oracle, gs, A, T, Z, w, tol=1e-7)


@pytest.mark.parametrize("cplx", [False, True])
def test_vectors_large_no_z(gs, oracle, cplx):
    """Eigenvectors of T itself (no Z) above order 1024: the finish takes the
    solutions directly, with the support-limited max scaling."""
    A, T, Z, w = _schur_lapack(1030, cplx, 77)
    _compare(oracle, gs, T, T, Z, w, tol=1e-7, use_z=False)


def test_batched_vectors_with_large_order(gs, oracle):
    """A batch holding one matrix above 1024 among small ones: the large one
    runs through the whole-GPU path, the rest through the batched kernel."""
    import ctypes
    from genericschur import _lib
    for cplx in (False, True):
        dt = np.complex128 if cplx else np.float64
        Ts, Zs = [], []
        for k, m in enumerate((40, 1050, 7, 300)):
            _, T, Z, _ = _schur_lapack(m, cplx, 900 + k)
            Ts.append(T)
            Zs.append(Z)
        ns, off, tflat = gs.pack(Ts, dt)
        _, _, zflat = gs.pack(Zs, dt)
        V = np.zeros(len(tflat), dtype=np.complex128)
        h = _lib.handle()
        f = _lib.lib().gs_trevc_batched_z if cplx else _lib.lib().gs_trevc_batched_d
        p = lambda x: ctypes.c_void_p(x.ctypes.data)
        for left in (0, 1):
            print(f"batched large cplx={cplx} left={left}", flush=True)
            rc = f(h, len(ns), p(ns), p(off), p(tflat), p(zflat), p(V), left, 1, 0)
            _lib.check(rc, h, "gs_trevc_batched")
            for k, (T, Z) in enumerate(zip(Ts, Zs)):
                m, o = int(ns[k]), int(off[k])
                Vg = V[o:o + m * m].reshape((m, m), order="F")
                Vo = oracle.eigvecs(T, Z, left=bool(left))
                assert np.abs(Vg - Vo).max() <= 1e-7, (cplx, left, m)


@pytest.mark.parametrize("n,cplx", [(50, False), (50, True), (1100, False), (1100, True)])
def test_vectors_unnormalized(gs, oracle, n, cplx):
    """normalize = 0: the solutions back-transformed and scaled to largest
    entry 1 (max abs, or abs1 for the real pairs and the complex kernel),
    before _enormalize! — the form eigen! feeds to the balancing
    back-transform (pirates.jl:86-110).  Small orders on the one-workgroup
    kernels, 1100 on the whole-GPU path; entrywise against the oracle."""
    A, T, Z, w = _schur_lapack(n, cplx, 500 + n)
    for left in (False, True):
        Vg = gs.geigvecs(S(T, Z, w), left=left, normalize=False)
        Vo = oracle.eigvecs(T, Z, left=left, normalize=False)
        assert np.abs(Vg - Vo).max() <= (1e-9 if n <= 64 else 1e-7), left
