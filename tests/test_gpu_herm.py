"""GPU parity of the Hermitian / real symmetric path (heev.hip, steqr.h)
against the CPU oracle.

  * gs_steqr (symtridiag.jl:77-427) is compared bit for bit: d, e and Z,
    real and complex Z, on random tridiagonals and the reference's own test
    classes (test/symtridiag.jl), including the floatmin/ulp- and
    floatmax*ulp-scaled blocks, with runtest's bounds (< 50) on top.
  * gs_hetrd (hessenberg.jl:19-232): d, e, Q against the oracle entry by
    entry for n <= 9, the spectrum of (d, e) against the oracle's for all
    n, both triangles, and hesstest's bounds (test/complex.jl:63-83) at 1
    and 100 floatmin (single tridiagonal entries are not forward-stable
    under the GPU's different summation order; the factorization is).
  * gs_heev (GenericSchur.jl:124-146): eigenvalues against the oracle to
    16 n eps |A|, eigenvectors column by column after aligning the unit
    factor (the iteration's signs follow the algorithm, not a convention),
    A V = V diag(w), V unitary, the sorted order, batches, the largest LDS
    order, and edge cases (n = 1, 2, zero, identity, diagonal).
"""
import numpy as np
import pytest

import symtri_cases as sc

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
TINY = np.finfo(np.float64).tiny


# largest orders of the LDS-resident kernels (one workgroup per matrix);
# above, up to gs.heev_max_n(), one matrix at a time over the whole GPU
LDS_MAX = {True: 96, False: 128}


def _herm(r, n, cplx, scale=1.0):
    A = r.standard_normal((n, n)) + (1j * r.standard_normal((n, n)) if cplx else 0)
    return (A + A.conj().T) * scale


@pytest.mark.parametrize("zkind", ["none", "real", "complex"])
def test_steqr_bitwise_random(gs, oracle, zkind):
    r = np.random.default_rng(60)
    for n in (1, 2, 3, 8, 31, 64, 130, 257):
        d = r.standard_normal(n)
        e = r.standard_normal(max(n - 1, 0))
        if zkind == "none":
            Z = None
        elif zkind == "real":
            Z = r.standard_normal((n + 3, n))
        else:
            Z = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
        print(f"steqr n={n} z={zkind}", flush=True)
        wo, zo, info_o, _ = oracle.steqr(d, e, Z)
        wg, zg, _, info_g = gs.steqr_(d, e, Z)
        assert info_g == info_o == 0
        assert np.array_equal(wg, wo), n
        if Z is not None:
            assert np.array_equal(zg, zo), n


def test_steqr_reference_classes_bitwise(gs, oracle):
    """The reference's classes 1 (types 1-9) and 2 (types 0-8, type 8 the
    scaled glued Wilkinson matrices) at n = 32 through geigen!(SymTridiagonal)
    (V = I): bitwise equal to the oracle and within runtest's bounds."""
    r = np.random.default_rng(61)
    tri = lambda A: oracle.hetrd(A, "L", wantq=False)[:2]
    mats = [("c1", t) + sc.class1(32, t, r, tri)[:2] for t in range(1, 10)]
    mats += [("c2", t) + sc.class2(32, t)[:2] for t in range(0, 9)]
    for cls, t, d, e in mats:
        n = d.shape[0]
        print(f"steqr class {cls} type {t} n={n}", flush=True)
        wo, zo, info_o, _ = oracle.steqr(d, e, np.eye(n))
        wg, zg, _, info_g = gs.steqr_(d, e, np.eye(n))
        assert info_g == info_o == 0
        assert np.array_equal(wg, wo) and np.array_equal(zg, zo), (cls, t)
        de, ve = sc.errors(d, e, wg, zg)
        assert de < 50 and ve < 50, (cls, t, de, ve)
        E = gs.geigen_symtridiag(d, e)
        assert np.all(np.diff(E.values) >= 0)
        assert np.array_equal(np.sort(wo), E.values)


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_hetrd_parity(gs, oracle, cplx, uplo):
    r = np.random.default_rng(62)
    nmax = LDS_MAX[cplx]
    for n in (1, 2, 3, 9, 40, 64, nmax):
        A = _herm(r, n, cplx)
        print(f"hetrd n={n} cplx={cplx} uplo={uplo}", flush=True)
        d, e, Q = gs.hessenberg_hermitian(A, uplo)
        do, eo, Qo = oracle.hetrd(A, uplo)
        scale = np.abs(A).max()
        if n <= 9:
            # entrywise only for small orders: single entries of the
            # tridiagonal form are not forward-stable under a different
            # rounding (the GPU sums in parallel), the factorization is
            tol = 32 * n * EPS
            assert np.abs(d - do).max() <= tol * scale
            if n > 1:
                assert np.abs(e - eo).max() <= tol * scale
            assert np.abs(Q - Qo).max() <= tol
        # the same spectrum as the oracle's tridiagonal form
        lam = np.linalg.eigvalsh(np.diag(d) + np.diag(e, 1) + np.diag(e, -1))
        lamo = np.linalg.eigvalsh(np.diag(do) + np.diag(eo, 1) + np.diag(eo, -1))
        assert np.abs(lam - lamo).max() <= 16 * n * EPS * scale
        for s in (1.0, 100 * TINY):
            As = A * s
            d2, e2, Q2 = gs.hessenberg_hermitian(As, uplo)
            T = np.diag(d2) + np.diag(e2, 1) + np.diag(e2, -1)
            nrm = np.linalg.norm(As / s)
            assert np.linalg.norm((As - Q2 @ T @ Q2.conj().T) / s) / (n * nrm * EPS) < 10
            assert np.linalg.norm(np.eye(n) - Q2 @ Q2.conj().T) / (n * EPS) < 10


def _check_eig(A, w, V, wo, Vo):
    n = A.shape[0]
    scale = max(np.abs(A).max(), TINY)
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - wo).max() <= 16 * n * EPS * scale
    assert np.abs(A @ V - V * w).max() <= 16 * n * EPS * scale
    assert np.abs(V.conj().T @ V - np.eye(n)).max() <= 16 * n * EPS
    # columns of well-separated eigenvalues agree with the oracle's after
    # aligning the unit factor
    gap = np.full(n, np.inf)
    if n > 1:
        dw = np.diff(wo)
        gap[:-1] = np.minimum(gap[:-1], dw)
        gap[1:] = np.minimum(gap[1:], dw)
    for j in range(n):
        if gap[j] < 1e-3 * scale:
            continue
        ph = np.vdot(Vo[:, j], V[:, j])
        ph = ph / abs(ph)
        assert np.abs(V[:, j] - ph * Vo[:, j]).max() <= 1e3 * n * EPS * scale / gap[j]


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_heev_parity(gs, oracle, cplx, uplo):
    r = np.random.default_rng(63)
    nmax = LDS_MAX[cplx]
    for n in (1, 2, 3, 10, 33, 64, nmax):
        A = _herm(r, n, cplx)
        print(f"heev n={n} cplx={cplx} uplo={uplo}", flush=True)
        E = gs.geigen_hermitian(A, uplo)
        wo, Vo, info = oracle.heev(A, uplo)
        assert info == 0
        _check_eig(A, E.values, E.vectors, wo, Vo)
        w = gs.geigvals_hermitian(A, uplo)
        assert np.array_equal(w, E.values)


@pytest.mark.parametrize("cplx", [False, True])
def test_heev_batched(gs, oracle, cplx):
    """A batch of 300 matrices in one launch against the oracle, unsorted
    (the iteration's own order) and sorted."""
    r = np.random.default_rng(64)
    n = 40
    A = np.stack([_herm(r, n, cplx) for _ in range(300)])
    W, V, info = gs.heev_batched(A)
    assert not info.any()
    Wo, Vo = oracle.heev_batched(A)
    for k in range(0, 300, 7):
        _check_eig(A[k], W[k], V[k], Wo[k], Vo[k])
    Wu, Vu, _ = gs.heev_batched(A, sortby=None)
    for k in range(0, 300, 37):
        wo, vo, _ = oracle.heev(A[k], "L", sort=False)
        assert np.abs(Wu[k] - wo).max() <= 16 * n * EPS * np.abs(A[k]).max()
        p = np.argsort(Wu[k], kind="stable")
        assert np.array_equal(Wu[k][p], W[k])


@pytest.mark.parametrize("cplx", [False, True])
def test_heev_edge_cases(gs, oracle, cplx):
    dt = np.complex128 if cplx else np.float64
    # zero, identity (all ties: no permutation), diagonal with repeats
    for A in (np.zeros((5, 5), dtype=dt), np.eye(6, dtype=dt), np.diag([3.0, -1.0, 2.0, -1.0]).astype(dt)):
        E = gs.geigen_hermitian(A)
        wo, Vo, _ = oracle.heev(A, "L")
        assert np.array_equal(E.values, wo)
        assert np.array_equal(E.vectors, Vo.astype(dt))
    # 1 x 1 with a non-real diagonal: the reference reads the real part only
    A = np.array([[2.5 + (0.7j if cplx else 0)]], dtype=dt)
    E = gs.geigen_hermitian(A)
    assert E.values[0] == 2.5 and E.vectors[0, 0] == 1
    # the unreferenced triangle is ignored
    r = np.random.default_rng(65)
    A = _herm(r, 12, cplx)
    B = np.tril(A) + np.triu(r.standard_normal((12, 12)), 1)
    assert np.array_equal(gs.geigvals_hermitian(B, "L"), gs.geigvals_hermitian(A, "L"))
    C = np.triu(A) + np.tril(r.standard_normal((12, 12)), -1)
    assert np.array_equal(gs.geigvals_hermitian(C, "U"), gs.geigvals_hermitian(A, "U"))
    # above the largest order: a clear error, no fallback
    assert gs.heev_max_n(cplx) == 4096
    with pytest.raises(gs.GSError):
        gs.geigen_hermitian(np.eye(gs.heev_max_n(cplx) + 1, dtype=dt))


@pytest.mark.parametrize("cplx", [False, True])
def test_heev_phases_bitwise(gs, cplx, monkeypatch):
    """The batched solver (tridiagonalization + Q, the lane-parallel QL/QR
    recurrence writing rotation streams, their replay onto Q, the sort)
    against the same steps one call each (gs_hetrd, then gs_steqr on
    (d, e, Q), then a stable sort): bit for bit.  Again with streams too
    short for the iteration, so that phase 3 finishes it in place."""
    r = np.random.default_rng(66)
    n, count = 40, 24
    A = np.stack([_herm(r, n, cplx) for _ in range(count)])
    W, V, info = gs.heev_batched(A)
    assert not info.any()
    for k in range(count):
        d, e, Q = gs.hessenberg_hermitian(A[k])
        w, Z, _, inf1 = gs.steqr_(d, e, Q)
        assert inf1 == 0
        p = np.argsort(w, kind="stable")
        assert np.array_equal(W[k], w[p]), k
        assert np.array_equal(V[k], Z[:, p]), k
    for cap in ("1", "300"):
        monkeypatch.setenv("GSCHUR_HEEV_LOGCAP", cap)
        W2, V2, info2 = gs.heev_batched(A)
        assert not info2.any()
        assert np.array_equal(W2, W) and np.array_equal(V2, V), cap


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_heev_large_orders(gs, oracle, cplx, uplo):
    """Orders above the LDS-resident kernels: the working matrix in HBM,
    each reflector step and each Q accumulation step as grid-wide launches,
    the QL/QR iteration on Q in HBM.  Against the oracle's heev as for the
    small orders (values, residual, unitarity, aligned columns), hetrd's
    factorization bounds, and values-only calls."""
    r = np.random.default_rng(70)
    for n in (LDS_MAX[cplx] + 1, 300):
        A = _herm(r, n, cplx)
        print(f"heev large n={n} cplx={cplx} uplo={uplo}", flush=True)
        E = gs.geigen_hermitian(A, uplo)
        wo, Vo, info = oracle.heev(A, uplo)
        assert info == 0
        _check_eig(A, E.values, E.vectors, wo, Vo)
        w = gs.geigvals_hermitian(A, uplo)
        assert np.abs(w - E.values).max() <= 16 * n * EPS * np.abs(A).max()
        d, e, Q = gs.hessenberg_hermitian(A, uplo)
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        nrm = np.linalg.norm(A)
        assert np.linalg.norm(A - Q @ T @ Q.conj().T) / (n * nrm * EPS) < 10
        assert np.linalg.norm(np.eye(n) - Q @ Q.conj().T) / (n * EPS) < 10
        do, eo, _ = oracle.hetrd(A, uplo)
        lam = np.linalg.eigvalsh(T)
        lamo = np.linalg.eigvalsh(np.diag(do) + np.diag(eo, 1) + np.diag(eo, -1))
        assert np.abs(lam - lamo).max() <= 16 * n * EPS * np.abs(A).max()


@pytest.mark.parametrize("cplx", [False, True])
def test_heev_large_batched_and_structured(gs, oracle, cplx):
    """A batch of two large matrices (one after the other on the GPU), and
    structured large inputs: diagonal (every reflector trivial: tau = 0)
    and a matrix with an exactly zero column below the diagonal."""
    r = np.random.default_rng(71)
    n = LDS_MAX[cplx] + 40
    As = np.stack([_herm(r, n, cplx) for _ in range(2)])
    Eb = gs.heev_batched(As, wantV=True)
    for k in range(2):
        wo, Vo, info = oracle.heev(As[k], "L")
        _check_eig(As[k], Eb[0][k], Eb[1][k], wo, Vo)
    D = np.diag(r.standard_normal(n)).astype(As.dtype)
    E = gs.geigen_hermitian(D)
    assert np.array_equal(E.values, np.sort(np.diag(D).real))
    B = _herm(r, n, cplx)
    B[5:, 3] = 0
    B[3, 5:] = 0
    wo, Vo, info = oracle.heev(B, "L")
    E = gs.geigen_hermitian(B)
    _check_eig(B, E.values, E.vectors, wo, Vo)


@pytest.mark.parametrize("cplx", [False, True])
def test_heev_large_phases_bitwise(gs, cplx, monkeypatch):
    """Above the LDS orders the eigenvectors come from a record stream
    written by one producer thread and replayed onto Q's rows over the whole
    GPU; against the same steps one call each (gs_hetrd, gs_steqr's
    one-workgroup iteration on (d, e, Q), a stable sort): bit for bit, also
    with streams too short for the iteration (the in-place finish)."""
    r = np.random.default_rng(72)
    n = LDS_MAX[cplx] + 70
    A = np.stack([_herm(r, n, cplx) for _ in range(2)])
    W, V, info = gs.heev_batched(A)
    assert not info.any()
    for k in range(2):
        d, e, Q = gs.hessenberg_hermitian(A[k])
        w, Z, _, inf1 = gs.steqr_(d, e, Q)
        assert inf1 == 0
        p = np.argsort(w, kind="stable")
        assert np.array_equal(W[k], w[p]), k
        assert np.array_equal(V[k], Z[:, p]), k
    for cap in ("1", "5000"):
        monkeypatch.setenv("GSCHUR_HEEV_LOGCAP", cap)
        W2, V2, info2 = gs.heev_batched(A)
        assert not info2.any()
        assert np.array_equal(W2, W) and np.array_equal(V2, V), cap


def test_steqr_above_lds_log_bitwise(gs, oracle):
    """gs_steqr above order 2048 (its LDS log's range): the rotations go
    through a record stream in HBM and are replayed onto Z's rows over the
    whole GPU, bit for bit the oracle's d and Z (a few rows of Z keep the
    oracle's time down)."""
    r = np.random.default_rng(73)
    n = 2100
    d = r.standard_normal(n)
    e = r.standard_normal(n - 1)
    Z = r.standard_normal((5, n))
    wo, zo, info_o, _ = oracle.steqr(d, e, Z)
    wg, zg, _, info_g = gs.steqr_(d, e, Z)
    assert info_g == info_o == 0
    assert np.array_equal(wg, wo)
    assert np.array_equal(zg, zo)
    wv, _, _, info_v = gs.steqr_(d, e, None)
    assert info_v == 0 and np.array_equal(wv, wo)


def test_heev_above_2048(gs):
    """A real symmetric eigensystem of order 2100 (above the one-workgroup
    log's range): residual, orthogonality and the sorted order."""
    r = np.random.default_rng(74)
    n = 2100
    A = _herm(r, n, False)
    E = gs.geigen_hermitian(A)
    scale = np.abs(A).max()
    assert np.all(np.diff(E.values) >= 0)
    assert np.abs(A @ E.vectors - E.vectors * E.values).max() <= 16 * n * EPS * scale
    assert np.abs(E.vectors.T @ E.vectors - np.eye(n)).max() <= 16 * n * EPS
    assert np.abs(np.linalg.eigvalsh(A) - E.values).max() <= 16 * n * EPS * scale


def _revsort(x):
    """test/wrappers.jl: revsort(λ::Real) = -λ; revsort(λ::Complex) = (-real, -imag)."""
    return (-x.real, -x.imag) if np.iscomplexobj(x) else -x


@pytest.mark.parametrize("kind", ["bare", "hermitian", "symmetric"])
def test_wrappers_sortby(gs, kind):
    """test/wrappers.jl's existence-and-sanity checks: eigen, eigvecs and
    eigvals with sortby = eigsortby and with revsort agree with each other;
    the reversed order reverses values and vectors (up to phases)."""
    r = np.random.default_rng(75)
    n = 24
    if kind == "bare":
        A = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
        E = gs.eigen_(A)
        Er = gs.eigen_(A, sortby=_revsort)
        lam = gs.eigvals_(A)
        lamr = gs.eigvals_(A, sortby=_revsort)
        V, Vr, w, wr = E.vectors, Er.vectors, E.values, Er.values
    else:
        cplx = kind == "hermitian"
        A = _herm(r, n, cplx)
        E = gs.geigen_hermitian(A)
        Er = gs.geigen_hermitian(A, sortby=_revsort)
        lam = gs.geigvals_hermitian(A)
        lamr = gs.geigvals_hermitian(A, sortby=_revsort)
        V, Vr, w, wr = E.vectors, Er.vectors, E.values, Er.values
    assert np.array_equal(lam, w) or np.abs(lam - w).max() <= 1e3 * n * EPS * np.abs(A).max()
    assert np.abs(lamr - lam[::-1]).max() <= 1e3 * n * EPS * np.abs(A).max()
    assert np.abs(wr - w[::-1]).max() <= 1e3 * n * EPS * np.abs(A).max()
    for j in range(n):
        a, b = V[:, j], Vr[:, n - 1 - j]
        p = np.vdot(a, b)
        assert abs(abs(p) - np.linalg.norm(a) * np.linalg.norm(b)) <= 1e-8, j
    assert np.abs(A @ V - V * w).max() <= 1e3 * n * EPS * np.abs(A).max() * max(1.0, np.abs(V).max())
