"""Pins the CPU oracle (oracle/, a restatement of GenericSchur.jl's Schur path).

What pins what (the reference cannot run here: Julia is absent):
  * LAPACK through scipy: xLARFG (householder.jl:12-102 is its translation),
    xGEHRD/xUNGHR (the Hessenberg/Q math), xGEES (for n < 75 it runs
    xLAHQR, the routine GenericSchur.jl:194-952 translates): the oracle must
    reproduce LAPACK's eigenvalue order and, for complex input, T and Z up to
    column phases.
  * mpmath at 256 bits (tests/golden/fixtures.json): eigenvalues.
  * closed forms and invariants: reflector identities, the reference's own
    schurtest/hesstest/vectest/checkblocks/checkeigvals, known-answer cases
    (Godunov, Jordan, the near-degenerate 2x2 of test/real.jl:311-317),
    normal matrices giving (block-)diagonal T.
Each test names the mistake it would catch.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.linalg.lapack as lap

import checks

HERE = os.path.dirname(os.path.abspath(__file__))
EPS = np.finfo(np.float64).eps
TINY = np.finfo(np.float64).tiny


def fixtures():
    with open(os.path.join(HERE, "golden", "fixtures.json")) as f:
        return json.load(f)


def dec(d):
    A = np.array(d["re"], dtype=np.float64)
    if "im" in d:
        A = A + 1j * np.array(d["im"], dtype=np.float64)
    return A


def snorm(x):
    """2-norm without underflow (numpy squares directly)."""
    a = np.abs(x)
    m = a.max()
    return 0.0 if m == 0 else m * float(np.sqrt(np.sum((a / m) ** 2)))


def rand_c(n, seed):
    r = np.random.default_rng(seed)
    return r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))


def rand_r(n, seed):
    return np.random.default_rng(seed).standard_normal((n, n))


# ------------------------------------------------------------ reflectors
@pytest.mark.parametrize("m", [2, 3, 7, 20])
def test_reflector_matches_lapack(oracle, m):
    """householder.jl:12-102 vs LAPACK xLARFG (same algorithm): a wrong sign
    of beta or a conjugated tau shows up here."""
    r = np.random.default_rng(m)
    x = r.standard_normal(m)
    xo, tau = oracle.reflector(x)
    beta, v, t = lap.dlarfg(m, x[0], x[1:].copy())
    assert abs(xo[0] - beta) <= 4 * EPS * abs(beta)
    assert np.allclose(xo[1:], v, rtol=1e-13, atol=1e-15)
    assert abs(tau - t) <= 4 * EPS
    z = r.standard_normal(m) + 1j * r.standard_normal(m)
    zo, tauz = oracle.reflector(z)
    beta, v, t = lap.zlarfg(m, z[0], z[1:].copy())
    assert abs(zo[0] - beta) <= 4 * EPS * abs(beta)
    assert np.allclose(zo[1:], v, rtol=1e-13, atol=1e-15)
    assert abs(tauz - t) <= 8 * EPS


@pytest.mark.parametrize("scale", [1.0, 1e-300, 100 * TINY, 1e300])
def test_reflector_closed_form(oracle, scale):
    """H^H x = beta e1 with beta real, |beta| = ||x||, sign(beta) = -sign(Re x1),
    H unitary; the tiny scales run the sfmin rescaling loop."""
    r = np.random.default_rng(7)
    for cplx in (False, True):
        x = r.standard_normal(6) + (1j * r.standard_normal(6) if cplx else 0)
        x = x * scale
        xo, tau = oracle.reflector(x)
        w = np.concatenate([[1.0], xo[1:]])
        H = np.eye(6) - tau * np.outer(w, w.conj())
        y = H.conj().T @ x
        beta = xo[0]
        assert abs(np.imag(beta)) == 0
        nx = snorm(x)
        assert abs(abs(beta) - nx) <= 8 * EPS * nx
        assert np.sign(np.real(beta)) == -np.sign(np.real(x[0]))
        assert snorm(y[1:]) <= 16 * EPS * nx
        assert np.linalg.norm(H.conj().T @ H - np.eye(6)) <= 16 * EPS


def test_reflector_complex_length_one(oracle):
    """The complex n = 1 reflector is not trivial (householder.jl:59-69): it
    makes a complex scalar real."""
    xo, tau = oracle.reflector(np.array([3.0 + 4.0j]))
    assert xo[0] == pytest.approx(-5.0) and xo[0].imag == 0
    assert (1 - np.conj(tau)) * (3 + 4j) == pytest.approx(-5.0)
    xo, tau = oracle.reflector(np.array([2.0 + 0.0j]))
    assert tau == 0 and xo[0] == 2.0


# -------------------------------------------------------- Hessenberg + Q
@pytest.mark.parametrize("cplx", [False, True])
def test_hessenberg_matches_lapack(oracle, cplx):
    """hessenberg.jl:3-17,150-166 vs xGEHRD/xUNGHR: a transposed update or a
    wrong reflector index makes H or Q differ."""
    n = 20
    A = rand_c(n, 3) if cplx else rand_r(n, 3)
    fac, tau, Q, H = oracle.hessenberg(A)
    if cplx:
        ht, tt, info = lap.zgehrd(A)
        q, info2 = lap.zunghr(ht, tt)
    else:
        ht, tt, info = lap.dgehrd(A)
        q, info2 = lap.dorghr(ht, tt)
    assert np.abs(np.triu(ht, -1) - H).max() <= 1e-12 * np.abs(A).max()
    assert np.abs(q - Q).max() <= 1e-13
    assert np.abs(tau - tt[: n - 1]).max() <= 1e-13


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("s", [1.0, 100 * TINY, np.finfo(np.float64).max / 100])
def test_hesstest(oracle, cplx, s):
    """hesstest (test/complex.jl:36-61, test/real.jl:76-99) incl. the tiny and
    huge scalings of those tests."""
    n = 32
    r = np.random.default_rng(1234)
    A = r.random((n, n)) + (1j * r.random((n, n)) if cplx else 0)
    A = A * s
    _, _, Q, H = oracle.hessenberg(A)
    assert np.all(np.tril(H, -2) == 0)
    na = checks.fnorm(A)
    assert checks.fnorm(A - Q @ H @ Q.conj().T) / (n * na * EPS) < 10
    assert checks.fnorm(np.eye(n) - Q @ Q.conj().T) / (n * EPS) < 10
    if cplx:
        assert np.all(np.diag(H, -1).imag == 0)


# ------------------------------------------------------------ Schur: LAPACK
@pytest.mark.parametrize("n", [4, 16, 40, 70])
def test_complex_schur_matches_zgees(oracle, n):
    """For n < 75 ZGEES runs ZLAHQR, which GenericSchur.jl:194-504 translates:
    same eigenvalue order, T equal up to the column phases of Z."""
    A = rand_c(n, 100 + n)
    t, sdim, w, vs, work, info = lap.zgees(lambda x: 0, A)
    T, Z, wo, inf, _ = oracle.gschur(A)
    assert inf == 0 and info == 0
    assert np.abs(np.diag(T) - np.diag(t)).max() <= 100 * n * EPS * np.abs(A).max()
    assert np.abs(np.abs(T) - np.abs(t)).max() <= 100 * n * EPS * np.abs(A).max()
    ph = np.array([np.vdot(Z[:, j], vs[:, j]) for j in range(n)])
    ph /= np.abs(ph)
    assert np.abs(Z * ph - vs).max() <= 100 * n * EPS


@pytest.mark.parametrize("n", [4, 16, 40, 70])
def test_real_schur_matches_dgees_order(oracle, n):
    """GenericSchur.jl:513-952 vs DGEES (DLAHQR for n < 75): same eigenvalue
    order; values within a few ulps of ||A|| (2x2 blocks may differ by a
    rotation inside the block, the reference keeps the older dlanv2)."""
    A = rand_r(n, 200 + n)
    t, sdim, wr, wi, vs, work, info = lap.dgees(lambda x, y: 0, A)
    T, Z, wo, inf, _ = oracle.gschur(A)
    assert inf == 0 and info == 0
    assert np.abs(wo - (wr + 1j * wi)).max() <= 100 * n * EPS * np.abs(A).max()
    assert checks.checkblocks(T)


# ----------------------------------------------- Schur: invariants, KATs
def _schurtest(oracle, A, tol=20.0, normal=False):
    T, Z, w, info, _ = oracle.gschur(A)
    assert info == 0
    checks.schur_invariants(A, T, Z, w, tol, normal=normal)
    VR = oracle.eigvecs(T, Z, left=False)
    VL = oracle.eigvecs(T, Z, left=True)
    checks.vectest(A, w, VR, VL, tol, normal=normal)
    if not np.iscomplexobj(A):
        Tc, Zc = oracle.triangularize(T, Z)
        checks.schur_invariants(A.astype(complex), Tc, Zc, np.diag(Tc), tol)
    return T, Z, w


@pytest.mark.parametrize("n", [4, 32])
def test_schurtest_random(oracle, n):
    """random general (test/complex.jl:124-129, test/real.jl:127-132)"""
    r = np.random.default_rng(1234 + n)
    _schurtest(oracle, r.random((n, n)) + 1j * r.random((n, n)), 10)
    _schurtest(oracle, r.random((n, n)), 10)


@pytest.mark.parametrize("n", [4, 32])
def test_schurtest_normal(oracle, n):
    """random unitary / orthogonal matrices are normal, so T is (block)
    diagonal (test/complex.jl:130-135, test/real.jl:133-138)"""
    r = np.random.default_rng(99 + n)
    Q, R = np.linalg.qr(r.standard_normal((n, n)) + 1j * r.standard_normal((n, n)))
    U = Q * np.sign(np.diag(R))
    _schurtest(oracle, U, 20, normal=True)
    Q, _ = np.linalg.qr(r.standard_normal((n, n)))
    _schurtest(oracle, Q, 10, normal=True)


@pytest.mark.parametrize("n", [4, 32])
def test_schurtest_jordan(oracle, n):
    """Jordan block (test/complex.jl:221-232, test/real.jl:180-191)"""
    A = np.diag(np.ones(n - 1), 1) + np.eye(n)
    _schurtest(oracle, A.astype(complex))
    _schurtest(oracle, A)


@pytest.mark.parametrize("mag", [1.0, np.finfo(float).max * EPS, TINY / EPS])
def test_schurtest_magnitudes(oracle, mag):
    """magnitudes {1, ovfl*ulp, unfl*ulpinv} of zdrves/ddrves
    (test/complex.jl:187, test/real.jl:178); the small one exercises _scale!"""
    n = 16
    r = np.random.default_rng(5)
    A = (r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))) * mag / n
    T, Z, w, info, _ = oracle.gschur(A)
    assert info == 0
    checks.schur_invariants(A, T, Z, w, 20)
    Ar = r.standard_normal((n, n)) * mag / n
    T, Z, w, info, _ = oracle.gschur(Ar)
    assert info == 0
    checks.schur_invariants(Ar, T, Z, w, 20)


def test_godunov(oracle):
    """Godunov (test/testfuncs.jl:127-142): in double only the backward error
    is meaningful (econd 7e16); the quad-double oracle (~212 bits) must hit the
    exact eigenvalues within 3 delta econd as test/real.jl:142-157 demands at
    80 bits."""
    fx = fixtures()["godunov"]
    A = np.array(fx["A"])
    T, Z, w, info, _ = oracle.gschur(A)
    assert info == 0
    assert np.linalg.norm(Z @ T @ Z.T - A) < 100 * EPS * np.linalg.norm(A)
    vals, w4, info = oracle.eig_hp(A, prec=2)
    assert info == 0
    delta = 2.0 ** -200 * np.linalg.norm(A)
    t = 3 * delta * fx["econd"]
    got = np.sort(vals.real)
    assert np.abs(got - np.array(fx["values"])).max() <= max(t, 1e-30)
    assert np.abs(vals.imag).max() <= max(t, 1e-30)


def test_gs2x2_near_degenerate(oracle):
    """_gs2x2! known-answer test (test/real.jl:311-317)."""
    for case in fixtures()["gs2x2_kat"]:
        B = np.array(case["B"])
        abcd, cssn, w1, w2 = oracle.gs2x2(B[0, 0], B[0, 1], B[1, 0], B[1, 1])
        l1, l2 = case["l1"], case["l2"]
        assert abs(max(w1.real, w2.real) - l1) < 2 * EPS
        assert abs(min(w1.real, w2.real) - l2) < 2 * EPS


def test_gs2x2_standard_form(oracle):
    """dlanv2 output: c = 0, or a = d with b c < 0; the rotation is a
    similarity; eigenvalues a +- i sqrt|b| sqrt|c|."""
    r = np.random.default_rng(11)
    for _ in range(200):
        a, b, c, d = r.standard_normal(4)
        (A, B, C, D), (cs, sn), w1, w2 = oracle.gs2x2(a, b, c, d)
        assert C == 0 or (A == D and B * C < 0)
        G = np.array([[cs, sn], [-sn, cs]])
        M = G @ np.array([[a, b], [c, d]]) @ G.T
        assert np.abs(M - np.array([[A, B], [C, D]])).max() <= 16 * EPS * max(abs(a), abs(b), abs(c), abs(d))
        ev = np.sort_complex(np.linalg.eigvals(np.array([[a, b], [c, d]])))
        got = np.sort_complex(np.array([w1, w2]))
        assert np.abs(ev - got).max() <= 1e-12 * max(1, np.abs(ev).max()) + 1e-7 * (C != 0)


def test_mp256_fixtures(oracle):
    """Eigenvalues against mpmath at 256 bits: the double oracle within
    10 n eps ||A|| cond, the quad-double oracle within 2^-190 ||A|| cond."""
    for case in fixtures()["mp256"]:
        A = dec(case["A"])
        n = case["n"]
        ref = np.array([complex(float(re), float(im)) for re, im in case["eig256"]])
        T, Z, w, info, _ = oracle.gschur(A)
        assert info == 0
        VR = oracle.eigvecs(T, Z, left=False)
        VL = oracle.eigvecs(T, Z, left=True)
        cond = checks.eig_conditions(VR, VL)
        checks.eigen_parity(A, w, ref, cond, factor=10.0)
        vals, w4, info = oracle.eig_hp(A, prec=2)
        perm, err = checks.match_eigenvalues(vals, ref)
        assert np.all(err <= 2.0 ** -190 * np.linalg.norm(A, 2) * cond + 1e-300), (n, err.max())


def test_values_slot_order_real(oracle):
    """values[j] corresponds to T[j, j], the +imag member of a pair first
    (GenericSchur.jl:677-679,794-799)."""
    A = rand_r(24, 17)
    T, Z, w, info, _ = oracle.gschur(A)
    assert checks.checkeigvals(T, w, 20)
    for j in range(23):
        if T[j + 1, j] != 0:
            assert w[j].imag > 0 and w[j + 1].imag < 0


# ----------------------------------------------------------- eigenvectors
@pytest.mark.parametrize("cplx", [False, True])
def test_eigvecs_match_lapack(oracle, cplx):
    """vectors.jl:12-698 vs LAPACK xGEEV vectors, phase-insensitive
    (_chkeigvecs, test/wrappers.jl:12-54)."""
    n = 30
    A = rand_c(n, 8) if cplx else rand_r(n, 8)
    T, Z, w, info, _ = oracle.gschur(A)
    for left in (False, True):
        V = oracle.eigvecs(T, Z, left=left)
        ev, vl, vr = sla.eig(A, left=True, right=True)
        ref = vl if left else vr
        perm, _ = checks.match_eigenvalues(ev, w)
        ref = ref[:, perm]
        p = np.abs(np.sum(ref.conj() * V, axis=0)) / (np.linalg.norm(ref, axis=0) * np.linalg.norm(V, axis=0))
        assert np.all(np.abs(p - 1) < np.sqrt(EPS))


def test_eigvecs_normalization(oracle):
    """_enormalize! (util.jl:572-592): unit 2-norm, largest entry real and
    positive (test/wrappers.jl:181-192)."""
    A = np.array(fixtures()["normalization"]["A"], dtype=float)
    for M in (A, A + 0.5j * A.T):
        T, Z, w, info, _ = oracle.gschur(M)
        for left in (False, True):
            V = oracle.eigvecs(T, Z, left=left)
            assert np.all(np.abs(np.linalg.norm(V, axis=0) - 1) / (3 * EPS) < 10)
            for j in range(3):
                i0 = np.argmax(np.abs(V[:, j]))
                assert V[i0, j].imag == 0 and V[i0, j].real > 0


@pytest.mark.parametrize("dmode", [0, 1, 2])
def test_eigvecs_triangular_repeated(oracle, dmode):
    """triangular, specified eigenvalues with an exact / next / previous
    repeat (test/complex.jl:327-364): exercises the smin perturbation of the
    scaled solves."""
    n = 32
    r = np.random.default_rng(31 + dmode)
    d = r.random(n) + 1j * r.random(n)
    if dmode == 0:
        d[1] = d[n - 2]
    elif dmode == 1:
        d[1] = d[n - 2] * np.nextafter(1.0, 2.0)
    else:
        d[1] = d[n - 2] * np.nextafter(1.0, 0.0)
    A = np.triu(r.random((n, n)) + 1j * r.random((n, n)), 1) + np.diag(d)
    VR = oracle.eigvecs(A, np.eye(n, dtype=complex), left=False)
    VL = oracle.eigvecs(A, np.eye(n, dtype=complex), left=True)
    checks.vectest(A, np.diag(A), VR, VL, 20)


def test_eigvecs_nonfinite_no_crash(oracle):
    """test/complex.jl:393-406: non-finite T must not throw."""
    A = rand_c(5, 3)
    T, Z, w, info, _ = oracle.gschur(A)
    for bad in (np.inf, np.nan):
        T2 = T.copy()
        T2[1, 2] = bad
        oracle.eigvecs(T2, Z, left=False)
        oracle.eigvecs(T2, Z, left=True)


# ------------------------------------------------------------ error paths
def test_errors(oracle):
    """UnconvergedException at maxiter (GenericSchur.jl:234-236,554-556) and
    ArgumentError for a non-real subdiagonal (test/errors.jl:1-5)."""
    A = rand_c(8, 1)
    assert oracle.gschur(A, maxiter=1)[3] == oracle.UNCONVERGED
    assert oracle.gschur(rand_r(8, 1), maxiter=1)[3] == oracle.UNCONVERGED
    An = A.copy()
    An[2, 3] = np.nan
    assert oracle.gschur(An)[3] == oracle.UNCONVERGED
    n = 5
    H = np.diag(np.full(n - 1, -1.0 + 1.0j), -1) + np.triu(np.random.default_rng(0).random((n, n)))
    assert oracle.schur_hess(H, checksd=True)[3] == oracle.NONREAL_SUB


def test_qd_arithmetic(oracle):
    """The quad-double type against mpmath at 256 bits."""
    import mpmath as mp
    mp.mp.prec = 256

    def val(r):
        return sum(mp.mpf(float(x)) for x in r)

    r = oracle.qd_op("sqrt", [2.0, 0, 0, 0])
    assert abs(val(r) - mp.sqrt(2)) < mp.mpf(2) ** -205
    r = oracle.qd_op("div", [1.0, 0, 0, 0], [3.0, 0, 0, 0])
    assert abs(val(r) - mp.mpf(1) / 3) < mp.mpf(2) ** -205
    a = oracle.qd_op("div", [1.0, 0, 0, 0], [7.0, 0, 0, 0])
    b = oracle.qd_op("sqrt", [3.0, 0, 0, 0])
    r = oracle.qd_op("mul", a, b)
    assert abs(val(r) - mp.sqrt(3) / 7) < mp.mpf(2) ** -204
    r = oracle.qd_op("add", a, [-a[0], -a[1], -a[2], 0.0])
    assert abs(val(r) - val(a) + val([a[0], a[1], a[2], 0.0])) < mp.mpf(2) ** -250


# -------------------------------------------------------- test families
@pytest.mark.parametrize("cplx", [False, True])
def test_families_vs_lapack(oracle, cplx):
    """Every re-created family of the reference's test suite (tests/
    families.py, at the reference's orders 4 and 32 and at 96): the oracle's
    eigenvalues against LAPACK's xGEEV on a copy scaled exactly (by a power of
    2) to norm 1, within 10 n eps ||A||_2 cond(lambda) (clusters of
    overlapping error discs as a whole), and the reference's schurtest.  A
    slip in the oracle's scaling (_scale!, util.jl:14-29), deflation or
    shift logic shows up on one of these families."""
    import families
    for n in (4, 32, 96):
        for name, A, normal in families.families(n, cplx, seed=5):
            T, Z, w, cond, _ = families.oracle_result(oracle, A)
            checks.schur_invariants(A, T, Z, w, 20.0, normal=normal)
            m = np.abs(A).max()
            s = 1.0 if m == 0 else 2.0 ** -np.round(np.log2(m))
            ev = sla.eigvals(A * s) / s
            checks.eigen_parity(A, ev, w, cond)


# ------------------------------------------------------------ balancing
def _bal_cases():
    r = np.random.default_rng(17)
    out = []
    for cplx in (False, True):
        for trial in range(6):
            n = (5, 12, 30)[trial % 3]
            A = r.standard_normal((n, n)) * np.exp(3 * r.standard_normal((n, 1))) * np.exp(3 * r.standard_normal((1, n)))
            if trial >= 3:
                A[np.abs(A) < 1.0] = 0.0  # zeros: row / column permutations
            if cplx:
                A = A + 1j * A[::-1]
            out.append(A)
    return out


def test_balance_matches_lapack(oracle):
    """balance.jl:33-199 vs LAPACK xGEBAL (the routine it translates): the
    balanced matrix, ilo/ihi, the scale factors and the permutation record
    bit for bit.  A wrong loop bound of the permutation search, a swapped
    row/column exchange or a wrong scaling test changes one of them."""
    for A in _bal_cases():
        ab, bal, sp, d, info = oracle.balance(A)
        assert info == 0
        f = lap.zgebal if np.iscomplexobj(A) else lap.dgebal
        ba, lo, hi, ps, inf = f(A, scale=1, permute=1)
        assert (bal[0], bal[1]) == (lo + 1, hi + 1)
        assert np.array_equal(ab, ba)
        assert np.array_equal(d[bal[0] - 1:bal[1]], ps[lo:hi + 1])
        outside = np.r_[0:lo, hi + 1:A.shape[0]]
        assert np.array_equal(sp[outside], ps[outside].astype(np.int64))


def test_balance_sanity_block_structure(oracle):
    """test/balance.jl:74-106: a permuted block upper-triangular matrix
    [T1 0 0; 0 A2 0; 0 0 T3] balances to upper triangular outside the
    ilo:ihi block."""
    r = np.random.default_rng(3)
    for n1 in (0, 1, 2):
        for n2 in (0, 1, 2):
            for n3 in (0, 1, 2):
                n = n1 + n2 + n3
                if n == 0:
                    continue
                A = np.zeros((n, n))
                A[:n1, :n1] = np.triu(r.random((n1, n1)))
                A[n1:n1 + n2, n1:n1 + n2] = r.random((n2, n2))
                A[n1 + n2:, n1 + n2:] = np.triu(r.random((n3, n3)))
                P = np.eye(n)[r.permutation(n)]
                Ax = P.T @ A @ P
                C, bal, sp, d, info = oracle.balance(Ax)
                ilo, ihi = bal[0], bal[1]
                assert ilo <= ihi
                for i in range(1, n + 1):
                    for j in range(1, i):
                        if i < ilo or j > ihi:
                            assert C[i - 1, j - 1] == 0


def unbal_classic(cplx):
    """test/balance.jl:28-42: a badly balanced 3 x 3 matrix with known
    asymptotic eigenvalues."""
    m = np.round(-2 * np.log2(EPS) / 5)
    x = 2.0 ** m
    A = np.array([[1, 0, x ** -2], [1, 1, x ** -1], [x ** 2, x, 1]], dtype=np.complex128 if cplx else np.float64)
    rt2 = np.sqrt(2.0)
    lam = np.array([1 - rt2 + 1 / (4 * x), 1 - 1 / (2 * x), 1 + rt2 + 1 / (4 * x)])
    return A, lam


@pytest.mark.parametrize("cplx", [False, True])
def test_balance_unbal_classic(oracle, cplx):
    """test/balance.jl:58-72: eigenvalues of the balanced matrix are far more
    accurate (ru > 100 rb), and the left eigenvectors of the balanced
    matrix, back-transformed by ldiv!(B, Vl), are left eigenvectors of A."""
    A, lam = unbal_classic(cplx)
    wu = np.sort(oracle.gschur(A)[2].real)
    Ab, bal, sp, d, info = oracle.balance(A)
    T, Z, wb, info2, _ = oracle.gschur(Ab)
    ru = np.abs(wu - lam).max()
    rb = np.abs(np.sort(wb.real) - lam).max()
    assert ru > 100 * rb
    if cplx:
        Vl = oracle.eigvecs(T, Z, left=True)
        Vl = oracle.balance_apply(Vl, bal, sp, d, inverse=True)
        assert np.allclose(A.conj().T @ Vl, Vl @ np.diag(wb.conj()), rtol=1e-8, atol=1e-8)


def test_eigen_composition_matches_lapack(oracle):
    """eigen!(A) (pirates.jl:63-160) composed from the restated pieces:
    sorted eigenvalues and unit eigenvectors (largest entry real) against
    scipy's xGEEV, phase-insensitive, for balanced and unbalanced input."""
    r = np.random.default_rng(5)
    for cplx in (False, True):
        for n in (6, 25):
            A = r.standard_normal((n, n)) * np.exp(2 * r.standard_normal((n, 1))) * np.exp(-2 * r.standard_normal((1, n)))
            if cplx:
                A = A + 1j * r.standard_normal((n, n))
            for bal in (True, False):
                w, V, VL, rc = oracle.eigen(A, permute=bal, scale=bal, jvl=True, jce=True)
                ev, vl, vr = sla.eig(A, left=True, right=True)
                perm, err = checks.match_eigenvalues(ev, w)
                assert err.max() <= 1e-9 * np.abs(ev).max()
                for ref, got in ((vr[:, perm], V), (vl[:, perm], VL)):
                    p = np.abs(np.sum(ref.conj() * got, axis=0)) / np.linalg.norm(ref, axis=0)
                    assert np.all(np.abs(p - 1) < 1e-7)
                assert np.all(np.abs(np.linalg.norm(V, axis=0) - 1) < 1e-13)
                assert np.all((rc > 0) & (rc <= 1 + 1e-12))
                # rconde (pirates.jl:112-120) is |y^H x| / (|y| |x|) of the
                # matrix the eigenvectors belong to, the balanced one with
                # balancing (as xGEEVX's RCONDE): the closed form from
                # LAPACK's eigenvectors of that matrix
                Ab = oracle.balance(A, scale=bal, permute=bal)[0] if bal else A
                eb, lb, rb = sla.eig(Ab, left=True, right=True)
                pb, _ = checks.match_eigenvalues(eb, w)
                ref_rc = np.abs(np.sum(lb.conj() * rb, axis=0)) / np.linalg.norm(lb, axis=0) / np.linalg.norm(rb, axis=0)
                assert np.allclose(rc, ref_rc[pb], rtol=1e-8, atol=1e-12)


def test_balance_nan_error(oracle):
    """A NaN entry is the reference's "NaN encountered while balancing"
    (balance.jl:160-162) rather than a sweep loop that never settles."""
    A = np.array([[1.0, 2.0, 0.5], [np.nan, 3.0, 1.0], [4.0, 1.0, 2.0]])
    assert oracle.balance(A)[4] == -3


# ------------------------------------------ reordering and conditioning
def test_givens_matches_zlartg(oracle):
    """LinearAlgebra.givensAlgorithm (the rotation of _trexchange!,
    ordschur.jl:62) against LAPACK's xLARTG: the same c, s, r to rounding,
    [c s; -conj(s) c] [f; g] = [r; 0], c^2 + |s|^2 = 1, including the
    scaled ranges and the zero cases."""
    r = np.random.default_rng(31)
    cases = [(1 + 2j, 3 - 1j), (0j, 2 + 1j), (2 - 1j, 0j), (1e-300 + 1e-300j, 3e-300j), (1e300 - 2e300j, 5e299j),
             (1e-160j, 1.0 + 0j), (3.0 + 0j, 1e-170 + 0j)]
    cases += [complex(*r.standard_normal(2)) * 10.0 ** r.uniform(-5, 5) for _ in range(20)]
    cases = [(c, complex(*r.standard_normal(2))) if not isinstance(c, tuple) else c for c in cases]
    for f, g in cases:
        c, s, rr = oracle.givens(f, g)
        cl, sl, rl = sla.lapack.zlartg(f, g)
        scale = max(abs(f), abs(g))
        assert abs(c * c + abs(s) ** 2 - 1) <= 8 * EPS
        assert abs(rr - (c * f + s * g)) <= 8 * EPS * scale
        assert abs(-np.conj(s) * f + c * g) <= 8 * EPS * scale
        assert abs(c - cl) <= 8 * EPS and abs(s - sl) <= 8 * EPS and abs(rr - rl) <= 8 * EPS * scale


@pytest.mark.parametrize("adjoint", [False, True])
def test_trsylvester_residual(oracle, adjoint):
    """trsylvester! / adjtrsylvester! (sylvester.jl:24-139): the residual of
    A X - X B = scale C (or A^H X - X B^H = scale C) at rounding level, and
    X against scipy's solve_sylvester on the same well-separated triangular
    blocks."""
    r = np.random.default_rng(32)
    for m, n in ((1, 1), (3, 5), (12, 7), (20, 20)):
        A = np.triu(r.standard_normal((m, m)) + 1j * r.standard_normal((m, m))) + 4 * np.eye(m)
        B = np.triu(r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))) - 4 * np.eye(n)
        C = r.standard_normal((m, n)) + 1j * r.standard_normal((m, n))
        X, sc, pert = oracle.trsylvester(A, B, C, adjoint=adjoint)
        assert sc == 1.0 and not pert
        Aop, Bop = (A.conj().T, B.conj().T) if adjoint else (A, B)
        res = np.linalg.norm(Aop @ X - X @ Bop - sc * C) / (np.linalg.norm(Aop) + np.linalg.norm(Bop)) / np.linalg.norm(X)
        assert res <= 20 * max(m, n) * EPS
        Xs = sla.solve_sylvester(Aop, -Bop, C)
        assert np.linalg.norm(X - Xs) <= 1e-10 * np.linalg.norm(Xs)


def _ztrsen(T, Z, select):
    n = T.shape[0]
    m = int(np.count_nonzero(select))
    out = sla.lapack.ztrsen(np.asarray(select, dtype=np.int32), T, Z, job="B", lwork=max(1, 2 * m * (n - m)))
    ts, qs, w, mm, s, sep, info = out
    assert info == 0
    return ts, qs, w, s, sep


@pytest.mark.parametrize("n", [6, 32])
def test_ordschur_matches_ztrsen(oracle, n):
    """gordschur! on a complex Schur form (the reference's checkord,
    test/ordschur.jl:1-29: triangular, backward error, unitary Z, the selected
    eigenvalues first) and against LAPACK's xTRSEN on the same (T, Z): the
    same diagonal order and T, Z equal up to diagonal phases; eigvalscond and
    subspacesep against xTRSEN's s and sep (test/ordschur.jl:31-56)."""
    r = np.random.default_rng(33 + n)
    A = r.standard_normal((n, n)) + 1j * r.standard_normal((n, n))
    T, Z, w, info, _ = oracle.gschur(A)
    assert info == 0
    for trial in range(3):
        if trial == 0:
            select = np.ones(n, dtype=bool)
            select[: n // 2] = False
        else:
            select = np.zeros(n, dtype=bool)
            select[r.choice(n, size=max(1, n // 4), replace=False)] = True
        T2, Z2, w2 = oracle.ordschur(T, Z, select)
        assert np.all(np.tril(T2, -1) == 0)
        assert np.linalg.norm(Z2 @ T2 @ Z2.conj().T - A) / (n * np.linalg.norm(A) * EPS) < 10
        assert np.linalg.norm(np.eye(n) - Z2 @ Z2.conj().T) / (n * EPS) < 10
        k = int(select.sum())
        want = w[select]
        assert np.all(np.abs(want - w2[:k]) <= 10 * EPS * np.abs(w).max())
        ts, qs, wl, s, sep = _ztrsen(np.asfortranarray(T), np.asfortranarray(Z), select)
        assert np.all(np.abs(wl - w2) <= 100 * n * EPS * np.abs(w).max())
        # up to diagonal phases: T2 = D^H Ts D, Z2 = Qs D
        ph = np.array([np.vdot(qs[:, j], Z2[:, j]) for j in range(n)])
        ph = ph / np.abs(ph)
        assert np.abs(Z2 - qs * ph).max() <= 1e-10
        assert np.abs(T2 - np.conj(ph)[:, None] * ts * ph[None, :]).max() <= 1e-10 * np.abs(T).max()
        s2 = oracle.eigvalscond(T2, k)
        sep2 = oracle.subspacesep(T2, k)
        assert abs(s2 - s) <= 1e-8 * s
        assert abs(sep2 - sep) <= 1e-8 * sep


# ------------------------------------------------------ Hermitian path pins
# The Hermitian / real symmetric path (hessenberg.jl:19-232, symtridiag.jl):
# LAPACK xSYTRD/xHETRD (the algorithm _hehessenberg! translates) pins d, e,
# tau and the reflector layout of both triangles; the reference's own
# symmetric tridiagonal test classes and error measures (test/symtridiag.jl)
# and scipy's tridiagonal eigensolver pin the QL/QR iteration; mpmath pins
# Julia's hypot; the old/new xLARTG sign relation pins givensAlgorithm.

import symtri_cases as sc


def _lapack_q(c, tau, lower):
    """Q of xHETRD's reflectors (LAPACK's documented layout)."""
    n = c.shape[0]
    Q = np.eye(n, dtype=c.dtype)
    if lower:
        for i in range(n - 1):
            v = np.zeros(n, dtype=c.dtype)
            v[i + 1] = 1
            v[i + 2:] = c[i + 2:, i]
            Q = Q @ (np.eye(n) - tau[i] * np.outer(v, v.conj()))
    else:
        for i in range(n - 2, -1, -1):
            v = np.zeros(n, dtype=c.dtype)
            v[i] = 1
            v[:i] = c[:i, i + 1]
            Q = Q @ (np.eye(n) - tau[i] * np.outer(v, v.conj()))
    return Q


def test_jhypot_correctly_rounded(oracle):
    """Julia's hypot (the fma branch) is correctly rounded: compare with
    mpmath at 200 bits over random, widely varying, near-overflow and
    subnormal operands (a dropped correction term or scaling fails)."""
    import mpmath
    mpmath.mp.prec = 200
    r = np.random.default_rng(40)
    cases = [(3.0, 4.0), (0.0, 0.0), (0.0, -2.5), (1e308, 1e308), (1e-310, 3e-310), (5e-324, 5e-324),
             (1.0, 1e-9), (1.0, 1e-8), (np.inf, 1.0), (-2.0, np.inf)]
    for _ in range(300):
        x, y = r.standard_normal(2) * 10.0 ** r.uniform(-300, 300, 2)
        cases.append((x, y))
        cases.append((x, x * (1 + r.uniform(-1e-3, 1e-3))))
    for x, y in cases:
        got = oracle.jhypot(x, y)
        if np.isinf(x) or np.isinf(y):
            assert got == np.inf
            continue
        want = float(mpmath.sqrt(mpmath.mpf(x) ** 2 + mpmath.mpf(y) ** 2))
        assert got == want, (x, y, got, want)


def test_givens_re_matches_dlartg(oracle):
    """givensAlgorithm(f, g) for real input (LAPACK 3.1 xLARTG): the rotation
    identities, c > 0 when |f| > |g|, r > 0 otherwise, and LAPACK's current
    xLARTG (c >= 0, sign(r) = sign(f)) equal up to that sign convention,
    including the floatmin2-scaled ranges and the zero cases."""
    r = np.random.default_rng(41)
    cases = [(1.0, 2.0), (-1.0, 2.0), (2.0, -1.0), (-2.0, 1.0), (0.0, 3.0), (3.0, 0.0), (0.0, -3.0), (-3.0, 0.0),
             (1e300, -3e300), (-2e-300, 1e-300), (1e-320, 2e-320), (5e299, 5e-200)]
    for _ in range(200):
        f, g = r.standard_normal(2) * 10.0 ** r.uniform(-200, 200, 2)
        cases.append((f, g))
    for f, g in cases:
        c, s, rr = oracle.givens_re(f, g)
        scale = max(abs(f), abs(g))
        assert abs(c * c + s * s - 1) <= 4 * EPS
        assert abs(c * f + s * g - rr) <= 4 * EPS * scale
        assert abs(-s * f + c * g) <= 4 * EPS * scale
        if g == 0:
            assert (c, s, rr) == (1.0, 0.0, f)
            continue
        if f == 0:
            assert (c, s, rr) == (0.0, 1.0, g)
            continue
        if abs(f) > abs(g):
            assert c > 0
        else:
            assert rr > 0
        cl, sl, rl = lap.dlartg(f, g)
        sign = -1.0 if (abs(f) <= abs(g) and f < 0) else 1.0
        assert abs(c - sign * cl) <= 4 * EPS and abs(s - sign * sl) <= 4 * EPS
        assert abs(rr - sign * rl) <= 4 * EPS * scale


def test_sym2x2_closed_form(oracle):
    """_dgz_sym2x2 (symtridiag.jl:375-427; xLAEV2): rt1 the eigenvalue of
    larger modulus, rt2 the other, (cs, sn) the unit eigenvector of rt1 --
    against numpy's eigvalsh, including b = 0, a = c and a + c = 0."""
    r = np.random.default_rng(42)
    cases = [(1.0, 0.0, 2.0), (2.0, 0.0, 1.0), (1.0, 1.0, 1.0), (1.0, 2.0, -1.0), (0.0, 1.0, 0.0),
             (-3.0, 1e-9, -3.0), (1e-200, 1e-200, 2e-200), (5.0, 0.0, -5.0)]
    cases += [tuple(r.standard_normal(3) * 10.0 ** r.uniform(-3, 3)) for _ in range(100)]
    for a, b, c in cases:
        rt1, rt2, cs, sn = oracle.sym2x2(a, b, c)
        M = np.array([[a, b], [b, c]])
        ev = np.linalg.eigvalsh(M)
        scale = np.abs(M).max()
        assert abs(rt1) >= abs(rt2) - 4 * EPS * scale
        assert np.abs(np.sort([rt1, rt2]) - ev).max() <= 8 * EPS * scale
        assert abs(cs * cs + sn * sn - 1) <= 4 * EPS
        x = np.array([cs, sn])
        assert np.abs(M @ x - rt1 * x).max() <= 8 * EPS * scale


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("uplo", ["L", "U"])
def test_hetrd_matches_lapack(oracle, cplx, uplo):
    """_hehessenberg! (hessenberg.jl:27-95) and _materializeQ (:133-232)
    against LAPACK xSYTRD/xHETRD on the same triangle: d, e and Q to
    rounding (the reflector conventions and the U form's reversed sweep),
    plus hesstest's decomposition and orthogonality bounds
    (test/complex.jl:63-83, tol 10) at scale 1 and 100 floatmin."""
    r = np.random.default_rng(43)
    for n in (1, 2, 3, 7, 24, 40):
        A = r.standard_normal((n, n)) + (1j * r.standard_normal((n, n)) if cplx else 0)
        A = A + A.conj().T
        d, e, Q = oracle.hetrd(A, uplo)
        f = lap.zhetrd if cplx else lap.dsytrd
        c, dl, el, tau, info = f(A, lower=int(uplo == "L"))
        assert info == 0
        scale = np.abs(A).max()
        assert np.abs(d - dl).max() <= 64 * n * EPS * scale
        if n > 1:
            assert np.abs(e - el).max() <= 64 * n * EPS * scale
            Ql = _lapack_q(c, tau, uplo == "L")
            assert np.abs(Q - Ql).max() <= 64 * n * EPS
        for s in (1.0, 100 * TINY):
            As = A * s
            d2, e2, Q2 = oracle.hetrd(As, uplo)
            T = np.diag(d2) + np.diag(e2, 1) + np.diag(e2, -1)
            # norms taken of the rescaled matrices (numpy's norm squares entries)
            nrm = np.linalg.norm(As / s)
            assert np.linalg.norm((As - Q2 @ T @ Q2.conj().T) / s) / (n * nrm * EPS) < 10
            assert np.linalg.norm(np.eye(n) - Q2 @ Q2.conj().T) / (n * EPS) < 10


def test_steqr_reference_classes(oracle):
    """_gschur!(SymTridiagonal, QRIteration(), Z) on the reference's own test
    classes (test/symtridiag.jl:47-190, n = 32): runtest's d_err and v_err
    below 50, and the eigenvalues against scipy's tridiagonal solver (a
    dropped shift term, a wrong rotation sign or a lost scaling fails)."""
    r = np.random.default_rng(44)
    tri = lambda A: oracle.hetrd(A, "L", wantq=False)[:2]
    mats = [sc.class1(32, t, r, tri)[:2] for t in range(1, 10)] + [sc.class2(32, t)[:2] for t in range(0, 9)]
    for d, e in mats:
        n = d.shape[0]
        w, Z, info, sweeps = oracle.steqr(d, e, np.eye(n))
        assert info == 0
        de, ve = sc.errors(d, e, w, Z)
        assert de < 50 and ve < 50
        ref = sla.eigh_tridiagonal(d, e, eigvals_only=True)
        scale = max(np.abs(d).max(), np.abs(e).max() if e.size else 0.0)
        assert np.abs(np.sort(w) - ref).max() <= 16 * n * EPS * scale
        # eigenvalues alone: the same values, bit for bit (Z does not steer the iteration)
        w2, _, _, _ = oracle.steqr(d, e, None)
        assert np.array_equal(w, w2)


def test_steqr_sweep_limit(oracle):
    """The iteration stops after maxsweeps sweeps in all (the reference's own
    counter never advances, symtridiag.jl:141) and reports the unconverged
    off-diagonal entries, LAPACK's info."""
    d, e, _ = sc.class2(20, 2)
    w, Z, info, sweeps = oracle.steqr(d, e, np.eye(20), maxsweeps=3)
    assert sweeps == 3 and info > 0
    w, Z, info, sweeps = oracle.steqr(d, e, np.eye(20))
    assert info == 0 and sweeps > 3


@pytest.mark.parametrize("cplx", [False, True])
def test_heev_eigensystem(oracle, cplx):
    """geigen! (GenericSchur.jl:124-135): eigenvalues ascending against
    numpy's eigh, A V = V diag(w), V unitary, both triangles giving the same
    eigenvalues, the iteration bit-identical with and without vectors, and
    sorteig!'s stable order on repeated eigenvalues (identity, diagonal)."""
    r = np.random.default_rng(45)
    for n in (1, 2, 5, 33, 64):
        A = r.standard_normal((n, n)) + (1j * r.standard_normal((n, n)) if cplx else 0)
        A = A + A.conj().T
        w, V, info = oracle.heev(A, "L")
        assert info == 0
        scale = np.abs(A).max()
        assert np.all(np.diff(w) >= 0)
        assert np.abs(w - np.linalg.eigvalsh(A)).max() <= 16 * n * EPS * scale
        assert np.abs(A @ V - V * w).max() <= 16 * n * EPS * scale
        assert np.abs(V.conj().T @ V - np.eye(n)).max() <= 16 * n * EPS
        wu, _, _ = oracle.heev(A, "U")
        assert np.abs(wu - w).max() <= 16 * n * EPS * scale
        w0, _, _ = oracle.heev(A, "L", wantv=False)
        assert np.array_equal(w0, w)
    w, V, _ = oracle.heev(np.eye(4), "L")
    assert np.array_equal(w, np.ones(4)) and np.array_equal(V, np.eye(4))
    D = np.diag([3.0, -1.0, 2.0, -1.0])
    w, V, _ = oracle.heev(D, "L")
    assert np.array_equal(w, [-1.0, -1.0, 2.0, 3.0])
    assert np.array_equal(np.abs(V), np.eye(4)[:, [1, 3, 2, 0]])


# ------------------------------------------------ real reordering pins
# gordschur! for real T (rordschur.jl:3-466): LAPACK xTRSEN (whose xTREXC /
# xLAEXC / xLASY2 the reference translates) on the same (T, Z) must give the
# same eigenvalue order and the same T and Z up to the signs the two xLARTG
# conventions give; the reference's own checkord (test/ordschur.jl:89-133);
# _syl1or2 against the Sylvester residual and scipy's solver.


def _real_schur_pairs(A):
    T, Z, *_ = sla.schur(A)
    return np.asfortranarray(T), np.asfortranarray(Z)


@pytest.mark.parametrize("n1,n2", [(1, 2), (2, 1), (2, 2)])
def test_syl1or2_residual(oracle, n1, n2):
    """_syl1or2 (xLASY2): Tl X - X Tr = scale B at rounding level and X
    against scale times scipy's solve_sylvester (a transposed operand, a
    wrong pivot table entry or a sign slip fails).  scale = 1 for the 1 x 2
    and 2 x 1 forms on well-scaled data; the 4 x 4 form compares 8 smallnum
    |b| with the signed pivot (rordschur.jl:406-411, where xLASY2 takes its
    absolute value), so a negative pivot gives scale = 1/(8 max |b|)."""
    r = np.random.default_rng(50 + 3 * n1 + n2)
    for _ in range(200):
        Tl = r.standard_normal((n1, n1)) * 2
        Tr = r.standard_normal((n2, n2)) * 2 + 6 * np.eye(n2)
        B = r.standard_normal((n1, n2))
        X, scale = oracle.syl1or2(Tl, Tr, B)
        res = Tl @ X - X @ Tr - scale * B
        assert np.abs(res).max() <= 64 * EPS * (np.abs(Tl).max() + np.abs(Tr).max()) * np.abs(X).max() + 64 * EPS
        if n1 + n2 == 3:
            assert scale == 1.0
        else:
            assert 0.0 < scale <= 1.0
        Xs = sla.solve_sylvester(Tl, -Tr, B)
        assert np.abs(X - scale * Xs).max() <= 1e-10 * max(1.0, np.abs(Xs).max())


@pytest.mark.parametrize("n", [6, 20, 40])
def test_ordschur_real_matches_dtrsen(oracle, n):
    """The real gordschur! against LAPACK xTRSEN: the same eigenvalue order
    and |T|, |Z| equal to rounding (LAPACK 3.10's xLARTG differs only in the
    rotation's sign), eigenvalues as the reference computes them from the
    standard blocks."""
    r = np.random.default_rng(46 + n)
    A = r.standard_normal((n, n))
    T, Z = _real_schur_pairs(A)
    for trial in range(4):
        sel = r.random(n) < 0.35
        T2, Z2, w, info = oracle.ordschur_re(T, Z, sel)
        assert info == 0
        out = lap.dtrsen(sel.astype(np.int32), T, Z, job="N", lwork=max(1, n), liwork=1)
        Tl, Zl, wr, wi, info_l = out[0], out[1], out[2], out[3], out[-1]
        assert info_l == 0
        assert np.abs(w - (wr + 1j * wi)).max() <= 100 * n * EPS * np.abs(T).max()
        assert np.abs(np.abs(T2) - np.abs(Tl)).max() <= 1e-12 * np.abs(T).max()
        assert np.abs(np.abs(Z2) - np.abs(Zl)).max() <= 1e-12


def test_ordschur_real_checkord(oracle):
    """test/ordschur.jl:89-133 (n = 20, Float64): two reals, one complex pair,
    a real and a pair, two pairs moved to the front; quasi-triangular,
    backward error and orthogonality below 10, checkeigvals, standard blocks,
    the wanted eigenvalues first."""
    ulp = EPS
    r = np.random.default_rng(47)
    n = 20
    A = r.random((n, n))
    T, Z, wv, info, _ = oracle.gschur(A)
    assert info == 0
    for nr, ni in ((2, 0), (0, 2), (1, 2), (0, 4)):
        select = np.zeros(n, dtype=bool)
        i = n
        while nr + ni > 0 and i > 1:
            i -= 1
            if nr > 0 and wv[i].imag == 0:
                select[i] = True
                nr -= 1
            if ni > 0:
                iwi = wv[i].imag
                if abs(iwi) > 1000 * ulp and wv[i - 1].imag == -iwi:
                    select[i] = True
                    select[i - 1] = True
                    ni -= 2
                    i -= 1
        if nr + ni > 0:
            continue
        T2, Z2, w2, info = oracle.ordschur_re(T, Z, select)
        assert info == 0
        assert np.all(np.tril(T2, -2) == 0)
        assert np.linalg.norm(Z2 @ T2 @ Z2.T - A) / (n * np.linalg.norm(A) * ulp) < 10
        assert np.linalg.norm(np.eye(n) - Z2 @ Z2.T) / (n * ulp) < 10
        assert checks.checkeigvals(T2, w2, 10)
        assert checks.checkblocks(T2)
        want = wv[select]
        wscale = np.abs(w2).max()
        for k in range(int(select.sum())):
            assert np.min(np.abs(want - w2[k])) / (ulp + wscale) < 10
