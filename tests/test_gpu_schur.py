"""GPU parity of the Schur kernels against the CPU oracle.

Each case runs the HIP path through the C ABI (via the genericschur package)
and checks (a) the reference's own schurtest invariants (test/complex.jl:1-34,
test/real.jl:24-74) at the reference's tolerance 20 (LAPACK's), and (b) the
eigenvalues against the oracle's within 10 n eps ||A||_2 cond(lambda), with
cond from the oracle's eigenvectors.  The north-star acceptance bounds
||A - Z T Z^H||_F / ||A||_F <= 50 n eps and ||Z^H Z - I||_F <= 50 n eps are
checked as well.
"""
import numpy as np
import pytest

import checks
from genericschur import gen

pytestmark = pytest.mark.gpu

TOL = 20.0


def _oracle_ref(oracle, A):
    T, Z, w, info, _ = oracle.gschur(A)
    assert info == 0
    VR = oracle.eigvecs(T, Z, left=False)
    VL = oracle.eigvecs(T, Z, left=True)
    return w, checks.eig_conditions(VR, VL)


def _accept(A, S):
    n = A.shape[0]
    eps = np.finfo(np.float64).eps
    assert checks.rel_backward(A, S.T, S.Z) <= 50 * n * eps
    assert checks.rel_orth(S.Z) <= 50 * n * eps


def _check(oracle, A, S, normal=False):
    print(f"case n={A.shape[0]} dtype={A.dtype}", flush=True)
    checks.schur_invariants(A, S.T, S.Z, S.values, TOL, normal=normal)
    if A.shape[0] > 0:
        _accept(A, S)
        ref_w, cond = _oracle_ref(oracle, A)
        checks.eigen_parity(A, S.values, ref_w, cond)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 8, 17, 32, 64, 96, 99])
def test_complex_single_lds(gs, oracle, n):
    rng = gen.stream(11, 100, n)
    A = gen.ginibre_complex(n, rng)
    S = gs.gschur(A)
    _check(oracle, A, S)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 8, 17, 32, 64, 127, 141])
def test_real_single_lds(gs, oracle, n):
    rng = gen.stream(12, 100, n)
    A = gen.ginibre_real(n, rng)
    S = gs.gschur(A)
    _check(oracle, A, S)


@pytest.mark.parametrize("n", [3, 40, 130])
def test_complex_single_plain(gs, oracle, n):
    A = gen.ginibre_complex(n, gen.stream(13, 100, n))
    S = gs.gschur(A, plain=True)
    _check(oracle, A, S)


@pytest.mark.parametrize("n", [3, 40, 160])
def test_real_single_plain(gs, oracle, n):
    A = gen.ginibre_real(n, gen.stream(14, 100, n))
    S = gs.gschur(A, plain=True)
    _check(oracle, A, S)


def test_workload1_n256(gs, oracle):
    """One 256 x 256 complex Ginibre matrix (the reference's CPU case)."""
    A = gen.workload1(seed=0)
    S = gs.gschur(A)
    _check(oracle, A, S)


def test_batched_complex(gs, oracle):
    mats = gen.workload2(seed=3, count=96)
    Ss, info = gs.gschur_batched(mats)
    assert np.all(info == 0)
    for A, S in zip(mats, Ss):
        _check(oracle, A, S)


def test_batched_real(gs, oracle):
    mats, kinds = gen.workload3(seed=4, count=96)
    Ss, info = gs.gschur_batched(mats)
    assert np.all(info == 0)
    for A, S in zip(mats, Ss):
        _check(oracle, A, S)


def test_batched_matches_single(gs):
    """A batch and per-matrix calls run the same kernel: identical bits."""
    mats = gen.workload2(seed=5, count=12)
    Ss, _ = gs.gschur_batched(mats)
    for A, S in zip(mats, Ss):
        S1 = gs.gschur(A)
        assert np.array_equal(S.T, S1.T) and np.array_equal(S.Z, S1.Z) and np.array_equal(S.values, S1.values)


def test_zlog_matches_direct(gs, oracle):
    """The Z log: Z updated after the QR sweeps from the logged reflectors
    and scalings (default), during the sweeps (GS_ZDIRECT), and with a log of
    one sweep, so that the QR kernel applies it to Z in HBM before nearly
    every sweep (GS_ZLOG_TINY).  T does not depend on how Z is updated, and
    the three forms apply the same operations to Z in the same order with
    the same rounding: all bitwise equal.  The log form also passes the
    Schur invariants."""
    eps = np.finfo(np.float64).eps
    mats = gen.workload2(seed=7, count=64)
    mats += [gen.toeplitz_grcar(40).astype(np.complex128), np.eye(12, dtype=np.complex128)]
    Sl, il = gs.gschur_batched(mats)
    Sd, idd = gs.gschur_batched(mats, zdirect=True)
    St, it = gs.gschur_batched(mats, zlog_tiny=True)
    assert np.all(il == 0) and np.all(idd == 0) and np.all(it == 0)
    for A, a, d, t in zip(mats, Sl, Sd, St):
        n = A.shape[0]
        assert np.array_equal(a.T, d.T) and np.array_equal(a.T, t.T)
        assert np.array_equal(a.Z, t.Z) and np.array_equal(a.Z, d.Z)
        checks.schur_invariants(A, a.T, a.Z, a.values, 20.0)
        assert checks.rel_backward(A, a.T, a.Z) <= 50 * n * eps


def test_dlog_matches_direct(gs):
    """The real Z log (sweeps' 3-element reflectors and the 2 x 2 blocks'
    rotations, applied to Z after the sweeps) against in-place updates
    (GS_ZDIRECT) and a one-sweep log that the QR kernel applies itself
    before nearly every sweep or rotation (GS_ZLOG_TINY): T, Z and the
    values bitwise equal; the log form also passes the real Schur checks."""
    eps = np.finfo(np.float64).eps
    mats, _ = gen.workload3(seed=8, count=48)
    rng = gen.stream(31, 100, 0)
    mats += [gen.haar_orthogonal(70, rng), gen.haar_orthogonal(100, rng), gen.toeplitz_grcar(90),
             gen.ginibre_real(128, rng)]
    Sl, il = gs.gschur_batched(mats)
    Sd, idd = gs.gschur_batched(mats, zdirect=True)
    St, it = gs.gschur_batched(mats, zlog_tiny=True)
    assert np.all(il == 0) and np.all(idd == 0) and np.all(it == 0)
    for A, a, d, t in zip(mats, Sl, Sd, St):
        n = A.shape[0]
        assert np.array_equal(a.T, d.T) and np.array_equal(a.T, t.T)
        assert np.array_equal(a.values, d.values) and np.array_equal(a.values, t.values)
        assert np.array_equal(a.Z, t.Z) and np.array_equal(a.Z, d.Z)
        checks.schur_invariants(A, a.T, a.Z, a.values, 20.0)
        assert checks.rel_backward(A, a.T, a.Z) <= 50 * n * eps


def test_windowed_matches_band_qr(gs):
    """The windowed QR kernels (H in HBM, chase window in LDS, chained updates
    outside the window) apply the same reflectors in the same order as the
    band-resident kernels.  The compiler may fuse multiply-adds differently
    in the two codes, and rounding-level differences can change a later
    deflation step, so the factorizations are compared as Schur forms: the
    same eigenvalues in the same diagonal order, and T, Z equal up to a
    diagonal unitary D (T_w = D^H T_b D, Z_w = Z_b D) for complex input; real
    2x2 standard-form blocks have further freedom, so real factorizations are
    compared by their eigenvalues (their invariants are checked elsewhere)."""
    eps = np.finfo(np.float64).eps
    for kind in ("complex", "real"):
        mats = gen.workload2(seed=6, count=48) if kind == "complex" else gen.workload3(seed=6, count=48)[0]
        Sw, iw = gs.gschur_batched(mats)
        Sb, ib = gs.gschur_batched(mats, qr_band=True)
        assert np.all(iw == 0) and np.all(ib == 0)
        nbit = nphase = 0
        for A, a, b in zip(mats, Sw, Sb):
            n = A.shape[0]
            anorm = np.linalg.norm(A)
            if np.array_equal(a.T, b.T) and np.array_equal(a.Z, b.Z):
                nbit += 1
                continue
            _, d = checks.match_eigenvalues(a.values, b.values)
            assert d.max() <= 1e3 * n * eps * anorm
            if kind == "complex" and np.max(np.abs(a.values - b.values)) <= 1e3 * n * eps * anorm:
                D = np.diag(np.diag(b.Z.conj().T @ a.Z))
                assert np.allclose(np.abs(np.diag(D)), 1.0, atol=1e-6)
                assert np.linalg.norm(a.Z - b.Z @ D) <= 1e-6 * n
                nphase += 1
        print(f"{kind}: bitwise equal {nbit}, equal up to diagonal phases {nphase}, of {len(mats)}", flush=True)
