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
