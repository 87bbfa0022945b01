"""GPU parity of eigen! (pirates.jl:63-160) and its pieces: balancing
(balance.jl:33-199) and the balancer's back-transforms (balance.jl:201-261),
triangularize (triang.jl:9-43), _enormalize! (util.jl:572-592), and the
reciprocal eigenvalue condition numbers of the jce option, each against the
oracle (whose balancing is pinned bit for bit to LAPACK's xGEBAL on the
CPU), plus the reference's own balancing tests (test/balance.jl)."""
import numpy as np
import pytest

import checks
from genericschur import gen

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _bal_cases():
    r = np.random.default_rng(23)
    out = []
    for cplx in (False, True):
        for trial in range(8):
            n = (5, 12, 30, 77)[trial % 4]
            A = r.standard_normal((n, n)) * np.exp(3 * r.standard_normal((n, 1))) * np.exp(3 * r.standard_normal((1, n)))
            if trial >= 4:
                A[np.abs(A) < 1.0] = 0.0
            if cplx:
                A = A + 1j * A[::-1]
            out.append(A)
    return out


def test_balance_matches_oracle(gs, oracle):
    """balance! on the GPU against the oracle: the balanced matrix, ilo, ihi,
    trivial, the permutation record and the scale factors bit for bit (the
    scale decisions compare norms with 0.95 and powers of two, far from
    their rounding), for dense and sparse (permutable) input, real and
    complex, without scaling and without permuting too."""
    for A in _bal_cases():
        for scale, permute in ((True, True), (True, False), (False, True)):
            print(f"case balance n={A.shape[0]} cplx={np.iscomplexobj(A)} scale={scale} permute={permute}",
                  flush=True)
            Ab, B = gs.balance_(A, scale=scale, permute=permute)
            ab, bal, sp, d, info = oracle.balance(A, scale=scale, permute=permute)
            assert info == 0
            assert (B.ilo, B.ihi, B.trivial) == bal
            assert np.array_equal(Ab, ab)
            assert np.array_equal(B.D, d)
            assert np.array_equal(B.sp, sp.astype(np.int32))


def test_balance_sanity_block_structure(gs):
    """test/balance.jl:74-106 on the GPU: a permuted block-triangular matrix
    balances to upper triangular outside the ilo:ihi block."""
    r = np.random.default_rng(4)
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
                C, B = gs.balance_(P.T @ A @ P)
                assert B.ilo <= B.ihi
                for i in range(1, n + 1):
                    for j in range(1, i):
                        if i < B.ilo or j > B.ihi:
                            assert C[i - 1, j - 1] == 0


@pytest.mark.parametrize("cplx", [False, True])
def test_balance_unbal_classic(gs, cplx):
    """test/balance.jl:58-72: balancing makes the eigenvalues of the badly
    balanced classic matrix far more accurate (ru > 100 rb), and ldiv!(B, Vl)
    turns the balanced matrix's left eigenvectors into A's."""
    m = np.round(-2 * np.log2(EPS) / 5)
    x = 2.0 ** m
    A = np.array([[1, 0, x ** -2], [1, 1, x ** -1], [x ** 2, x, 1]], dtype=np.complex128 if cplx else np.float64)
    rt2 = np.sqrt(2.0)
    lam = np.array([1 - rt2 + 1 / (4 * x), 1 - 1 / (2 * x), 1 + rt2 + 1 / (4 * x)])
    eu = gs.eigen_(A, permute=False, scale=False)
    eb = gs.eigen_(A)
    ru = np.abs(eu.values.real - lam).max()
    rb = np.abs(eb.values.real - lam).max()
    print(f"unbalanced error {ru:.3e}, balanced {rb:.3e}", flush=True)
    assert ru > 100 * rb
    if cplx:
        Ab, B = gs.balance_(A)
        S = gs.gschur(Ab)
        Vl = gs.ldiv_(B, gs.geigvecs(S, left=True))
        assert np.allclose(A.conj().T @ Vl, Vl @ np.diag(S.values.conj()), rtol=1e-8, atol=1e-8)


def test_balance_nan_raises(gs, oracle):
    """balance.jl:160-162: a NaN met while scaling is an error.  (A NaN
    entry gives a NaN norm, which the reference's sweep loop never settles
    on; both the kernel and the oracle report it as this error.)"""
    for A in (np.array([[1.0, 2.0, 0.5], [np.nan, 3.0, 1.0], [4.0, 1.0, 2.0]]),
              np.array([[1.0, 2.0j], [3.0, np.nan]])):
        assert oracle.balance(A)[4] == -3
        with pytest.raises(RuntimeError, match="NaN"):
            gs.balance_(A)


@pytest.mark.parametrize("n", [2, 9, 40, 130, 300])
def test_triangularize_matches_oracle(gs, oracle, n):
    """triangularize on the GPU against the oracle on the same (T, Z): equal
    to a few ulps (the rotation's hypot and square roots may round
    differently); the result is a complex Schur factorization of A."""
    A = gen.ginibre_real(n, gen.stream(41, 90, n))
    S = gs.gschur(A)
    C = gs.triangularize(S)
    Tc, Zc = oracle.triangularize(S.T, S.Z)
    scale = np.abs(S.T).max()
    assert np.abs(C.T - Tc).max() <= 64 * EPS * scale
    assert np.abs(C.Z - Zc).max() <= 64 * EPS
    assert np.all(np.tril(C.T, -1) == 0)
    assert np.array_equal(C.values, np.diag(C.T))
    assert checks.rel_backward(A.astype(complex), C.T, C.Z) <= 50 * n * EPS
    assert checks.rel_orth(C.Z) <= 50 * n * EPS
    _, err = checks.match_eigenvalues(C.values, S.values)
    assert err.max() <= 64 * EPS * scale


def test_enormalize_matches_oracle(gs, oracle):
    """_enormalize! on the GPU against the oracle: unit 2-norm columns with the
    first largest entry real and positive, ties to the first index."""
    r = np.random.default_rng(8)
    for n, m in ((3, 3), (70, 70), (200, 200)):
        V = r.standard_normal((n, m)) + 1j * r.standard_normal((n, m))
        V[:, 0] = V[0, 0]  # all entries tied: the first one is rotated real
        got = gs.enormalize_(V)
        ref = oracle.enormalize(V)
        assert np.abs(got - ref).max() <= 16 * EPS
        i0 = np.argmax(np.abs(got) ** 2, axis=0)
        assert np.all(got[i0, np.arange(m)].imag == 0) and np.all(got[i0, np.arange(m)].real > 0)
        assert got[0, 0].imag == 0


@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [5, 33, 100, 200])
def test_eigen_matches_oracle(gs, oracle, cplx, n):
    """eigen!(A; jvl, jce) on the GPU against the oracle's composition of the
    same steps: sorted eigenvalues within 10 n eps ||A|| cond, right and left
    vectors (unit, phase fixed by _enormalize!) entry by entry, and rconde;
    with and without balancing, on a badly scaled matrix."""
    rng = gen.stream(7, 91, n + int(cplx))
    A = rng.standard_normal((n, n)) * np.exp(2 * rng.standard_normal((n, 1))) * np.exp(-2 * rng.standard_normal((1, n)))
    if cplx:
        A = A + 1j * rng.standard_normal((n, n))
    for bal in (True, False):
        print(f"case eigen n={n} cplx={cplx} balance={bal}", flush=True)
        E = gs.eigen_(A, permute=bal, scale=bal, jvl=True, jce=True)
        w, V, VL, rc = oracle.eigen(A, permute=bal, scale=bal, jvl=True, jce=True)
        cond = 1.0 / rc
        checks.eigen_parity(A, E.values, w, cond)
        same = np.abs(E.values - w) <= 1e3 * n * EPS * np.linalg.norm(A, 2) * cond
        for got, ref in ((E.vectors, V), (E.vectorsl, VL)):
            assert np.all(np.abs(np.linalg.norm(got, axis=0) - 1) < 1e-12)
            # vectors of eigenvalues in the same sorted slot, weighted by the
            # eigenvalue's sensitivity
            d = np.abs(got - ref).max(axis=0)
            assert np.all(d[same] <= 1e-6 * np.maximum(1.0, cond[same]))
        assert np.abs(E.rconde - rc).max() <= 1e-6
        # A x = lambda x for the returned right vectors
        r = np.linalg.norm(A @ E.vectors - E.vectors * E.values, axis=0)
        assert np.all(r <= 1e3 * n * EPS * np.linalg.norm(A, 2) * np.maximum(1.0, cond))


@pytest.mark.parametrize("cplx", [False, True])
def test_eigen_large_order(gs, cplx):
    """eigen!(A; jvl, jce) above order 1024, where every piece takes its
    whole-GPU path (the large real Schur path or the complex windowed path,
    the order > 1024 eigenvector solves, balancing and its back-transforms):
    unit right and left vectors, A x = lambda x and y^H A = lambda y^H within
    the conditions the call returns, and the eigenvalues against LAPACK's."""
    n = 1100
    rng = gen.stream(8, 92, n + int(cplx))
    A = rng.standard_normal((n, n))
    if cplx:
        A = A + 1j * rng.standard_normal((n, n))
    print(f"case eigen n={n} cplx={cplx}", flush=True)
    E = gs.eigen_(A, jvl=True, jce=True)
    cond = 1.0 / E.rconde
    nrm = np.linalg.norm(A)  # Frobenius: bounds ||A||_2
    for V in (E.vectors, E.vectorsl):
        assert np.all(np.abs(np.linalg.norm(V, axis=0) - 1) < 1e-12)
    r = np.linalg.norm(A @ E.vectors - E.vectors * E.values, axis=0)
    assert np.all(r <= 1e3 * n * EPS * nrm * np.maximum(1.0, cond))
    rl = np.linalg.norm(E.vectorsl.conj().T @ A - E.values[:, None] * E.vectorsl.conj().T, axis=1)
    assert np.all(rl <= 1e3 * n * EPS * nrm * np.maximum(1.0, cond))
    perm, d = checks.match_eigenvalues(E.values, np.linalg.eigvals(A))
    assert np.all(d <= 1e3 * n * EPS * nrm * np.maximum(1.0, cond))
