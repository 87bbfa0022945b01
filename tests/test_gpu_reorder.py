"""GPU parity of the reordering and conditioning of complex Schur forms:
gordschur! (ordschur.jl:23-73) against the oracle (and so, through the
oracle's pinning, LAPACK's xTRSEN), with the reference's own checks
(test/ordschur.jl:1-29: triangular, backward error, unitary Z, the selected
eigenvalues first); eigvalscond / subspacesep (ordschur.jl:85-131) against
the oracle (test/ordschur.jl:31-56 compares them with xTRSEN's s and sep)."""
import numpy as np
import pytest

from genericschur import gen

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


def _selects(n, r):
    s0 = np.ones(n, dtype=bool)
    s0[: n // 2] = False
    s1 = np.zeros(n, dtype=bool)
    s1[r.choice(n, size=max(1, n // 4), replace=False)] = True
    s2 = np.zeros(n, dtype=bool)
    s2[-1] = True
    return [s0, s1, s2]


@pytest.mark.parametrize("n", [2, 7, 32, 100, 200])
def test_ordschur_matches_oracle(gs, oracle, n):
    r = np.random.default_rng(50 + n)
    A = gen.ginibre_complex(n, gen.stream(9, 50, n))
    S = gs.gschur(A)
    for sel in _selects(n, r):
        print(f"case ordschur n={n} k={int(sel.sum())}", flush=True)
        S2 = gs.ordschur_(S, sel)
        T2, Z2, w2 = oracle.ordschur(S.T, S.Z, sel)
        assert np.all(np.tril(S2.T, -1) == 0)
        assert np.linalg.norm(S2.Z @ S2.T @ S2.Z.conj().T - A) / (n * np.linalg.norm(A) * EPS) < 10
        assert np.linalg.norm(np.eye(n) - S2.Z @ S2.Z.conj().T) / (n * EPS) < 10
        k = int(sel.sum())
        want = S.values[sel]
        scale = np.abs(S.values).max()
        assert np.all(np.abs(want - S2.values[:k]) <= 10 * n * EPS * scale)
        assert np.array_equal(S2.values, np.diag(S2.T))
        # the same swaps as the oracle on the same (T, Z): equal to rounding
        assert np.abs(S2.values - w2).max() <= 100 * n * EPS * scale
        assert np.abs(S2.T - T2).max() <= 1e3 * n * EPS * scale
        assert np.abs(S2.Z - Z2).max() <= 1e3 * n * EPS


def test_ordschur_without_z(gs, oracle):
    n = 24
    A = gen.ginibre_complex(n, gen.stream(9, 51, n))
    S = gs.gschur(A)
    sel = np.zeros(n, dtype=bool)
    sel[[3, 10, 20]] = True
    S0 = gs.Schur(S.T, np.zeros((0, 0), complex), S.values)
    S1 = gs.ordschur_(S0, sel)
    S2 = gs.ordschur_(S, sel)
    assert np.array_equal(S1.T, S2.T)


@pytest.mark.parametrize("n,nsub", [(2, 1), (10, 3), (32, 8), (64, 40)])
def test_eigcond_matches_oracle(gs, oracle, n, nsub):
    A = gen.ginibre_complex(n, gen.stream(9, 52, n))
    S = gs.gschur(A)
    print(f"case eigcond n={n} nsub={nsub}", flush=True)
    s, sep = gs.eigcond(S, nsub)
    so = oracle.eigvalscond(S.T, nsub)
    sepo = oracle.subspacesep(S.T, nsub)
    assert abs(s - so) <= 1e-10 * so
    assert abs(sep - sepo) <= 1e-8 * sepo
    assert 0 < s <= 1


def test_eigcond_rejects_bad_nsub(gs):
    S = gs.gschur(gen.ginibre_complex(6, gen.stream(9, 53, 6)))
    for k in (0, 6):
        with pytest.raises(gs.GSError):
            gs.eigcond(S, k)
