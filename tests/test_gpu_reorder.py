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


# ------------------------------------------------ real Schur forms
# gordschur! for real T (rordschur.jl:3-466) against the oracle bit for bit
# (the kernel evaluates the oracle's operations in its order, without
# contraction), with the reference's real checkord (test/ordschur.jl:89-133).


@pytest.mark.parametrize("n", [1, 2, 5, 20, 64, 150])
def test_ordschur_real_bitwise(gs, oracle, n):
    import checks
    r = np.random.default_rng(70 + n)
    A = r.standard_normal((n, n))
    S = gs.gschur(A)
    T, Z = np.asarray(S.T), np.asarray(S.Z)
    sels = [np.zeros(n, dtype=bool), np.ones(n, dtype=bool)]
    for _ in range(3):
        sels.append(r.random(n) < 0.4)
    s_last = np.zeros(n, dtype=bool)
    s_last[-1] = True
    sels.append(s_last)
    for k, sel in enumerate(sels):
        print(f"case ordschur_d n={n} sel={k}", flush=True)
        S2 = gs.ordschur_(S, sel)
        T2, Z2, w2, info = oracle.ordschur_re(T, Z, sel)
        assert info == 0
        assert np.array_equal(S2.T, T2), (n, k)
        assert np.array_equal(S2.Z, Z2), (n, k)
        assert np.array_equal(S2.values, w2), (n, k)
        if n > 1:
            assert np.all(np.tril(S2.T, -2) == 0)
            assert checks.checkblocks(S2.T)
            assert np.linalg.norm(S2.Z @ S2.T @ S2.Z.T - A) / (n * np.linalg.norm(A) * EPS) < 10
            assert np.linalg.norm(np.eye(n) - S2.Z @ S2.Z.T) / (n * EPS) < 10
    # without Z: the same T
    S0 = gs.Schur(T, np.zeros((0, 0)), S.values)
    sel = sels[2]
    assert np.array_equal(gs.ordschur_(S0, sel).T, gs.ordschur_(S, sel).T)


def test_ordschur_real_checkord(gs, oracle):
    """test/ordschur.jl:89-133 on the GPU: two reals, a pair, a real and a
    pair, two pairs moved to the front of an n = 20 real Schur form."""
    import checks
    n = 20
    r = np.random.default_rng(48)
    A = r.random((n, n))
    S = gs.gschur(A)
    wv = np.asarray(S.values)
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
                if abs(iwi) > 1000 * EPS and wv[i - 1].imag == -iwi:
                    select[i] = True
                    select[i - 1] = True
                    ni -= 2
                    i -= 1
        if nr + ni > 0:
            continue
        S2 = gs.ordschur_(S, select)
        T2, Z2, w2 = np.asarray(S2.T), np.asarray(S2.Z), np.asarray(S2.values)
        assert np.all(np.tril(T2, -2) == 0)
        assert np.linalg.norm(Z2 @ T2 @ Z2.T - A) / (n * np.linalg.norm(A) * EPS) < 10
        assert np.linalg.norm(np.eye(n) - Z2 @ Z2.T) / (n * EPS) < 10
        assert checks.checkeigvals(T2, w2, 10)
        want = wv[select]
        wscale = np.abs(w2).max()
        for k in range(int(select.sum())):
            assert np.min(np.abs(want - w2[k])) / (EPS + wscale) < 10


def test_eigcond_real(gs, oracle):
    """eigvalscond / subspacesep of a real Schur factorization: with real
    eigenvalues (T triangular) against LAPACK xTRSEN's s and sep (the
    reference's checkecond, test/ordschur.jl:31-56); with complex pairs the
    reference's generic solvers read T as triangular, as the oracle does on
    the same T in complex form."""
    import scipy.linalg.lapack as lap
    r = np.random.default_rng(49)
    n, k = 24, 5
    T = np.triu(r.standard_normal((n, n))) + np.diag(np.arange(n) * 0.5)
    S = gs.Schur(T, np.eye(n), np.diag(T).astype(complex))
    s, sep = gs.eigcond(S, k)
    sel = np.zeros(n, dtype=np.int32)
    sel[:k] = 1
    out = lap.dtrsen(sel, T, np.eye(n), job="B", lwork=max(1, 2 * k * (n - k)), liwork=max(1, k * (n - k)))
    s_l, sep_l, info = out[5], out[6], out[-1]
    assert info == 0
    assert abs(s - s_l) <= 1e-8 * s_l and abs(sep - sep_l) <= 1e-8 * sep_l
    A = r.standard_normal((n, n))
    Sq = gs.gschur(A)
    Tq = np.asarray(Sq.T)
    s2, sep2 = gs.eigcond(Sq, k)
    assert s2 == oracle.eigvalscond(Tq.astype(complex), k)
    assert abs(sep2 - oracle.subspacesep(Tq.astype(complex), k)) <= 1e-12 * sep2
