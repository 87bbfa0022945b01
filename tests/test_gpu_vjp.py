"""GPU eigenvalue VJP (the backward of the eigenvalue rrule).

Loss of the differentiation workload: L = sum_i w_i Re(lambda_i) + v_i Im(lambda_i)^2,
so gbar_i = w_i + 2 i v_i Im(lambda_i).  Checked against
  * the exact invariant w = 1, v = 0: L = tr A, Abar = I;
  * the same formula evaluated with the oracle's eigenvectors;
  * central finite differences, h = 1e-6, random directions, eigenvalue slots
    matched by nearest neighbour (separations >= 1e-3), relative error 1e-6;
  * the formula evaluated with mpmath eigenvectors at 256 bits.
"""
import numpy as np
import pytest

from genericschur import gen
from genericschur.autodiff import eig_vjp_batched

pytestmark = pytest.mark.gpu


def _loss(vals, w, v):
    return float(np.sum(w * vals.real + v * vals.imag ** 2))


def _gbar(vals, w, v):
    return w + 2j * v * vals.imag


def _forward(gs, mats):
    Ss, info = gs.gschur_batched(mats)
    assert np.all(info == 0)
    return Ss


def _vjp(gs, mats, Ss, gbars):
    ns, off, Tf = gs.pack([S.T for S in Ss], np.float64)
    _, _, Zf = gs.pack([S.Z for S in Ss], np.float64)
    g = np.concatenate(gbars)
    Ab, info = eig_vjp_batched(ns, off, Tf, Zf, g)
    assert np.all(info == 0)
    return gs.unpack(ns, off, Ab)


def test_trace_invariant(gs):
    mats, _, _ = gen.workload5(seed=1, count=6)
    Ss = _forward(gs, mats)
    Abs = _vjp(gs, mats, Ss, [np.ones(A.shape[0], dtype=complex) for A in mats])
    for A, Ab in zip(mats, Abs):
        assert np.abs(Ab - np.eye(A.shape[0])).max() <= 1e-11


def test_matches_oracle_formula(gs, oracle):
    mats, ws, vs = gen.workload5(seed=2, count=6)
    Ss = _forward(gs, mats)
    gbars = [_gbar(S.values, w, v) for S, w, v in zip(Ss, ws, vs)]
    Abs = _vjp(gs, mats, Ss, gbars)
    for A, S, g, Ab in zip(mats, Ss, gbars, Abs):
        X = oracle.eigvecs(S.T, S.Z, left=False)
        Y = oracle.eigvecs(S.T, S.Z, left=True)
        d = np.sum(Y.conj() * X, axis=0)
        ref = np.real((Y * (g / np.conj(d))) @ X.conj().T)
        assert np.abs(Ab - ref).max() <= 1e-10 * max(1.0, np.abs(ref).max())


def test_central_differences(gs):
    mats, ws, vs = gen.workload5(seed=3, count=4)
    Ss = _forward(gs, mats)
    gbars = [_gbar(S.values, w, v) for S, w, v in zip(Ss, ws, vs)]
    Abs = _vjp(gs, mats, Ss, gbars)
    h = 1e-6
    rng = np.random.default_rng(5)
    worst = 0.0
    for A, S, w, v, Ab in zip(mats, Ss, ws, vs, Abs):
        n = A.shape[0]
        base = S.values
        dirs = [rng.standard_normal((n, n)) for _ in range(64)]  # the workload's 64 directions per matrix
        pert = []
        for E in dirs:
            pert += [A + h * E, A - h * E]
        Sp = _forward(gs, pert)

        def L(vals):
            # slot of each base eigenvalue in the perturbed spectrum
            idx = [int(np.argmin(np.abs(vals - b))) for b in base]
            assert len(set(idx)) == n
            return _loss(vals[idx], w, v)

        for k, E in enumerate(dirs):
            fd = (L(Sp[2 * k].values) - L(Sp[2 * k + 1].values)) / (2 * h)
            an = float(np.sum(Ab * E))
            worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
    print("worst FD relative error", worst)
    assert worst <= 1e-6


def test_mp256_gradient(gs):
    import mpmath as mp
    mp.mp.prec = 256
    rng = np.random.default_rng(7)
    n = 8
    A = rng.standard_normal((n, n))
    w = rng.standard_normal(n)
    v = rng.standard_normal(n)
    S = _forward(gs, [A])[0]
    g = _gbar(S.values, w, v)
    Ab = _vjp(gs, [A], [S], [g])[0]
    M = mp.matrix(A.tolist())
    E, L, R = mp.eig(M, left=True, right=True)
    ev = np.array([complex(e) for e in E])
    ref = mp.matrix(n, n)
    for k in range(n):
        slot = int(np.argmin(np.abs(S.values - ev[k])))
        gk = mp.mpc(g[slot].real, g[slot].imag)
        x = R[:, k]
        y = L[k, :].T  # mpmath returns left vectors as rows: y^H A = lambda y^H
        y = mp.matrix([mp.conj(y[i]) for i in range(n)])
        d = sum(mp.conj(y[i]) * x[i] for i in range(n))
        c = gk / mp.conj(d)
        for r in range(n):
            for s in range(n):
                ref[r, s] += c * y[r] * mp.conj(x[s])
    R64 = np.array([[float(mp.re(ref[r, s])) for s in range(n)] for r in range(n)])
    assert np.abs(Ab - R64).max() <= 1e-10 * max(1.0, np.abs(R64).max())


def test_autograd_function(gs):
    import torch
    from genericschur.autodiff import make_schur_eigvals
    SchurEigvals = make_schur_eigvals(torch)
    mats, ws, vs = gen.workload5(seed=4, count=3)
    ns, off, flat = gs.pack(mats, np.float64)
    dev = torch.device("cuda:0")
    A = torch.tensor(flat, device=dev, requires_grad=True)
    lam = SchurEigvals.apply(A, ns, off)
    wt = torch.tensor(np.concatenate(ws), device=dev)
    vt = torch.tensor(np.concatenate(vs), device=dev)
    loss = (wt * lam.real + vt * lam.imag ** 2).sum()
    loss.backward()
    Ss = _forward(gs, mats)
    gbars = [_gbar(S.values, w, v) for S, w, v in zip(Ss, ws, vs)]
    ref = np.concatenate([a.reshape(-1, order="F") for a in _vjp(gs, mats, Ss, gbars)])
    assert np.abs(A.grad.cpu().numpy() - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def _sep_ginibre(n, seed):
    k = 0
    while True:
        A = gen.ginibre_real(n, gen.stream(seed, 95, n + (k << 16)))
        if gen.min_separation(A) >= 1e-3:
            return A
        k += 1


def test_large_orders_gemm_assembly(gs, oracle):
    """Orders above the LDS-resident assembly (n > 70) form Abar as one FP64
    MFMA GEMM of the split factors: a mixed batch against the trace invariant
    and the oracle-eigenvector formula, and n = 100 against central finite
    differences."""
    mats = [_sep_ginibre(n, 9) for n in (40, 100, 250, 71, 64)]
    Ss = [gs.gschur(A) for A in mats]  # n = 250 takes the large single-matrix path
    Abs = _vjp(gs, mats, Ss, [np.ones(A.shape[0], dtype=complex) for A in mats])
    for A, Ab in zip(mats, Abs):
        print(f"case vjp trace n={A.shape[0]}", flush=True)
        assert np.abs(Ab - np.eye(A.shape[0])).max() <= 1e-9
    rng = np.random.default_rng(11)
    gbars = [rng.standard_normal(A.shape[0]) + 1j * rng.standard_normal(A.shape[0]) for A in mats]
    Abs = _vjp(gs, mats, Ss, gbars)
    for A, S, g, Ab in zip(mats, Ss, gbars, Abs):
        X = oracle.eigvecs(S.T, S.Z, left=False)
        Y = oracle.eigvecs(S.T, S.Z, left=True)
        d = np.sum(Y.conj() * X, axis=0)
        ref = np.real((Y * (g / np.conj(d))) @ X.conj().T)
        print(f"case vjp formula n={A.shape[0]}", flush=True)
        assert np.abs(Ab - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())
    A = mats[1]
    n = A.shape[0]
    w, v = rng.standard_normal(n), rng.standard_normal(n)
    S = Ss[1]
    Ab = _vjp(gs, [A], [S], [_gbar(S.values, w, v)])[0]
    h = 1e-6
    dirs = [rng.standard_normal((n, n)) for _ in range(8)]
    pert = []
    for E in dirs:
        pert += [A + h * E, A - h * E]
    Sp = _forward(gs, pert)
    worst = 0.0
    for k, E in enumerate(dirs):
        def L(vals):
            idx = [int(np.argmin(np.abs(vals - b))) for b in S.values]
            assert len(set(idx)) == n
            return _loss(vals[idx], w, v)
        fd = (L(Sp[2 * k].values) - L(Sp[2 * k + 1].values)) / (2 * h)
        an = float(np.sum(Ab * E))
        worst = max(worst, abs(fd - an) / max(1.0, abs(an)))
    print("n=100 worst FD relative error", worst, flush=True)
    assert worst <= 1e-6


def test_order_above_1024(gs, oracle):
    """An order above 1024: the right and left eigenvectors recomputed on the
    whole-GPU path, Abar as the FP64 GEMM of the split factors; the trace
    invariant (gbar = 1 gives the identity) and the oracle-eigenvector
    formula."""
    n = 1100
    A = gen.ginibre_real(n, gen.stream(12, 1, n))
    S = gs.gschur(A)
    print(f"case vjp n={n}", flush=True)
    Ab = _vjp(gs, [A], [S], [np.ones(n, dtype=complex)])[0]
    assert np.abs(Ab - np.eye(n)).max() <= 1e-8
    rng = np.random.default_rng(13)
    g = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Ab = _vjp(gs, [A], [S], [g])[0]
    X = oracle.eigvecs(S.T, S.Z, left=False)
    Y = oracle.eigvecs(S.T, S.Z, left=True)
    d = np.sum(Y.conj() * X, axis=0)
    ref = np.real((Y * (g / np.conj(d))) @ X.conj().T)
    assert np.abs(Ab - ref).max() <= 1e-7 * max(1.0, np.abs(ref).max())
