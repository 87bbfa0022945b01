"""Invariant checks restated from the reference's test suite.

schurtest (test/complex.jl:1-34, test/real.jl:24-74), hesstest
(test/complex.jl:36-61, test/real.jl:76-99), vectest (test/testfuncs.jl:5-54),
checkblocks (test/real.jl:3-22), checkeigvals (test/testfuncs.jl:57-86), plus
eigenvalue matching against a reference set weighted by condition numbers.
All residuals are in the reference's units: ||.||_F / (n ||A|| ulp).
"""
import numpy as np
from scipy.optimize import linear_sum_assignment

ULP = np.finfo(np.float64).eps
TINY = np.finfo(np.float64).tiny


def fnorm(M):
    """Frobenius / 2-norm of a vector without over- or underflow (numpy's
    norm squares entries directly; Julia's norm rescales)."""
    a = np.abs(np.asarray(M))
    if a.size == 0:
        return 0.0
    m = a.max()
    if m == 0 or not np.isfinite(m):
        return float(m)
    return float(m * np.sqrt(np.sum((a / m) ** 2)))


def backward_error(A, T, Z):
    n = A.shape[0]
    na = fnorm(A)
    if n == 0 or na == 0:
        return 0.0
    return fnorm(A - Z @ T @ Z.conj().T) / (n * na * ULP)


def orth_error(Z):
    n = Z.shape[0]
    if n == 0:
        return 0.0
    return fnorm(np.eye(n) - Z @ Z.conj().T) / (n * ULP)


def rel_backward(A, T, Z):
    """||A - Z T Z^H||_F / ||A||_F (the north-star acceptance metric)."""
    na = fnorm(A)
    return 0.0 if na == 0 else fnorm(A - Z @ T @ Z.conj().T) / na


def rel_orth(Z):
    return fnorm(Z.conj().T @ Z - np.eye(Z.shape[0]))


def csort(v):
    v = np.asarray(v)
    return v[np.lexsort((v.imag, v.real))]


def checkblocks(T):
    """test/real.jl:3-22: standard-form 2x2 blocks."""
    n = T.shape[0]
    for j in range(n - 1):
        if T[j + 1, j] != 0:
            if not (T[j, j] == T[j + 1, j + 1]):
                return False
            if not (T[j, j + 1] != 0):
                return False
            if not (np.sign(T[j, j + 1]) * np.sign(T[j + 1, j]) < 0):
                return False
    return True


def checkeigvals(T, w, tol):
    """test/testfuncs.jl:57-86."""
    n = T.shape[0]
    ok = True
    for j in range(n):
        ok &= (T[j, j] == w[j].real)
    if n > 1:
        if T[1, 0] == 0:
            ok &= (w[0].imag == 0)
        if T[n - 1, n - 2] == 0:
            ok &= (w[n - 1].imag == 0)
    for j in range(n - 1):
        if T[j + 1, j] != 0:
            t = np.sqrt(abs(T[j + 1, j])) * np.sqrt(abs(T[j, j + 1]))
            cmp = max(ULP * t, TINY)
            ok &= (abs(w[j].imag - t) / cmp < tol)
            ok &= (abs(w[j + 1].imag + t) / cmp < tol)
        elif j > 0 and T[j + 1, j] == 0 and T[j, j - 1] == 0:
            ok &= (w[j].imag == 0)
    return bool(ok)


def schur_invariants(A, T, Z, w, tol, normal=False):
    """Returns a dict of the schurtest checks; raises AssertionError on failure."""
    n = A.shape[0]
    res = {}
    cplx = np.iscomplexobj(T)
    if cplx:
        assert np.all(np.tril(T, -1) == 0), "T not upper triangular"
        assert np.all(csort(w) == csort(np.diag(T))), "values differ from diag(T)"
    else:
        assert np.all(np.tril(T, -2) == 0), "T not quasi-triangular"
        assert checkblocks(T), "2x2 blocks not in standard form"
        assert checkeigvals(T, w, tol), "values inconsistent with the 2x2 blocks"
    res["backward"] = backward_error(A, T, Z)
    res["orth"] = orth_error(Z)
    assert res["backward"] < tol, f"backward error {res['backward']} >= {tol}"
    assert res["orth"] < tol, f"orthogonality error {res['orth']} >= {tol}"
    if normal:
        k = 1 if cplx else 2
        d = fnorm(np.triu(T, k)) / (n * fnorm(A) * ULP)
        assert d < tol, f"normal matrix not diagonalised: {d}"
    return res


def vectest(A, w, VR, VL, vtol, normal=False):
    """test/testfuncs.jl:5-54."""
    n = A.shape[0]
    na = fnorm(A)
    r = fnorm(A @ VR - VR @ np.diag(w)) / (n * na * ULP)
    l = fnorm(A.conj().T @ VL - VL @ np.diag(w.conj())) / (n * na * ULP)
    assert r < vtol, f"right eigenvector residual {r}"
    assert l < vtol, f"left eigenvector residual {l}"
    if normal:
        votol = vtol if np.iscomplexobj(A) else 5 * vtol
        assert fnorm(VR @ VR.conj().T - np.eye(n)) / (n * ULP) < votol
        assert fnorm(VL @ VL.conj().T - np.eye(n)) / (n * ULP) < votol
    return r, l


def match_eigenvalues(w, ref):
    """Optimal assignment of w to ref; returns (perm, |w[perm] - ref|)."""
    w = np.asarray(w)
    ref = np.asarray(ref)
    C = np.abs(w[:, None] - ref[None, :])
    rows, cols = linear_sum_assignment(C)
    perm = np.empty(len(ref), dtype=int)
    perm[cols] = rows
    return perm, np.abs(w[perm] - ref)


def eig_conditions(VR, VL):
    """cond(lambda_k) = ||x|| ||y|| / |y^H x| for unit-norm vectors."""
    d = np.abs(np.sum(VL.conj() * VR, axis=0))
    nr = np.linalg.norm(VR, axis=0) * np.linalg.norm(VL, axis=0)
    with np.errstate(divide="ignore"):
        return nr / d


def eigen_parity(A, w, ref_w, cond, factor=10.0):
    """|lambda - lambda_ref| <= factor * n * eps * ||A||_2 * cond(lambda)."""
    n = A.shape[0]
    if n == 0:
        return 0.0
    perm, err = match_eigenvalues(w, ref_w)
    bound = factor * n * ULP * np.linalg.norm(A, 2) * cond
    worst = float(np.max(err / np.maximum(bound, TINY)))
    assert worst <= 1.0, f"eigenvalue mismatch: worst err/bound = {worst}"
    return worst
