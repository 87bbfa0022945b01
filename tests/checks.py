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
        na = fnorm(A)
        d = 0.0 if na == 0 and fnorm(np.triu(T, k)) == 0 else fnorm(np.triu(T, k)) / (n * na * ULP)
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


def eigen_clusters(ref_w, bound):
    """Groups of reference eigenvalues whose error discs |z - ref_w[i]| <=
    bound[i] overlap (transitive closure).  Inside such a cluster the
    individual condition numbers do not bound the error of each member (a
    nearly multiple eigenvalue splits arbitrarily under perturbation); only
    the cluster as a whole is determined."""
    n = len(ref_w)
    parent = list(range(n))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i

    d = np.abs(ref_w[:, None] - ref_w[None, :])
    ov = d <= bound[:, None] + bound[None, :]
    for i in range(n):
        for j in np.nonzero(ov[i, i + 1:])[0] + i + 1:
            parent[find(i)] = find(j)
    return np.array([find(i) for i in range(n)])


def eigen_parity(A, w, ref_w, cond, factor=10.0):
    """|lambda - lambda_ref| <= factor * n * eps * ||A||_2 * cond(lambda).

    The computed values are matched to the reference values by an optimal
    assignment.  A value whose reference belongs to a cluster of overlapping
    error discs (eigen_clusters) passes if it lies in the disc of any member
    of that cluster; the number of such values is returned with the worst
    ratio err/bound over the isolated ones."""
    n = A.shape[0]
    if n == 0:
        return 0.0
    ref_w = np.asarray(ref_w)
    perm, err = match_eigenvalues(w, ref_w)
    bound = factor * n * ULP * np.linalg.norm(A, 2) * np.asarray(cond, dtype=np.float64)
    bound = np.where(np.isnan(bound), np.inf, bound)
    ratio = err / np.maximum(bound, TINY)
    bad = np.nonzero(ratio > 1.0)[0]
    if len(bad):
        cl = eigen_clusters(ref_w, bound)
        wp = np.asarray(w)[perm]
        for i in bad:
            members = np.nonzero(cl == cl[i])[0]
            ok = len(members) > 1 and np.any(np.abs(wp[i] - ref_w[members]) <= bound[members])
            assert ok, (f"eigenvalue mismatch: err/bound = {ratio[i]} at {ref_w[i]} "
                        f"(computed {wp[i]}, cluster of {len(members)})")
    iso = np.setdiff1d(np.arange(n), bad)
    return float(np.max(ratio[iso])) if len(iso) else 0.0


def well_conditioned(A, ref_w, cond, margin=1e6):
    """Whether every eigenvalue is simple and separated from the others by
    more than margin times its error bound n eps ||A||_2 cond: the Schur form
    is then determined up to phases to about 1/margin."""
    n = A.shape[0]
    if n < 2:
        return True
    b = n * ULP * np.linalg.norm(A, 2) * np.asarray(cond)
    if not np.all(np.isfinite(b)):
        return False
    d = np.abs(ref_w[:, None] - ref_w[None, :])
    np.fill_diagonal(d, np.inf)
    return bool(np.all(d > margin * (b[:, None] + b[None, :])))


def diag_blocks(T):
    """Diagonal blocks of a (quasi-)triangular T as (start, size) pairs."""
    n = T.shape[0]
    out = []
    j = 0
    while j < n:
        if not np.iscomplexobj(T) and j + 1 < n and T[j + 1, j] != 0:
            out.append((j, 2))
            j += 2
        else:
            out.append((j, 1))
            j += 1
    return out


def schur_form_match(A, T, Z, w, To, Zo, wo, cond, ztol=1e-6):
    """Compares a Schur factorization (T, Z, w) of A with the oracle's
    (To, Zo, wo) up to the factorization's own freedom.

    Returns one of
      "bitwise"  T, Z and the values bit for bit equal;
      "phases"   the same eigenvalue in every diagonal slot (within the
                 eigenvalue bound 10 n eps ||A||_2 cond) and T = D^H To D,
                 Z = Zo D for a block-diagonal unitary D over To's diagonal
                 blocks (complex: unit phases; real: signs of 1 x 1 blocks,
                 orthogonal 2 x 2 blocks), within ztol relative;
      "order"    the same slot order, T or Z outside ztol (e.g. a multiple
                 eigenvalue, whose invariant subspace has no unique basis);
      "differs"  another eigenvalue order.
    """
    n = A.shape[0]
    if np.array_equal(T, To) and np.array_equal(Z, Zo) and np.array_equal(w, wo):
        return "bitwise"
    if n == 0:
        return "bitwise"
    bound = 10.0 * n * ULP * np.linalg.norm(A, 2) * np.asarray(cond) + TINY
    if np.any(np.abs(np.asarray(w) - np.asarray(wo)) > bound):
        return "differs"
    M = Zo.conj().T @ Z
    D = np.zeros_like(M)
    for j, k in diag_blocks(To):
        D[j:j + k, j:j + k] = M[j:j + k, j:j + k]
    na = fnorm(A)
    zerr = fnorm(Z - Zo @ D)
    terr = fnorm(T - D.conj().T @ To @ D)
    if zerr <= ztol * np.sqrt(n) and terr <= ztol * max(na, TINY):
        return "phases"
    return "order"
