"""Test matrices and error measures of the reference's symmetric tridiagonal
tests (test/symtridiag.jl), restated in numpy.

  * wilkinson, clement, legendre, laguerre, hermite: symtridiag.jl:11-44
  * class1(n, itype): st_testmat1 (symtridiag.jl:47-98), the LAPACK-style
    spectra 1..9 placed in Q' diag(λ) Q and reduced to tridiagonal form
  * class2(n, itype): st_testmat2 (symtridiag.jl:103-142), special
    tridiagonal matrices; type 8 glues floatmin/ulp-, 1- and floatmax*ulp-
    scaled Wilkinson matrices to exercise the scaling (gluemats, :146-155)
  * errors(D, E, lam, V): runtest (symtridiag.jl:195-222): the decomposition
    and orthogonality errors in units of eps * n, both required < 50.
"""
import numpy as np

EPS = np.finfo(np.float64).eps
TINY = np.finfo(np.float64).tiny
HUGE = np.finfo(np.float64).max


def wilkinson(n):
    if n % 2 == 1:
        nh = (n - 1) // 2
        D = np.concatenate([np.arange(nh, 0, -1), [0], np.arange(1, nh + 1)]).astype(float)
    else:
        D = np.abs(np.arange(-(n - 1), n, 2, dtype=float) / 2)
    return D, np.ones(n - 1)


def clement(n):
    return np.zeros(n), np.array([np.sqrt(float(i * (n + 1 - i))) for i in range(1, n)])


def legendre(n):
    return np.zeros(n), np.array([i / np.sqrt(float((2 * i - 1) * (2 * i) + 1)) for i in range(2, n + 1)])


def laguerre(n):
    return np.arange(3, 2 * n + 2, 2, dtype=float), np.arange(2, n + 1, dtype=float)


def hermite(n):
    return np.zeros(n), np.sqrt(np.arange(1, n, dtype=float))


def spectrum1(n, itype, rng):
    """The λ of st_testmat1 (symtridiag.jl:50-92)."""
    ulp = EPS
    kpar = 1 / ulp
    lam = np.zeros(n)
    j = np.arange(1, n + 1, dtype=float)
    if itype == 1:
        lam = 1 / j
    elif itype == 2:
        lam[:] = 1
        lam[-1] = 1 / n
    elif itype == 3:
        lam = 1 / kpar ** ((j - 1) / (n - 1))
    elif itype == 4:
        lam = 1 - ((j - 1) / (n - 1)) * (1 - 1 / kpar)
    elif itype == 5:
        lam = np.exp(-np.log(n) * rng.random(n))
    elif itype == 6:
        lam = rng.standard_normal(n)
    elif itype == 7:
        lam = np.arange(1, n + 1, dtype=float) * ulp
        lam[-1] = 1
    elif itype == 8:
        lam = 1 + np.sqrt(ulp) * j
        lam[0] = ulp
        lam[-1] = 2
    elif itype == 9:
        lam = 1 + 100 * ulp * (j - 1)
    return lam


def class1(n, itype, rng, tridiag):
    """st_testmat1: Q' diag(λ) Q for a random orthogonal Q, reduced to a
    symmetric tridiagonal (d, e) by `tridiag` (the reference uses its own
    Hessenberg reduction)."""
    lam = spectrum1(n, itype, rng)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q.T @ np.diag(lam) @ Q
    A = (A + A.T) / 2
    d, e = tridiag(A)
    return d, e, lam


def gluemats(*mats, gval=0.0):
    D = np.concatenate([m[0] for m in mats])
    E = []
    for k, m in enumerate(mats):
        if k:
            E.append(gval)
        E.extend(m[1])
    return D, np.array(E)


def class2(n, itype):
    """st_testmat2 (symtridiag.jl:103-142): (d, e, known eigenvalues or None)."""
    if itype == 0:
        return np.zeros(n), np.zeros(n - 1), np.zeros(n)
    if itype == 1:
        return np.ones(n), np.zeros(n - 1), np.ones(n)
    if itype == 2:
        return np.full(n, 2.0), np.ones(n - 1), None
    if itype == 3:
        return (*wilkinson(n), None)
    if itype == 4:
        # the reference lists -n:2:n here, the spectrum of the order-(n+1)
        # Clement matrix; its runtest does not use it
        return (*clement(n), None)
    if itype == 5:
        return (*legendre(n), None)
    if itype == 6:
        return (*laguerre(n), None)
    if itype == 7:
        return (*hermite(n), None)
    if itype == 8:
        small, large = TINY / EPS, HUGE * EPS
        D1, E1 = wilkinson(n)
        D, E = gluemats((small * D1, small * E1), (D1, E1), (large * D1, large * E1))
        return D, E, None
    raise ValueError(itype)


def _maxnorm(A):
    return np.max(np.abs(A)) if A.size else 0.0


def errors(D, E, lam, V):
    """runtest (symtridiag.jl:195-222) for m = n: (d_err, v_err)."""
    n = D.shape[0]
    S = np.diag(D) + np.diag(E, 1) + np.diag(E, -1)
    ntol = n
    resnorm = _maxnorm(V @ np.diag(lam) @ V.T - S)
    Tnorm = max(np.max(np.sum(np.abs(S), axis=0)) if n else 0.0, TINY)
    if resnorm < Tnorm:
        d_err = (resnorm / Tnorm) / (EPS * ntol)
    elif Tnorm < 1:
        d_err = (min(resnorm, Tnorm * ntol) / Tnorm) / (EPS * ntol)
    else:
        d_err = min(resnorm / Tnorm, ntol) / (EPS * ntol)
    v_err = _maxnorm(V @ V.T - np.eye(n)) / (EPS * ntol)
    return d_err, v_err
