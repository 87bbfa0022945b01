"""Test-matrix families of the reference's test suite, re-created with the
seeded generators of genericschur.gen (the reference draws them from
LAPACK's TMG library and Julia's RNG, neither of which exists here), and the
oracle's result for a matrix with eigenvalue condition numbers.

  Jordan blocks (test/complex.jl:221-232, test/real.jl:180-191); random
  unitary / orthogonal matrices (test/complex.jl:130-135, test/real.jl:133-138);
  general matrices with specified eigenvalues, modes 4/3/1/5, eigenvector
  condition 1 and 1/sqrt(ulp), magnitudes 1, ovfl*ulp, unfl/ulp
  (test/complex.jl:235-255, test/real.jl:192-212); diagonal, Hermitian /
  symmetric, general and triangular random-entry matrices
  (test/complex.jl:258-380); triangular matrices with a repeated diagonal
  entry (test/complex.jl:327-364); 100 floatmin and floatmax / 100 scalings
  (test/real.jl:104-116); zero, identity, and the cyclic shift, which only the
  exceptional shifts move (GenericSchur.jl:302-307, :614-627).
"""
import numpy as np

import checks
from genericschur import gen

EPS = np.finfo(np.float64).eps
TINY = np.finfo(np.float64).tiny
HUGE = np.finfo(np.float64).max


def families(n, cplx, seed=0):
    """(name, matrix, normal) for one order and element type."""
    rng = gen.stream(seed, 77, 2 * n + int(cplx))
    dt = np.complex128 if cplx else np.float64
    out = [("jordan", gen.jordan(n, 1.0, cplx), False),
           ("normal", gen._haar(n, rng, cplx), True)]
    for mode in (4, 3, 1, 5):
        for ci, simcond in enumerate((1.0, gen.RTULPI)):
            out.append((f"latme-m{mode}-c{ci + 1}", gen.latme_like(n, mode, 1.0 / EPS, simcond, 1.0, rng, cplx),
                        False))
    for mi in (1, 2):
        out.append((f"latme-m5-magn{mi}", gen.latme_like(n, 5, 1.0 / EPS, gen.RTULPI, gen.MAGNS[mi], rng, cplx),
                    False))
    for mi in range(3):
        out.append((f"latmr-general-m{mi}", gen.latmr_like(n, gen.MAGNS[mi], rng, cplx, "general"), False))
    out.append(("latmr-diagonal", gen.latmr_like(n, 1.0, rng, cplx, "diagonal"), True))
    out.append(("latmr-hermitian", gen.latmr_like(n, 1.0, rng, cplx, "hermitian"), True))
    out.append(("latmr-triangular", gen.latmr_like(n, 1.0, rng, cplx, "triangular"), False))
    G = gen.ginibre_complex(n, rng) if cplx else gen.ginibre_real(n, rng)
    out.append(("tiny", G * (100 * TINY), False))
    out.append(("huge", G * (HUGE / 100 / max(1.0, np.abs(G).max())), False))
    out.append(("zero", np.zeros((n, n), dtype=dt), True))
    out.append(("identity", np.eye(n, dtype=dt), True))
    out.append(("cyclic", gen.cyclic_shift(n, cplx), True))
    if cplx and n >= 4:
        for dm in range(3):
            out.append((f"triangular-repeated-{dm}", gen.triangular_repeated(n, dm, rng), False))
    return out


_oracle_cache = {}


def oracle_result(oracle, A):
    key = (A.dtype.str, A.shape, A.tobytes())
    r = _oracle_cache.get(key)
    if r is None:
        T, Z, w, info, stats = oracle.gschur(A)
        assert info == 0
        # eigenvectors do not depend on the scale of T, but the solver's
        # perturbation floor (safemin n / ulp, vectors.jl:146) swamps a T of
        # norm near 1e-292: compute them on T scaled by a power of 2 to O(1)
        m = np.abs(T).max()
        Ts = T * 2.0 ** -np.round(np.log2(m)) if m > 0 else T
        VR = oracle.eigvecs(Ts, Z, left=False)
        VL = oracle.eigvecs(Ts, Z, left=True)
        cond = checks.eig_conditions(VR, VL)
        cond = np.where(np.isfinite(cond), cond, np.inf)
        r = (T, Z, w, cond, stats)
        _oracle_cache[key] = r
    return r
