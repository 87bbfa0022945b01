"""Seeded synthetic inputs for the five workloads of the Schur hot path.

Every matrix is drawn from its own counter-based Philox4x32-10 stream
(numpy.random.Philox; Salmon et al., SC'11) keyed by (seed, workload << 32 |
index), so a matrix is reproducible on its own, on any machine with the same
numpy, independent of batch size.  Normal deviates come from numpy's
Generator.standard_normal on that stream.

  workload 1  one n x n ComplexF64 Ginibre matrix, Re, Im ~ N(0, 1/2), n = 256
  workload 2  batch of ComplexF64 Ginibre matrices, n ~ U{8..96}
  workload 3  batch of Float64 matrices, n ~ U{16..128}, three kinds:
              1/2 real Ginibre N(0,1); 1/4 Toeplitz with -1 on the
              subdiagonal and +1 on the diagonal and superdiagonals 1..3;
              1/4 Q (D + N) Q^T with Haar Q, D block diagonal (1x1 ~ U[-1,1]
              or 2x2 rotations with angle ~ U[0, pi], 50/50), N strictly
              block-upper N(0, 1e-2)
  workload 4  one n x n Float64 A = Q (Lambda + U) Q^T, n = 4096: half real
              eigenvalues +-10^U[-3,3], half 2x2 blocks r rot(theta) with
              r = 10^U[-3,3], theta ~ U(0, pi); U block-strictly-upper N(0,1)
  workload 5  batch of Float64 real Ginibre N(0,1), n ~ U{32..64}, redrawn
              until the minimum eigenvalue separation is >= 1e-3
"""
import numpy as np


def stream(seed, workload, index):
    key = np.array([int(seed) & 0xFFFFFFFFFFFFFFFF, ((int(workload) << 32) | (int(index) & 0xFFFFFFFF))],
                   dtype=np.uint64)
    return np.random.Generator(np.random.Philox(key=key))


def ginibre_complex(n, rng):
    s = np.sqrt(0.5)
    re = rng.standard_normal((n, n)) * s
    im = rng.standard_normal((n, n)) * s
    return re + 1j * im


def ginibre_real(n, rng):
    return rng.standard_normal((n, n))


def haar_orthogonal(n, rng):
    """Haar-distributed orthogonal matrix: QR of a Gaussian matrix with the
    signs of diag(R) absorbed into Q (Mezzadri, Notices AMS 54, 2007)."""
    G = rng.standard_normal((n, n))
    Q, R = np.linalg.qr(G)
    d = np.sign(np.diag(R))
    d[d == 0] = 1.0
    return Q * d


def toeplitz_grcar(n):
    A = np.zeros((n, n))
    for k in range(4):
        A += np.diag(np.ones(n - k), k)
    A -= np.diag(np.ones(n - 1), -1)
    return A


def block_structure(n, rng, p2=0.5):
    """Random partition of 1..n into 1x1 and 2x2 diagonal blocks."""
    blocks = []
    i = 0
    while i < n:
        if i + 1 < n and rng.random() < p2:
            blocks.append((i, 2))
            i += 2
        else:
            blocks.append((i, 1))
            i += 1
    return blocks


def qdnq(n, rng):
    """Q (D + N) Q^T of workload 3."""
    Q = haar_orthogonal(n, rng)
    blocks = block_structure(n, rng, 0.5)
    D = np.zeros((n, n))
    mask = np.triu(np.ones((n, n)), 1)
    for (i, k) in blocks:
        if k == 1:
            D[i, i] = rng.uniform(-1.0, 1.0)
        else:
            th = rng.uniform(0.0, np.pi)
            c, s = np.cos(th), np.sin(th)
            D[i:i + 2, i:i + 2] = [[c, -s], [s, c]]
            mask[i, i + 1] = 0.0
    N = rng.standard_normal((n, n)) * 0.1 * mask
    return Q @ (D + N) @ Q.T


def workload1(seed=0, n=256):
    return ginibre_complex(n, stream(seed, 1, 0))


def sizes(seed, workload, count, nmin, nmax):
    return stream(seed, workload, 0xFFFFFFFF).integers(nmin, nmax + 1, size=count).astype(np.int32)


def workload2(seed=0, count=8192, nmin=8, nmax=96):
    ns = sizes(seed, 2, count, nmin, nmax)
    return [ginibre_complex(int(n), stream(seed, 2, i)) for i, n in enumerate(ns)]


def workload3_kinds(seed, count):
    k = np.zeros(count, dtype=np.int32)
    k[count // 2: count // 2 + count // 4] = 1
    k[count // 2 + count // 4:] = 2
    stream(seed, 3, 0xFFFFFFFE).shuffle(k)
    return k


def workload3(seed=0, count=4096, nmin=16, nmax=128):
    ns = sizes(seed, 3, count, nmin, nmax)
    kinds = workload3_kinds(seed, count)
    out = []
    for i, (n, kd) in enumerate(zip(ns, kinds)):
        rng = stream(seed, 3, i)
        n = int(n)
        if kd == 0:
            out.append(ginibre_real(n, rng))
        elif kd == 1:
            out.append(toeplitz_grcar(n))
        else:
            out.append(qdnq(n, rng))
    return out, kinds


def workload4_parts(seed=0, n=4096):
    """Returns (A, Lambda + U, exact eigenvalues of Lambda + U)."""
    rng = stream(seed, 4, 0)
    nblk2 = n // 4
    nreal = n - 2 * nblk2
    kinds = np.array([1] * nreal + [2] * nblk2)
    rng.shuffle(kinds)
    M = np.zeros((n, n))
    mask = np.triu(np.ones((n, n)), 1)
    evs = []
    i = 0
    for k in kinds:
        if k == 1:
            lam = rng.choice([-1.0, 1.0]) * 10.0 ** rng.uniform(-3.0, 3.0)
            M[i, i] = lam
            evs.append(lam)
            i += 1
        else:
            r = 10.0 ** rng.uniform(-3.0, 3.0)
            th = rng.uniform(0.0, np.pi)
            c, s = r * np.cos(th), r * np.sin(th)
            M[i:i + 2, i:i + 2] = [[c, -s], [s, c]]
            mask[i, i + 1] = 0.0
            evs.extend([complex(c, s), complex(c, -s)])
            i += 2
    M += rng.standard_normal((n, n)) * mask
    Q = haar_orthogonal(n, rng)
    A = Q @ M @ Q.T
    return A, M, np.array(evs)


def workload4(seed=0, n=4096):
    return workload4_parts(seed, n)[0]


def min_separation(A):
    ev = np.linalg.eigvals(A)
    d = np.abs(ev[:, None] - ev[None, :])
    np.fill_diagonal(d, np.inf)
    return float(d.min())


def workload5(seed=0, count=2048, nmin=32, nmax=64, min_sep=1e-3):
    """Returns (matrices, w, v): loss weights per eigenvalue slot."""
    ns = sizes(seed, 5, count, nmin, nmax)
    mats, ws, vs = [], [], []
    for i, n in enumerate(ns):
        n = int(n)
        redraw = 0
        while True:
            rng = stream(seed, 5, i + (redraw << 20))
            A = ginibre_real(n, rng)
            if min_separation(A) >= min_sep:
                break
            redraw += 1
        mats.append(A)
        ws.append(rng.standard_normal(n))
        vs.append(rng.standard_normal(n))
    return mats, ws, vs


def godunov():
    """Godunov's 7x7 matrix with eigenvalues {-4,-2,-1,0,1,2,4}
    (test/testfuncs.jl:127-142)."""
    A = np.array([
        [289, 2064, 336, 128, 80, 32, 16],
        [1152, 30, 1312, 512, 288, 128, 32],
        [-29, -2000, 756, 384, 1008, 224, 48],
        [512, 128, 640, 0, 640, 512, 128],
        [1053, 2256, -504, -384, -756, 800, 208],
        [-287, -16, 1712, -128, 1968, -30, 2032],
        [-2176, -287, -1565, -512, -541, -1152, -289],
    ], dtype=np.float64)
    return A, np.array([-4.0, -2.0, -1.0, 0.0, 1.0, 2.0, 4.0]), 7.0e16


# ------------------------------------------------- test-matrix families
# Re-creations of the families the reference's tests draw from LAPACK's
# matrix generators (test/complex.jl:162-380, test/real.jl:160-300 through
# test/TMGlib.jl's latme!/latmr!).  The reference's TMG binaries and Julia's
# RNG stream are not available, so the families (eigenvalue modes,
# magnitudes, eigenvector conditioning) are rebuilt here with the seeded
# streams above; the matrices themselves differ from the reference's.

ULP = float(np.finfo(np.float64).eps)
UNFL = float(np.finfo(np.float64).tiny)
OVFL = 1.0 / UNFL
MAGNS = (1.0, OVFL * ULP, UNFL / ULP)  # test/complex.jl:183, test/real.jl:174
RTULPI = 1.0 / np.sqrt(ULP)


def mode_values(n, mode, cond, rng):
    """Magnitudes of LAPACK xLATM1's eigenvalue modes: 1 one large, rest
    1/cond; 2 one small; 3 geometric 1 .. 1/cond; 4 arithmetic; 5 random
    log-uniform in [1/cond, 1]; 6 uniform in [-1, 1] (sign included)."""
    if n == 0:
        return np.zeros(0)
    i = np.arange(n, dtype=np.float64)
    if mode == 1:
        d = np.full(n, 1.0 / cond)
        d[0] = 1.0
    elif mode == 2:
        d = np.ones(n)
        d[-1] = 1.0 / cond
    elif mode == 3:
        d = cond ** (-i / max(n - 1, 1))
    elif mode == 4:
        d = 1.0 - i / max(n - 1, 1) * (1.0 - 1.0 / cond)
    elif mode == 5:
        d = np.exp(-rng.random(n) * np.log(cond))
    elif mode == 6:
        d = rng.uniform(-1.0, 1.0, n)
    else:
        raise ValueError(mode)
    return d


def _random_phases(n, rng, cplx):
    if cplx:
        return np.exp(2j * np.pi * rng.random(n))
    return np.where(rng.random(n) < 0.5, -1.0, 1.0)


def _haar(n, rng, cplx):
    if not cplx:
        return haar_orthogonal(n, rng)
    G = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * np.sqrt(0.5)
    Q, R = np.linalg.qr(G)
    d = np.diag(R)
    ph = np.where(d == 0, 1.0, d / np.abs(np.where(d == 0, 1.0, d)))
    return Q * ph


def _scale_max(A, anorm):
    m = np.abs(A).max() if A.size else 0.0
    if m == 0 or anorm == 1.0 and m == 1.0:
        return A
    # two steps so that neither the intermediate nor the result over/underflows
    return (A / m) * anorm


def latme_like(n, mode, cond, simcond, anorm, rng, cplx):
    """General matrix with specified eigenvalues (xLATME's role in
    zdrves/ddrves, test/complex.jl:235-255): X (D + U) X^-1 with D's
    magnitudes from `mode` and `cond` (random signs / phases), U strictly
    upper uniform in [-1, 1] (for real input, paired 2x2 rotation blocks turn
    half of the eigenvalues complex), X = Q1 S Q2 with singular values S
    arithmetic from 1 to 1/simcond (simcond = 1: X unitary), then scaled to
    max |a_ij| = anorm."""
    d = mode_values(n, mode, cond, rng)
    M = np.zeros((n, n), dtype=np.complex128 if cplx else np.float64)
    if cplx:
        M[np.diag_indices(n)] = d * _random_phases(n, rng, True)
        M += np.triu(rng.uniform(-1, 1, (n, n)) + 1j * rng.uniform(-1, 1, (n, n)), 1)
    else:
        dd = d * _random_phases(n, rng, False)
        i = 0
        mask = np.triu(np.ones((n, n)), 1)
        while i < n:
            if i + 1 < n and rng.random() < 0.5:
                r = abs(dd[i])
                th = rng.uniform(0.0, np.pi)
                M[i:i + 2, i:i + 2] = [[r * np.cos(th), -r * np.sin(th)], [r * np.sin(th), r * np.cos(th)]]
                mask[i, i + 1] = 0.0
                i += 2
            else:
                M[i, i] = dd[i]
                i += 1
        M += rng.uniform(-1, 1, (n, n)) * mask
    if n > 0:
        s = mode_values(n, 4, simcond, rng) if simcond != 1.0 else np.ones(n)
        X = (_haar(n, rng, cplx) * s) @ _haar(n, rng, cplx)
        M = X @ M @ np.linalg.inv(X)
    return _scale_max(M, anorm)


def latmr_like(n, anorm, rng, cplx, kind="general"):
    """Random-entry families of xLATMR (test/complex.jl:258-380): entries
    uniform in [-1, 1] (complex: both parts), shaped as `kind` = "general",
    "diagonal", "hermitian" (symmetric for real input) or "triangular",
    scaled to max |a_ij| = anorm."""
    A = rng.uniform(-1, 1, (n, n))
    if cplx:
        A = A + 1j * rng.uniform(-1, 1, (n, n))
    if kind == "diagonal":
        A = np.diag(np.diag(A))
    elif kind == "hermitian":
        A = np.triu(A)
        A = A + np.triu(A, 1).conj().T
        if cplx:
            A[np.diag_indices(n)] = A.diagonal().real
    elif kind == "triangular":
        A = np.triu(A)
    elif kind != "general":
        raise ValueError(kind)
    return _scale_max(A, anorm)


def jordan(n, anorm=1.0, cplx=False):
    """diagm(1 => ones(n-1)) + anorm I (test/complex.jl:221-232,
    test/real.jl:180-191)."""
    A = np.diag(np.ones(max(n - 1, 0)), 1) + anorm * np.eye(n)
    return A.astype(np.complex128) if cplx else A


def cyclic_shift(n, cplx=False):
    """The cyclic down-shift permutation, already Hessenberg with a zero
    diagonal: the standard shifts of its trailing 2 x 2 leave it unchanged,
    so the iteration reaches the exceptional shifts (GenericSchur.jl:302-307,
    :614-627) before it deflates."""
    A = np.diag(np.ones(max(n - 1, 0)), -1)
    if n > 0:
        A[0, n - 1] = 1.0
    return A.astype(np.complex128) if cplx else A


def triangular_repeated(n, dmode, rng):
    """Upper triangular complex matrix with a repeated diagonal entry:
    d[1] = d[n-2] exactly (dmode 0), times nextfloat(1) (1) or prevfloat(1)
    (2) (test/complex.jl:327-364)."""
    d = rng.random(n) + 1j * rng.random(n)
    if dmode == 0:
        d[1] = d[n - 2]
    elif dmode == 1:
        d[1] = d[n - 2] * np.nextafter(1.0, 2.0)
    else:
        d[1] = d[n - 2] * np.nextafter(1.0, 0.0)
    return np.triu(rng.random((n, n)) + 1j * rng.random((n, n)), 1) + np.diag(d)
