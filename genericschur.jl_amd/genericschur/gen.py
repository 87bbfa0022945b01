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
