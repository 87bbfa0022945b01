"""Writes tests/golden/fixtures.json: known-answer data for the Schur path.

Contents
  godunov        Godunov's 7x7 matrix and its exact eigenvalues with their
                 condition (test/testfuncs.jl:127-142), data only.
  gs2x2_kat      the near-degenerate 2x2 case of test/real.jl:311-317:
                 lambda = 1 +- 2 eps plus (eps/4) * U(0,1) perturbations, drawn
                 here from a seeded generator (the reference draws rand(T,2,2)).
  normalization  the 3x3 matrix of test/wrappers.jl:181-192 at m = 5 (Float64).
  mp256          seeded real and complex matrices (n = 2..12) with eigenvalues
                 computed by mpmath at 256-bit precision (mp.eig), as decimal
                 strings.  These are independent of the reference and of the
                 oracle, and pin both.

Run:  python tests/golden/make_golden.py   (takes about a minute)
"""
import json
import os

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def mp_eigs(A):
    mp.mp.prec = 256
    n = A.shape[0]
    M = mp.matrix(n, n)
    for i in range(n):
        for j in range(n):
            v = A[i, j]
            M[i, j] = mp.mpc(float(v.real), float(v.imag)) if np.iscomplexobj(A) else mp.mpf(float(v))
    E = mp.eig(M, left=False, right=False)
    return [(mp.nstr(mp.re(e), 70), mp.nstr(mp.im(e), 70)) for e in E]


def enc(A):
    if np.iscomplexobj(A):
        return {"re": A.real.tolist(), "im": A.imag.tolist()}
    return {"re": A.tolist()}


def main():
    out = {}
    G = np.array([
        [289, 2064, 336, 128, 80, 32, 16],
        [1152, 30, 1312, 512, 288, 128, 32],
        [-29, -2000, 756, 384, 1008, 224, 48],
        [512, 128, 640, 0, 640, 512, 128],
        [1053, 2256, -504, -384, -756, 800, 208],
        [-287, -16, 1712, -128, 1968, -30, 2032],
        [-2176, -287, -1565, -512, -541, -1152, -289],
    ], dtype=np.float64)
    out["godunov"] = {"A": G.tolist(), "values": [-4, -2, -1, 0, 1, 2, 4], "econd": 7.0e16}
    eps = np.finfo(np.float64).eps
    rng = np.random.default_rng(1234)
    kat = []
    for _ in range(8):
        l1, l2 = 1.0 + 2 * eps, 1.0 - 2 * eps
        B = np.diag([l1, l2]) + (eps / 4) * rng.random((2, 2))
        kat.append({"B": B.tolist(), "l1": l1, "l2": l2})
    out["gs2x2_kat"] = kat
    m = 5
    out["normalization"] = {"A": [[1, 0, 10.0 ** (-2 * m)], [1, 1, 10.0 ** (-m)], [10.0 ** (2 * m), 10.0 ** m, 1]]}
    cases = []
    for n in (2, 3, 5, 8, 12):
        for cplx in (False, True):
            r = np.random.default_rng(1000 + 10 * n + int(cplx))
            A = r.standard_normal((n, n))
            if cplx:
                A = A + 1j * r.standard_normal((n, n))
            cases.append({"n": n, "complex": cplx, "A": enc(A), "eig256": mp_eigs(A)})
    out["mp256"] = cases
    with open(os.path.join(HERE, "fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", os.path.join(HERE, "fixtures.json"))


if __name__ == "__main__":
    main()
