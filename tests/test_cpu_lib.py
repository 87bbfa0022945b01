"""CPU-side checks of the product library and its host layer (no GPU needed).

* the built library loads and exports every entry point the header declares;
* without a GPU the library fails loudly (no CPU fallback);
* host-side argument validation raises the reference's exceptions
  (DimensionMismatch for non-square input, test/errors.jl:12-13; no capture of
  non-float element types, test/errors.jl:16-28);
* the seeded generators are deterministic and produce the workloads' shapes.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import genericschur as gs
from genericschur import _lib, gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gschur_mi355x.h")


def header_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gs_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    assert set(syms) <= set(_lib.EXPORTS) | set(syms)
    assert set(_lib.EXPORTS) == set(syms)


def _gpu_visible():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-GPU failure path")
def test_fails_loudly_without_gpu():
    hp = ctypes.c_void_p()
    rc = _lib.lib().gs_create(ctypes.byref(hp), 0)
    assert rc != 0
    with pytest.raises(gs.GSError):
        gs.gschur(np.eye(3))


def test_argument_validation():
    with pytest.raises(gs.DimensionMismatch):
        gs.gschur(np.zeros((5, 4)))
    with pytest.raises(gs.DimensionMismatch):
        gs.gschur(np.zeros((5, 4), dtype=complex))
    with pytest.raises(TypeError):
        gs.gschur(np.ones((4, 4), dtype=np.int64))
    with pytest.raises(TypeError):
        gs.gschur(np.ones((4, 4), dtype=np.float32))
    with pytest.raises(gs.DimensionMismatch):
        gs.gschur_hessenberg_(np.triu(np.ones((5, 5)), -1), Z=np.ones((4, 4)))


def test_header_constants_match_host_layer():
    txt = open(HEADER).read()
    for name in ("GS_WANTZ", "GS_SCALE", "GS_HESSENBERG", "GS_CHECK_SUBDIAG", "GS_DEVICE_PTRS", "GS_ASYNC",
                 "GS_PLAIN"):
        m = re.search(r"#define %s (0x[0-9a-fA-F]+)" % name, txt)
        assert m and int(m.group(1), 16) == getattr(_lib, name), name


def test_pack_roundtrip():
    r = np.random.default_rng(0)
    mats = [r.standard_normal((k, k)) + 1j * r.standard_normal((k, k)) for k in (1, 5, 3, 8)]
    ns, off, flat = gs.pack(mats, np.complex128)
    assert list(ns) == [1, 5, 3, 8] and list(off) == [0, 1, 26, 35]
    for a, b in zip(mats, gs.unpack(ns, off, flat)):
        assert np.array_equal(a, b)
    assert list(gs.eig_offsets(ns)) == [0, 1, 6, 9]


def test_generators_deterministic():
    a = gen.workload2(seed=3, count=20)
    b = gen.workload2(seed=3, count=20)
    c = gen.workload2(seed=4, count=20)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not all(x.shape == y.shape and np.array_equal(x, y) for x, y in zip(a, c))
    assert all(8 <= m.shape[0] <= 96 and m.dtype == np.complex128 for m in a)
    # a matrix does not depend on the batch size
    d = gen.workload2(seed=3, count=5)
    assert all(np.array_equal(x, y) for x, y in zip(a[:5], d))
    # Re, Im ~ N(0, 1/2)
    big = np.concatenate([m.ravel() for m in gen.workload2(seed=1, count=200)])
    assert abs(big.real.var() - 0.5) < 0.02 and abs(big.imag.var() - 0.5) < 0.02


def test_workload3_kinds():
    mats, kinds = gen.workload3(seed=2, count=64)
    assert (kinds == 0).sum() == 32 and (kinds == 1).sum() == 16 and (kinds == 2).sum() == 16
    for A, k in zip(mats, kinds):
        n = A.shape[0]
        assert 16 <= n <= 128 and A.dtype == np.float64
        if k == 1:
            assert np.array_equal(A, gen.toeplitz_grcar(n))
            assert A[1, 0] == -1 and A[0, 3] == 1 and A[0, 4] == 0


def test_qdnq_eigenvalues():
    """Q (D + N) Q^T keeps the eigenvalues of the blocks of D."""
    rng = gen.stream(0, 99, 0)
    n = 24
    A = gen.qdnq(n, rng)
    ev = np.linalg.eigvals(A)
    assert np.all(np.abs(np.abs(ev) - 1) < 1e-8) | np.all(np.abs(ev) <= 1 + 1e-8)


def test_workload4_construction():
    A, M, ev = gen.workload4_parts(seed=0, n=64)
    assert len(ev) == 64
    got = np.linalg.eigvals(M)
    from scipy.optimize import linear_sum_assignment
    C = np.abs(got[:, None] - ev[None, :])
    r, c = linear_sum_assignment(C)
    assert C[r, c].max() <= 1e-9 * np.abs(ev).max()
    mags = np.abs(ev)
    assert mags.min() >= 1e-3 * (1 - 1e-12) and mags.max() <= 1e3 * (1 + 1e-12)


def test_workload5_separation():
    mats, w, v = gen.workload5(seed=0, count=8)
    for A in mats:
        assert 32 <= A.shape[0] <= 64
        assert gen.min_separation(A) >= 1e-3


def test_single_hip_runtime_with_torch(gs):
    """Loading the library before torch must not put a second HIP runtime in the process."""
    import importlib.util
    if importlib.util.find_spec("torch") is None:
        pytest.skip("torch not installed")
    from genericschur import _lib
    _lib.lib()
    import torch  # noqa: F401
    hips = set()
    with open("/proc/self/maps") as f:
        for line in f:
            path = line.split()[-1]
            if "libamdhip64" in path:
                hips.add(os.path.realpath(path))
    assert len(hips) == 1, hips


def test_tuning_rows_format():
    """The tuning table's conversion into gs_set_tuning rows {kind, a, b,
    value}, and the committed table's bins covering disjoint order ranges."""
    from genericschur import tuning
    r = tuning.rows({"bin_w": 24, "queue": 1, "bins": [{"nlo": 25, "nhi": 48, "nt_hq": 1024, "zlog": 0},
                                                       {"nlo": 49, "nhi": 72, "zlog": -1}]})
    assert r.tolist() == [[0, 1, 0, 24], [0, 3, 0, 1], [1, 25, 48, 1024], [2, 25, 48, 0]]
    table = tuning.load(tuning.DEFAULT_PATH)
    if table is None:
        return
    for key in ("complex", "real"):
        bins = sorted((b["nlo"], b["nhi"]) for b in (table.get(key) or {}).get("bins", []))
        for (a0, b0), (a1, b1) in zip(bins, bins[1:]):
            assert a0 <= b0 < a1 <= b1


def test_hermitian_algorithm_argument():
    """alg mirrors the reference: QRIteration runs, RobustRepresentations
    falls back to it with a one-time warning (symtridiag.jl:21-28), and
    DivideAndConquer (symtri_dc.jl) is refused; checked before any GPU work."""
    import warnings
    from genericschur import api
    api._check_alg("QRIteration")
    api._MRRR_WARNED[0] = False
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        api._check_alg("RobustRepresentations")
        api._check_alg("RobustRepresentations")
    assert len(w) == 1
    with pytest.raises(NotImplementedError):
        api.geigen_symtridiag(np.ones(3), np.ones(2), alg="DivideAndConquer")
    with pytest.raises(api.ArgumentError):
        api._check_alg("Bisection")


def test_lds_resident_order_limits():
    """The largest orders whose Hessenberg phase fits the 160 KiB of LDS of
    a CU: H with leading dimension n | 1 (odd), tau and two work vectors.
    Complex: 16 (99 * 99 + 3 * 99) + reduction slots <= 163840 < 16 (101 * 100
    + 3 * 100); real: 8 (141 * 141 + 3 * 141) + slots <= 163840 <
    8 (143 * 142 + 3 * 142).  The Hessenberg pass's partial sums only take
    LDS at orders where they fit, so they do not lower either limit."""
    def need(n, es):
        ld = n | 1
        return es * (ld * n + 3 * n) + 8 * 5

    assert gs.batched_max_n(True) == 99 and gs.batched_max_n(False) == 141
    assert need(99, 16) <= 160 * 1024 < need(100, 16)
    assert need(141, 8) <= 160 * 1024 < need(142, 8)
