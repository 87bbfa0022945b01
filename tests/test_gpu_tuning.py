"""The launch-parameter table of batched calls (gs_set_tuning, written by
genericschur/autotune.py): a non-default table gives results that pass the
parity checks, the same bits for a matrix alone and in a batch, the work-queue
QR launch the same bits as the per-bin launches, and bad entries are
rejected.  Every test restores the committed table afterwards."""
import ctypes

import numpy as np
import pytest

import checks
from genericschur import _lib, gen, tuning

pytestmark = pytest.mark.gpu

TABLE = {
    "complex": {"bin_w": 24, "lead": 1, "bins": [{"nlo": 25, "nhi": 48, "nt_hq": 1024, "zlog": 1},
                                                   {"nlo": 49, "nhi": 72, "nt_hq": 256, "zlog": 0},
                                                   {"nlo": 73, "nhi": 96, "nt_hq": 512, "zlog": 1}]},
    "real": {"bin_w": 16, "lead": 0, "bins": [{"nlo": 33, "nhi": 48, "nt_hq": 1024, "zlog": 0},
                                              {"nlo": 113, "nhi": 128, "nt_hq": 256, "zlog": 0}]},
}


@pytest.fixture
def tuned(gs):
    h = _lib.handle()
    tuning.apply(h, TABLE)
    yield h
    tuning.apply(h, tuning.load())


@pytest.mark.parametrize("cplx", [True, False])
def test_table_parity_and_single_match(gs, oracle, tuned, cplx):
    """Batched results under a non-default table: the parity checks against
    the oracle, and the same bits as single-matrix calls on the same handle."""
    mats = gen.workload2(seed=3, count=24) if cplx else gen.workload3(seed=3, count=24)[0]
    print(f"case tuned batch cplx={cplx}", flush=True)
    Ss, info = gs.gschur_batched(mats)
    assert np.all(info == 0)
    for A, S in zip(mats[:10], Ss[:10]):
        n = A.shape[0]
        eps = np.finfo(np.float64).eps
        checks.schur_invariants(A, S.T, S.Z, S.values, 20.0)
        assert checks.rel_backward(A, S.T, S.Z) <= 50 * n * eps
        T, Z, w, inf, _ = oracle.gschur(A)
        cond = checks.eig_conditions(oracle.eigvecs(T, Z, left=False), oracle.eigvecs(T, Z, left=True))
        checks.eigen_parity(A, S.values, w, cond)
        S1 = gs.gschur(A)
        assert np.array_equal(S.T, S1.T) and np.array_equal(S.Z, S1.Z)


def test_queue_matches_bins(gs, tuned):
    """One work-queue QR launch over all bins runs the same per-matrix code as
    the per-bin launches: identical bits (with the Z log on some bins)."""
    h = tuned
    mats = gen.workload2(seed=4, count=300)
    print("case queue vs bins", flush=True)
    tuning.set_section(h, True, dict(TABLE["complex"], queue=0))
    S0, i0 = gs.gschur_batched(mats)
    tuning.set_section(h, True, dict(TABLE["complex"], queue=1, slots=2))
    S1, i1 = gs.gschur_batched(mats)
    assert np.array_equal(i0, i1) and np.all(i0 == 0)
    for a, b in zip(S0, S1):
        assert np.array_equal(a.T, b.T) and np.array_equal(a.Z, b.Z) and np.array_equal(a.values, b.values)


def test_bad_entries_rejected(gs, tuned):
    h = tuned
    f = _lib.lib().gs_set_tuning
    for row in ((9, 0, 0, 0), (0, 1, 0, 2), (1, 33, 48, 100), (1, 1, 64, 64), (2, 10, 5, 1), (2, 1, 16, 3)):
        r = np.array(row, dtype=np.int32)
        rc = f(h, 1, r.ctypes.data_as(ctypes.c_void_p), 1)
        assert rc == _lib.GS_E_ARG, row
