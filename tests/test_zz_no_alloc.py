"""The no-allocation batched entry point (gs_schur_batched_z_noalloc)
against gs_schur_batched_z on the same device data: T, Z, the values and
info bit for bit equal (both run the same kernels; the no-allocation form
updates Z in place, which the Z-log tests show equal to the Z log bit for
bit), on two calls in a row, the second with a smaller batch of other
orders, so that the scratch array's leftovers from the first call must not
matter.  Also: more matrices than the scratch array holds are rejected."""
import ctypes

import numpy as np
import pytest

from genericschur import gen

pytestmark = pytest.mark.gpu


def _dev(ns, off, flat):
    import torch
    dev = torch.device("cuda:0")
    A = torch.from_numpy(flat.view(np.float64).copy()).to(dev)
    T = torch.full_like(A, 3.0)
    Z = torch.full_like(A, 3.0)
    W = torch.full((2 * int(ns.sum()),), 3.0, dtype=torch.float64, device=dev)
    info = torch.full((len(ns),), 7, dtype=torch.int32, device=dev)
    return A, T, Z, W, info


def _call(fn_name, ns, off, A, T, Z, W, info, flags):
    from genericschur import _lib
    L = _lib.lib()
    h = _lib.handle()
    a = lambda t: ctypes.c_void_p(t.data_ptr())
    ns = np.ascontiguousarray(ns, dtype=np.int32)
    off = np.ascontiguousarray(off, dtype=np.int64)
    rc = getattr(L, fn_name)(h, len(ns), ctypes.c_void_p(ns.ctypes.data), ctypes.c_void_p(off.ctypes.data), a(A),
                             a(T), a(Z), a(W), a(info), flags)
    return rc, (L.gs_last_error(h) or b"").decode()


def test_noalloc_matches_batched(gs):
    from genericschur import _lib
    import torch
    for step, (seed, count, nmin, nmax) in enumerate(((21, 300, 8, 96), (22, 97, 3, 40))):
        mats = gen.workload2(seed=seed, count=count, nmin=nmin, nmax=nmax)
        ns, off, flat = gs.pack(mats, np.complex128)
        print(f"case noalloc call {step}: {count} matrices, n in [{ns.min()}, {ns.max()}]", flush=True)
        ref = _dev(ns, off, flat)
        rc, msg = _call("gs_schur_batched_z", ns, off, *ref,
                        _lib.GS_WANTZ | _lib.GS_SCALE | _lib.GS_DEVICE_PTRS)
        assert rc == 0, msg
        got = _dev(ns, off, flat)
        rc, msg = _call("gs_schur_batched_z_noalloc", ns, off, *got, _lib.GS_SCALE)
        assert rc == 0, msg
        torch.cuda.synchronize()
        for x, y, name in zip(ref[1:], got[1:], ("T", "Z", "W", "info")):
            assert torch.equal(x, y), name
        assert int(got[4].abs().sum()) == 0
        # and the result is a Schur factorization
        Th = got[1].cpu().numpy().view(np.complex128)
        Zh = got[2].cpu().numpy().view(np.complex128)
        eps = np.finfo(np.float64).eps
        for i in range(0, count, 17):
            k, o = int(ns[i]), int(off[i])
            Ti = Th[o:o + k * k].reshape((k, k), order="F")
            Zi = Zh[o:o + k * k].reshape((k, k), order="F")
            assert np.linalg.norm(mats[i] - Zi @ Ti @ Zi.conj().T) <= 50 * k * eps * np.linalg.norm(mats[i])


def test_noalloc_rejects_oversized_batch(gs):
    from genericschur import _lib
    count = 16385
    ns = np.full(count, 2, dtype=np.int32)
    off = (np.arange(count, dtype=np.int64) * 4)
    import torch
    dev = torch.device("cuda:0")
    A = torch.zeros(8 * count, dtype=torch.float64, device=dev)
    T = torch.full_like(A, 5.0)
    Z = torch.full_like(A, 5.0)
    W = torch.zeros(4 * count, dtype=torch.float64, device=dev)
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    rc, msg = _call("gs_schur_batched_z_noalloc", ns, off, A, T, Z, W, info, 0)
    assert rc == _lib.GS_E_ARG and "16384" in msg
    torch.cuda.synchronize()
    assert bool((T == 5.0).all())
