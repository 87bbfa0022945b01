"""ctypes binding of libgschur_mi355x.so (the C ABI in include/gschur_mi355x.h).

The product path has no CPU fallback: if the library is missing, or no GPU is
visible when a compute entry point needs one, the call raises.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libgschur_mi355x.so"))

GS_WANTZ = 0x01
GS_SCALE = 0x02
GS_HESSENBERG = 0x04
GS_CHECK_SUBDIAG = 0x08
GS_DEVICE_PTRS = 0x10
GS_ASYNC = 0x20
GS_PLAIN = 0x40
GS_TIMING = 0x80
GS_QR_BAND = 0x100
GS_ZDIRECT = 0x200
GS_ZLOG_TINY = 0x400
GS_SORT = 0x800
GS_LAUNCH_REC = 6

GS_OK = 0
GS_E_DIM = -1
GS_E_ARG = -3
GS_E_UNSUPPORTED = -4
GS_E_HIP = -100

# every symbol the header declares
EXPORTS = (
    "gs_version", "gs_create", "gs_destroy", "gs_last_error", "gs_synchronize", "gs_set_stream",
    "gs_schur_z", "gs_schur_d", "gs_schur_batched_z", "gs_schur_batched_d", "gs_batched_max_n",
    "gs_trevc_z", "gs_trevc_d", "gs_trevc_batched_z", "gs_trevc_batched_d",
    "gs_eig_vjp_batched_d", "gs_last_stats", "gs_last_launches", "gs_dgemm", "gs_schur_batched_z_noalloc",
    "gs_balance_z", "gs_balance_d", "gs_balance_apply", "gs_triangularize_d", "gs_enormalize", "gs_eig_rcond",
    "gs_set_tuning", "gs_ordschur_z", "gs_ordschur_d", "gs_eigcond_z",
    "gs_heev_z", "gs_heev_d", "gs_heev_max_n", "gs_hetrd_z", "gs_hetrd_d", "gs_steqr",
)

_lib = None
_lock = threading.RLock()
_handles = {}

c_i64 = ctypes.c_int64
c_int = ctypes.c_int
c_vp = ctypes.c_void_p
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_i64p = ctypes.POINTER(ctypes.c_int64)


class GSError(RuntimeError):
    pass


def _share_torch_hip_runtime():
    """Make the library bind to the HIP runtime that torch ships, when torch is installed.

    torch's wheel carries its own libamdhip64/libhsa-runtime64.  If this library
    were loaded first it would pull in /opt/rocm's copies, the process would
    hold two HIP runtimes, and the second one to start finds no GPU (torch
    then reports "No HIP GPUs are available"); streams and device pointers
    could not be shared either.  Loading torch's runtime first (globally,
    without importing torch) makes the library's libamdhip64.so.7 dependency
    resolve to that same copy whatever the import order.  Without torch the
    system runtime is used.
    """
    if os.environ.get("GSCHUR_SYSTEM_HIP") == "1":
        return None
    try:
        import importlib.util
        spec = importlib.util.find_spec("torch")
    except (ImportError, ValueError):
        return None
    if spec is None or not spec.submodule_search_locations:
        return None
    tlib = os.path.join(list(spec.submodule_search_locations)[0], "lib")
    loaded = None
    for name in ("libhsa-runtime64.so", "libamdhip64.so"):
        p = os.path.join(tlib, name)
        if os.path.exists(p):
            ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
            loaded = p
    return loaded


def lib():
    """Load the library (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIBPATH):
            raise GSError(f"HIP library not built: {LIBPATH} is missing (run __graft_entry__.build())")
        _share_torch_hip_runtime()
        L = ctypes.CDLL(LIBPATH)
        L.gs_version.restype = c_int
        L.gs_create.argtypes = [ctypes.POINTER(c_vp), c_int]
        L.gs_destroy.argtypes = [c_vp]
        L.gs_last_error.argtypes = [c_vp]
        L.gs_last_error.restype = ctypes.c_char_p
        L.gs_synchronize.argtypes = [c_vp]
        L.gs_set_stream.argtypes = [c_vp, c_vp]
        sig_single = [c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_int, c_i64, c_i64, c_i32p]
        L.gs_schur_z.argtypes = sig_single
        L.gs_schur_d.argtypes = sig_single
        sig_b = [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int]
        L.gs_schur_batched_z.argtypes = sig_b
        L.gs_schur_batched_d.argtypes = sig_b
        L.gs_schur_batched_z_noalloc.argtypes = sig_b
        L.gs_batched_max_n.argtypes = [c_int]
        L.gs_set_tuning.argtypes = [c_vp, c_int, c_vp, c_int]
        L.gs_ordschur_z.argtypes = [c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_int]
        L.gs_eigcond_z.argtypes = [c_vp, c_i64, c_vp, c_i64, c_i64, c_vp, c_int]
        L.gs_ordschur_d.argtypes = [c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp, c_vp, c_int]
        sig_heev = [c_vp, c_i64, c_i64, c_vp, c_i64, c_i64, c_int, c_vp, c_vp, c_i64, c_i64, c_vp, c_int]
        L.gs_heev_z.argtypes = sig_heev
        L.gs_heev_d.argtypes = sig_heev
        L.gs_heev_max_n.argtypes = [c_int]
        sig_hetrd = [c_vp, c_i64, c_vp, c_i64, c_int, c_vp, c_vp, c_vp, c_i64, c_int]
        L.gs_hetrd_z.argtypes = sig_hetrd
        L.gs_hetrd_d.argtypes = sig_hetrd
        L.gs_steqr.argtypes = [c_vp, c_i64, c_vp, c_vp, c_vp, c_i64, c_i64, c_int, c_vp, c_int]
        sig_tv = [c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_int, c_int, c_int]
        L.gs_trevc_z.argtypes = sig_tv
        L.gs_trevc_d.argtypes = sig_tv
        sig_tvb = [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_int, c_int, c_int]
        L.gs_trevc_batched_z.argtypes = sig_tvb
        L.gs_trevc_batched_d.argtypes = sig_tvb
        L.gs_eig_vjp_batched_d.argtypes = [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_int]
        sig_bal = [c_vp, c_i64, c_vp, c_i64, c_int, c_int, c_vp, c_vp, c_vp, c_int]
        L.gs_balance_z.argtypes = sig_bal
        L.gs_balance_d.argtypes = sig_bal
        L.gs_balance_apply.argtypes = [c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_i64, c_int, c_int]
        L.gs_triangularize_d.argtypes = [c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_int]
        L.gs_enormalize.argtypes = [c_vp, c_i64, c_i64, c_vp, c_i64, c_int]
        L.gs_eig_rcond.argtypes = [c_vp, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_int]
        L.gs_last_stats.argtypes = [c_vp, c_i64p, c_int]
        L.gs_last_launches.argtypes = [c_vp, c_vp, c_int]
        L.gs_dgemm.argtypes = [c_vp, c_int, c_int, c_i64, c_i64, c_i64, ctypes.c_double, c_vp, c_i64, c_vp, c_i64,
                               ctypes.c_double, c_vp, c_i64]
        _lib = L
        return _lib


def handle(device=0):
    """The process-wide handle for `device` (created on first use)."""
    h = _handles.get(device)
    if h is not None:
        return h
    with _lock:
        h = _handles.get(device)
        if h is None:
            hp = c_vp()
            rc = lib().gs_create(ctypes.byref(hp), int(device))
            if rc != GS_OK:
                raise GSError(f"gs_create(device={device}) failed with status {rc}: no usable GPU")
            h = hp
            _handles[device] = h
            # the committed launch-parameter table of the autotuner (GSCHUR_TUNING
            # names another table, or "none" for the built-in defaults)
            from . import tuning
            tuning.apply(h, tuning.load())
    return h


def check(rc, h, what):
    if rc != GS_OK:
        msg = lib().gs_last_error(h)
        msg = msg.decode() if msg else ""
        raise GSError(f"{what} failed with status {rc}: {msg}")


def last_stats(device=0):
    """(sweeps, bulge steps, kernel microseconds) of the last Schur call."""
    out = (ctypes.c_int64 * 3)()
    k = lib().gs_last_stats(handle(device), out, 3)
    return tuple(int(out[i]) for i in range(k))


def last_launches(device=0):
    """Records of the last call made with GS_TIMING: a list of dicts.  Batched
    calls give kind 'hq' / 'qr' / 'zlog' (Z log applied after the complex QR
    sweeps) with (nmin, nmax, count, sum_n3, us); the large
    single-matrix path gives 'gemm' (m, n, k, flops, us) and 'phase' (phase:
    0 copy+scale, 1 Hessenberg, 2 Q formation, 3 QR, 4 finish; us)."""
    h = handle(device)
    k = lib().gs_last_launches(h, None, 0)
    if k <= 0:
        return []
    buf = (ctypes.c_double * (k * GS_LAUNCH_REC))()
    lib().gs_last_launches(h, buf, k)
    recs = []
    for r in range(k):
        v = buf[r * GS_LAUNCH_REC:(r + 1) * GS_LAUNCH_REC]
        kind = int(v[0])
        if kind in (0, 1, 4):
            recs.append({"kind": {0: "hq", 1: "qr", 4: "zlog"}[kind], "nmin": int(v[1]), "nmax": int(v[2]),
                         "count": int(v[3]), "sum_n3": v[4], "us": v[5]})
        elif kind == 2:
            recs.append({"kind": "gemm", "m": int(v[1]), "n": int(v[2]), "k": int(v[3]), "flops": v[4],
                         "us": v[5]})
        else:
            recs.append({"kind": "phase", "phase": int(v[1]), "us": v[5]})
    return recs
