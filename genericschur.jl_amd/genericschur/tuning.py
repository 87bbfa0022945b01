"""Launch-parameter tables of the batched calls (gs_set_tuning).

A table is the JSON the autotuner (genericschur/autotune.py) writes:

    {"device": ..., "workload": ..., "complex": SECTION, "real": SECTION, "measured": ...}
    SECTION = {"bin_w": int, "lead": int, "queue": 0|1, "slots": int,
               "bins": [{"nlo": int, "nhi": int, "nt_hq": int, "zlog": -1|0|1}, ...]}

Every key is optional; what a section leaves out keeps the library's
built-in default.  The table committed beside this module
(tuning_mi355x.json) is applied to every handle the package creates;
GSCHUR_TUNING names another file, or "none" for the built-in defaults."""
import ctypes
import json
import os

import numpy as np

DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_mi355x.json")
_GLOBAL_KEYS = {"bin_w": 1, "lead": 2, "queue": 3, "slots": 4}


def load(path=None):
    """The table at `path` (default: GSCHUR_TUNING, else the committed one),
    or None."""
    if path is None:
        env = os.environ.get("GSCHUR_TUNING")
        if env and env.lower() == "none":
            return None
        path = env or DEFAULT_PATH
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def rows(section):
    """The int32 rows {kind, a, b, value} of gs_set_tuning for one section."""
    out = []
    for key, code in _GLOBAL_KEYS.items():
        if section.get(key) is not None:
            out.append((0, code, 0, int(section[key])))
    for b in section.get("bins", []):
        if b.get("nt_hq"):
            out.append((1, int(b["nlo"]), int(b["nhi"]), int(b["nt_hq"])))
        if b.get("zlog", -1) != -1:
            out.append((2, int(b["nlo"]), int(b["nhi"]), int(b["zlog"])))
    return np.array(out, dtype=np.int32).reshape(-1, 4)


def set_section(h, is_complex, section):
    """Applies one section (None: the built-in defaults) to handle h."""
    from . import _lib
    r = np.ascontiguousarray(rows(section) if section else np.zeros((0, 4), dtype=np.int32))
    ptr = r.ctypes.data_as(ctypes.c_void_p) if r.size else None
    _lib.check(_lib.lib().gs_set_tuning(h, int(bool(is_complex)), ptr, int(r.shape[0])), h, "gs_set_tuning")


def apply(h, table):
    """Applies both sections of `table` (None: built-in defaults) to h."""
    table = table or {}
    set_section(h, True, table.get("complex"))
    set_section(h, False, table.get("real"))
