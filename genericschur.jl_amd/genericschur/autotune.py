"""Autotuner of the batched launch parameters (gs_set_tuning).

Times whole batched Schur decompositions (device-resident inputs, graph
replays, median of several) on the GPU it runs on and picks, by coordinate
descent from the library's built-in defaults:
  - the size-bin width;
  - per size bin, the Hessenberg + Q workgroup size and whether the bin's
    QR launch logs its Z updates;
  - how many of the largest bins' Hessenberg launches run first, alone;
  - complex batches: one work-queue QR launch over all bins or one per bin,
    and the work-queue launch's workgroups per CU.
A change is kept only when it beats the incumbent by more than `margin`
(run-to-run noise is a few tenths of a percent).  The result is a table in
the format of genericschur/tuning.py, with every trial recorded:

    python -m genericschur.autotune --out table.json     (from genericschur.jl_amd/)

None of the parameters changes a matrix's results between its batched and
single-matrix calls; the Hessenberg workgroup size changes the rounding of
the reduction (the parity tests hold at any of them)."""
import argparse
import datetime
import json
import sys
import time

import numpy as np

from . import _lib, gen, tuning
from .api import gschur_batched_device, pack, set_stream


def _workload(cplx, seed, count):
    if cplx:
        return gen.workload2(seed=seed, count=count)
    return gen.workload3(seed=seed, count=count)[0]


class _Batch:
    """A batch resident on the GPU and the timing of its graph replays."""

    def __init__(self, cplx, mats):
        import torch
        self.torch = torch
        self.cplx = cplx
        self.ns, self.off, flat = pack(mats, np.complex128 if cplx else np.float64)
        dev = torch.device("cuda:0")
        self.A = torch.from_numpy(flat.view(np.float64)).to(dev)
        self.T = torch.zeros_like(self.A)
        self.Z = torch.zeros_like(self.A)
        self.W = torch.zeros(2 * int(self.ns.sum()), dtype=torch.float64, device=dev)
        self.info = torch.zeros(len(self.ns), dtype=torch.int32, device=dev)
        self.stream = torch.cuda.Stream(device=dev)

    def run(self):
        gschur_batched_device(self.ns, self.off, self.A, self.T, self.Z, self.W, self.info,
                              complex_input=self.cplx, wantZ=True, scale=True, asynchronous=True)

    def time_ms(self, reps):
        torch = self.torch
        set_stream(self.stream.cuda_stream)
        try:
            self.run()  # plan and graph for the current parameters
            self.run()
            torch.cuda.synchronize()
            ts = []
            for _ in range(reps):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(self.stream)
                self.run()
                e1.record(self.stream)
                torch.cuda.synchronize()
                ts.append(e0.elapsed_time(e1))
        finally:
            set_stream(None)
        if int(self.info.cpu().numpy().max(initial=0)) != 0 or int(self.info.cpu().numpy().min(initial=0)) != 0:
            raise RuntimeError("nonzero info during tuning")
        return float(np.median(ts))


def _bins(bw, nmax):
    return [(k * bw + 1, (k + 1) * bw) for k in range((nmax + bw - 1) // bw)]


def tune(cplx, mats, reps=9, margin=0.003, log=print):
    """Coordinate descent over the parameters above for one element type;
    returns (section, trials)."""
    h = _lib.handle()
    batch = _Batch(cplx, mats)
    nmax = int(batch.ns.max())
    trials = []

    def measure(section, what):
        tuning.set_section(h, cplx, section)
        ms = batch.time_ms(reps)
        trials.append({"what": what, "ms": ms})
        log(f"  {what}: {ms:.3f} ms")
        return ms

    best = {}
    best_ms = measure(best, "defaults")

    def consider(cand, what):
        # a candidate that beats the incumbent is timed again beside it, and
        # kept only if it wins both times
        nonlocal best, best_ms
        ms = measure(cand, what)
        if ms < best_ms * (1.0 - margin):
            inc = measure(best, "incumbent (confirm)")
            ms2 = measure(cand, what + " (confirm)")
            if ms2 < inc * (1.0 - margin):
                best, best_ms = cand, min(ms, ms2)
                log(f"    kept ({best_ms:.3f} ms)")
            else:
                best_ms = min(best_ms, inc)

    for bw in ((16, 24, 32) if cplx else (16, 24, 32)):
        cand = dict(best, bin_w=bw)
        consider(cand, f"bin_w={bw}")
    bw = best.get("bin_w", 16 if cplx else 32)
    for lo, hi in _bins(bw, nmax):
        if hi > 32:
            for nt in (256, 512, 1024):
                rows = [b for b in best.get("bins", []) if (b["nlo"], b["nhi"]) != (lo, hi)]
                old = next((b for b in best.get("bins", []) if (b["nlo"], b["nhi"]) == (lo, hi)), {})
                row = dict(old, nlo=lo, nhi=hi, nt_hq=nt)
                consider(dict(best, bins=rows + [row]), f"bin {lo}-{hi} nt_hq={nt}")
        zl_ok = (hi > 48) if cplx else (32 < hi <= 128)
        if zl_ok:
            for z in (0, 1):
                rows = [b for b in best.get("bins", []) if (b["nlo"], b["nhi"]) != (lo, hi)]
                old = next((b for b in best.get("bins", []) if (b["nlo"], b["nhi"]) == (lo, hi)), {})
                row = dict(old, nlo=lo, nhi=hi, zlog=z)
                consider(dict(best, bins=rows + [row]), f"bin {lo}-{hi} zlog={z}")
    for lead in (0, 1, 2, 3):
        consider(dict(best, lead=lead), f"lead={lead}")
    if cplx:
        for slots in (4, 5, 6, 7, 8, 10):
            consider(dict(best, queue=1, slots=slots), f"queue slots={slots}")
    # the incumbent again, next to the defaults, for the record
    final = measure(best, "tuned (again)")
    base = measure({}, "defaults (again)")
    tuning.set_section(h, cplx, None)
    return best, {"trials": trials, "tuned_ms": final, "default_ms": base}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", required=True)
    ap.add_argument("--count-complex", type=int, default=8192)
    ap.add_argument("--count-real", type=int, default=4096)
    ap.add_argument("--reps", type=int, default=9)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--only", choices=["complex", "real"], default=None)
    args = ap.parse_args(argv)
    import torch
    props = torch.cuda.get_device_properties(0)
    table = {"device": f"{torch.cuda.get_device_name(0)} ({getattr(props, 'gcnArchName', '?')}, "
                       f"{props.multi_processor_count} CUs)",
             "date": datetime.date.today().isoformat(),
             "workload": {"complex": f"{args.count_complex} ComplexF64 Ginibre matrices, n ~ U{{8..96}}",
                          "real": f"{args.count_real} Float64 matrices of the mixed real workload, n ~ U{{16..128}}"},
             "measured": {}}
    for cplx, key, count in ((True, "complex", args.count_complex), (False, "real", args.count_real)):
        if args.only and args.only != key:
            continue
        print(f"tuning {key} batches ({count} matrices)", flush=True)
        t0 = time.time()
        section, rec = tune(cplx, _workload(cplx, args.seed, count), reps=args.reps,
                            log=lambda s: print(s, flush=True))
        table[key] = section
        rec["seconds"] = time.time() - t0
        table["measured"][key] = rec
        print(f"{key}: defaults {rec['default_ms']:.3f} ms, tuned {rec['tuned_ms']:.3f} ms", flush=True)
    with open(args.out, "w") as f:
        json.dump(table, f, indent=1)
    print(f"wrote {args.out}", flush=True)


if __name__ == "__main__":
    sys.exit(main())
