#!/usr/bin/env python3
"""Benchmark of the batched Schur hot path on one MI355X.

A step is one batched Schur decomposition (T, Z and eigenvalues) of 8192
ComplexF64 Ginibre matrices with n ~ U{8..96} (workload 2), inputs already
resident in HBM.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--count 8192]

Alongside the throughput it reports:
  roofline      for the dominant kernel (the QR-sweep kernel, one launch per
                size bin): its model flops per launch (100 n^3 per complex
                matrix for Hessenberg + Q + sweeps, Golub & Van Loan 4th ed.
                7.5.6, minus the 56/3 n^3 of Hessenberg + Q done by the
                other kernel) over its average launch duration, measured with
                HIP events around every launch on its own stream in one
                instrumented (eager, graph-free) pass after the timed steps,
                against the 78.6 TFLOP/s FP64 peak; HBM traffic per launch
                from rocprofv3 FETCH_SIZE/WRITE_SIZE runs when a summary of
                them is present in profiles/;
  cpu_baseline  the CPU oracle (a C++ restatement of GenericSchur's gschur!)
                on a bounded sample of the same workload, OpenMP over
                matrices on the host cores;
  dropin        per-matrix host-memory gs_schur_z calls on the same GPU (the
                drop-in schur! path) on a sample, and the batched/drop-in ratio;
  check         acceptance residuals on a sample of the batch.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "genericschur.jl_amd"))

PEAK_FP64 = 78.6e12
PEAK_HBM = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_probe(mib):
    import torch
    dev = torch.device("cuda:0")
    nbytes = mib << 20
    src = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes}), flush=True)


def load_traffic():
    """HBM bytes per QR-kernel launch from the committed rocprofv3 summary, if any."""
    path = os.path.join(ROOT, "profiles", "traffic_zqr_workload2.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--count", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--cpu-sample", type=int, default=8192)
    ap.add_argument("--dropin-sample", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-dropin", action="store_true")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only copy a MIB-MiB device buffer (calibrates rocprof FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported (batches shard with no exchange; "
              "run one process per GPU on a slice of the batch)", file=sys.stderr)
        sys.exit(2)

    import genericschur as gs
    from genericschur import gen
    from genericschur import _lib

    t0 = time.time()
    mats = gen.workload2(seed=args.seed, count=args.count)
    ns, off, flat = gs.pack(mats, np.complex128)
    log(f"generated {len(mats)} matrices in {time.time() - t0:.1f}s; sum n^3 = {int((ns.astype(np.int64) ** 3).sum())}")

    import torch
    dev = torch.device("cuda:0")
    A = torch.from_numpy(flat.view(np.float64)).to(dev)
    T = torch.zeros_like(A)
    Z = torch.zeros_like(A)
    W = torch.zeros(2 * int(ns.sum()), dtype=torch.float64, device=dev)
    info = torch.zeros(len(ns), dtype=torch.int32, device=dev)
    torch.cuda.synchronize()

    def step():
        gs.gschur_batched_device(ns, off, A, T, Z, W, info, complex_input=True, wantZ=True, scale=True,
                                 asynchronous=True)

    # the library runs on a torch stream of its own (default priority), so
    # torch's HIP events on that stream bracket exactly the batches' kernels
    stream = torch.cuda.Stream(device=dev)
    gs.set_stream(stream.cuda_stream)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t1 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    wall = t2 - t1
    kern_s = ev0.elapsed_time(ev1) * 1e-3 / args.steps
    # one instrumented pass (not part of the timed steps): events around every
    # launch on the stream it runs on
    gs.gschur_batched_device(ns, off, A, T, Z, W, info, complex_input=True, wantZ=True, scale=True, timing=True)
    recs = _lib.last_launches()
    gs.set_stream(None)
    ms_per_step = 1e3 * wall / args.steps
    value = args.count * args.steps / wall
    info_h = info.cpu().numpy()
    stats = _lib.last_stats()
    n64 = ns.astype(np.float64)
    flops = float((100.0 * n64 ** 3).sum())
    bytes_alg = float((3.0 * n64 ** 2 * 16 + 16 * n64).sum())
    batch_rate = flops / kern_s
    # dominant kernel: the QR-sweep launches
    qr = [r for r in recs if r["kind"] == "qr"]
    hq = [r for r in recs if r["kind"] == "hq"]
    qr_flops = (100.0 - 56.0 / 3.0) * sum(r["sum_n3"] for r in qr)
    qr_us = sum(r["us"] for r in qr)
    achieved = qr_flops / (qr_us * 1e-6) if qr_us > 0 else 0.0
    traffic = load_traffic()

    # acceptance residuals on a sample of the batch
    Th = T.cpu().numpy().view(np.complex128)
    Zh = Z.cpu().numpy().view(np.complex128)
    samp = np.linspace(0, len(ns) - 1, min(256, len(ns))).astype(int)
    worst_b, worst_o = 0.0, 0.0
    eps = np.finfo(np.float64).eps
    for i in samp:
        k, o = int(ns[i]), int(off[i])
        Ai = mats[i]
        Ti = Th[o:o + k * k].reshape((k, k), order="F")
        Zi = Zh[o:o + k * k].reshape((k, k), order="F")
        rb = np.linalg.norm(Ai - Zi @ Ti @ Zi.conj().T) / np.linalg.norm(Ai) / (k * eps)
        ro = np.linalg.norm(Zi.conj().T @ Zi - np.eye(k)) / (k * eps)
        worst_b, worst_o = max(worst_b, rb), max(worst_o, ro)

    out = {
        "metric": "Schur decomps/s (batched n<=128) + n=4096 time vs plain GPU, bwd err <=50*n*eps",
        "value": value,
        "unit": "matrices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"batched complex Schur (T, Z, values): {args.count} ComplexF64 Ginibre matrices, "
                        "entries N(0,1/2) per component, n ~ U{8..96}, seeded Philox",
            "count": args.count,
            "n_range": [8, 96],
            "seed": args.seed,
            "wantZ": True,
            "parallelism": "single GPU, one workgroup per matrix",
        },
        "roofline": {
            "bound": "mfma",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "kernel": "gs::zqr_kernel<64, true> (QR sweeps + Z, one launch per size bin)",
            "launches": len(qr),
            "avg_launch_us": qr_us / max(len(qr), 1),
            "flops_per_launch": qr_flops / max(len(qr), 1),
            "note": "FP64 vector-ALU work (the bulge chase has no dense contraction; FP64 vector and FP64 MFMA "
                    "peaks are both 78.6 TFLOP/s). achieved = model flops per launch (81.3 n^3 per complex "
                    "matrix) / average launch duration (HIP events on the launch's stream, instrumented pass)",
            "traffic_source": traffic["source"] if traffic else None,
            "hq_kernel_avg_launch_us": sum(r["us"] for r in hq) / max(len(hq), 1),
            "batch_ms": kern_s * 1e3,
            "batch_tflops": batch_rate / 1e12,
            "batch_frac": batch_rate / PEAK_FP64,
            "batch_hbm_frac": bytes_alg / kern_s / PEAK_HBM,
        },
        "check": {
            "sample": int(len(samp)),
            "max_backward_over_n_eps": worst_b,
            "max_orth_over_n_eps": worst_o,
            "bound_over_n_eps": 50.0,
            "info_nonzero": int((info_h != 0).sum()),
            "sweeps": stats[0] if stats else None,
            "bulge_steps": stats[1] if len(stats) > 1 else None,
        },
    }

    if not args.no_dropin and args.dropin_sample > 0:
        m = min(args.dropin_sample, len(mats))
        gs.gschur(mats[0])
        td = time.perf_counter()
        for i in range(m):
            gs.gschur(mats[i])
        dt = time.perf_counter() - td
        rate = m / dt
        out["dropin"] = {"value": rate, "unit": "matrices/s", "sample": f"first {m} matrices, one host-memory "
                         "gs_schur_z call each", "batched_over_dropin": value / rate}

    if not args.no_cpu and args.cpu_sample > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        m = min(args.cpu_sample, len(mats))
        sub = mats[:m]
        sn, so, sf = gs.pack(sub, np.complex128)
        cores = min(16, os.cpu_count() or 1)
        tc = time.perf_counter()
        _, _, _, cinfo = pyoracle.schur_batched(sn, so, sf, True, nthreads=cores)
        dt = time.perf_counter() - tc
        out["cpu_baseline"] = {"value": m / dt, "unit": "matrices/s", "cores": cores, "kind": "port",
                               "sample": f"first {m} matrices of the same workload; C++ restatement of gschur! "
                                         f"(oracle/), OpenMP over matrices, {cores} threads, {dt:.1f}s"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
