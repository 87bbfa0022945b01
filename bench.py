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
                HIP events around every launch in one instrumented (eager,
                graph-free) pass after the timed steps that runs each launch
                alone, in order, on one stream,
                against the 78.6 TFLOP/s FP64 peak; HBM traffic per launch
                from rocprofv3 FETCH_SIZE/WRITE_SIZE runs when a summary of
                them is present in profiles/;
  cpu_baseline  the CPU oracle (a C++ restatement of GenericSchur's gschur!)
                on a bounded sample of the same workload, OpenMP over
                matrices on the host cores;
  dropin        per-matrix host-memory gs_schur_z calls on the same GPU (the
                drop-in schur! path) on a sample, and the batched/drop-in ratio;
  check         acceptance residuals on a sample of the batch;
  single        the metric's single-matrix case: one 256 x 256 ComplexF64
                Ginibre matrix, seconds per solve and its CPU oracle time;
  large         the metric's n = 4096 half: one 4096 x 4096 Float64 solve on
                the large path against the plain unblocked kernel (timed at
                n = 1024, scaled by n^3), its GEMM-phase fraction of the FP64
                peak, residuals;
  ok            false (and exit status 1) when a residual exceeds 50 n eps or
                an info is nonzero.
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


def _tuning_note():
    """Which launch parameters the batched calls ran with."""
    from genericschur import tuning
    t = tuning.load()
    if not t:
        return "built-in defaults"
    return (f"autotuned table {os.path.relpath(os.environ.get('GSCHUR_TUNING') or tuning.DEFAULT_PATH, ROOT)} "
            f"({t.get('date', '?')}; complex {json.dumps(t.get('complex', {}), sort_keys=True)}; "
            f"real {json.dumps(t.get('real', {}), sort_keys=True)})")


def load_traffic():
    """HBM bytes per size bin's QR phase (QR-sweep launch + Z-log launch) from
    the committed rocprofv3 summary, if any."""
    path = os.path.join(ROOT, "profiles", "traffic_qrphase_workload2.json")
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
    ap.add_argument("--workload", type=int, default=2, choices=[2, 3, 4, 5, 6, 7, 8],
                    help="2: batched complex (the headline line); 3: batched real mix; 4: one large real matrix "
                         "(blocked path vs plain kernel); 5: eigenvalue-loss forward + backward; 6: batched "
                         "Hermitian eigensystems (QR iteration); 7: one large Hermitian eigensystem; 8: the "
                         "eigenvectors of one large real Schur factorization")
    ap.add_argument("--herm-big-n", type=int, default=2048, help="order of the Hermitian matrix of workload 7")
    ap.add_argument("--vec-n", type=int, default=4096, help="order of the Schur factorization of workload 8")
    ap.add_argument("--herm-n", type=int, default=64, help="order of the Hermitian matrices of workload 6")
    ap.add_argument("--herm-real", action="store_true", help="workload 6 on real symmetric matrices")
    ap.add_argument("--large-n", type=int, default=4096)
    ap.add_argument("--plain-n", type=int, default=1024,
                    help="order at which the plain kernel is timed (scaled by n^3 to --large-n)")
    ap.add_argument("--large-cpu-n", type=int, default=512,
                    help="order at which the oracle is timed for the large matrix (scaled by n^3)")
    ap.add_argument("--no-large", action="store_true", help="leave the large-matrix object out of the default line")
    ap.add_argument("--no-single", action="store_true",
                    help="leave the single 256 x 256 complex object out of the default line")
    ap.add_argument("--single-n", type=int, default=256)
    ap.add_argument("--single-reps", type=int, default=5)
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.workload == 3:
        return bench_workload3(args)
    if args.workload == 4:
        return bench_workload4(args)
    if args.workload == 5:
        return bench_workload5(args)
    if args.workload == 6:
        return bench_workload6(args)
    if args.workload == 7:
        return bench_workload7(args)
    if args.workload == 8:
        return bench_workload8(args)
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
    # one instrumented pass (not part of the timed steps): every launch alone,
    # in order on one stream, between two events
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
    # dominant kernel: the QR phase of each size bin, the QR-sweep launch
    # plus (the bin above n = 80) the launch applying its Z log
    qr = [r for r in recs if r["kind"] == "qr"]
    zl = [r for r in recs if r["kind"] == "zlog"]
    hq = [r for r in recs if r["kind"] == "hq"]
    # fewer QR records than size bins: the work-queue launch of the tuned
    # configuration (one for the batch); else one per size bin
    queue = 0 < len(qr) < len(hq)
    qr_flops = (100.0 - 56.0 / 3.0) * sum(r["sum_n3"] for r in qr)
    qr_us = sum(r["us"] for r in qr) + sum(r["us"] for r in zl)
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
            "tuning": _tuning_note(),
        },
        "roofline": {
            "bound": "mfma",
            "engine": "valu-fp64",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP64 / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64,
            "traffic": traffic["bytes_per_launch"] if traffic else None,
            "kernel": ("QR phase of the batch: gs::zqrw2_queue_kernel<true> (two-wavefront windowed QR sweeps, "
                       "one work-queue launch over all size bins, largest matrix first) then gs::zlog_kernel "
                       "(the Z log of the largest matrices applied with Z in LDS, one launch per logging bin)"
                       if queue else
                       "QR phase of a size bin: gs::zqrw2_kernel<true> (two-wavefront windowed QR sweeps) "
                       "then, for the bin above n = 80, gs::zlog_kernel (the sweeps' Z log applied with Z in LDS)"),
            "launches": len(qr),
            "avg_launch_us": qr_us / max(len(qr), 1),
            "qr_sweep_avg_launch_us": sum(r["us"] for r in qr) / max(len(qr), 1),
            "zlog_launches": len(zl),
            "zlog_avg_launch_us": sum(r["us"] for r in zl) / max(len(zl), 1),
            "flops_per_launch": qr_flops / max(len(qr), 1),
            "note": "FP64 vector-ALU work (engine valu-fp64: the bulge chase has no dense contraction); the compute "
                    "roof is the FP64 peak, 78.6 TFLOP/s for the vector ALUs and for MFMA alike. achieved = model "
                    "flops of the QR phase (81.3 n^3 per complex matrix: H and Z updates; of the whole batch for "
                    "the work-queue launch, of a bin otherwise) / its duration (HIP events around the QR-sweep "
                    "and Z-log launches, instrumented pass with every launch alone)",
            "traffic_kind": "calibrated (committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary, not measured in this "
                            "run)" if traffic else None,
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
        out["cpu_baseline"] = _cpu_batched(mats, True, min(args.cpu_sample, len(mats)))
    # the metric's other configurations on the same GPU, after the timed steps
    if not args.no_single:
        out["single"] = single_object(args, with_cpu=not args.no_cpu)
    if not args.no_large:
        out["large"] = large_object(args, with_cpu=not args.no_cpu)[0]
    ok = (out["check"]["info_nonzero"] == 0 and worst_b <= 50.0 and worst_o <= 50.0)
    ok = ok and all(out[k]["ok"] for k in ("single", "large") if k in out)
    out["ok"] = bool(ok)
    print(json.dumps(out), flush=True)
    if not ok:
        log("acceptance check failed: residual above 50 n eps or info != 0 (see check / single / large)")
        sys.exit(1)


def _common(args, metric_value, unit, ms, workload, extra_cfg, dtype="f64"):
    cfg = {"workload": workload}
    cfg.update(extra_cfg)
    return {"metric": "Schur decomps/s (batched n<=128) + n=4096 time vs plain GPU, bwd err <=50*n*eps",
            "value": metric_value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic", "config": cfg}


def _timed(stream, step, steps, warmup):
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t1 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    return wall, ev0.elapsed_time(ev1) * 1e-3 / steps


def _residuals(mats, ns, off, Th, Zh, sample=128):
    eps = np.finfo(np.float64).eps
    idx = np.linspace(0, len(ns) - 1, min(sample, len(ns))).astype(int)
    wb = wo = 0.0
    for i in idx:
        k, o = int(ns[i]), int(off[i])
        Ai = mats[i]
        Ti = Th[o:o + k * k].reshape((k, k), order="F")
        Zi = Zh[o:o + k * k].reshape((k, k), order="F")
        wb = max(wb, np.linalg.norm(Ai - Zi @ Ti @ Zi.conj().T) / np.linalg.norm(Ai) / (k * eps))
        wo = max(wo, np.linalg.norm(Zi.conj().T @ Zi - np.eye(k)) / (k * eps))
    return {"sample": int(len(idx)), "max_backward_over_n_eps": wb, "max_orth_over_n_eps": wo,
            "bound_over_n_eps": 50.0}


def bench_workload3(args):
    """Batched real Schur of 4096 matrices n ~ U{16..128}: half real Ginibre,
    a quarter Grcar-type Toeplitz, a quarter Q (D + N) Q^T."""
    import genericschur as gs
    from genericschur import gen, _lib
    import torch
    count = 4096 if args.count == 8192 else args.count
    mats, _ = gen.workload3(seed=args.seed, count=count)
    ns, off, flat = gs.pack(mats, np.float64)
    dev = torch.device("cuda:0")
    A = torch.from_numpy(flat).to(dev)
    T = torch.zeros_like(A)
    Z = torch.zeros_like(A)
    W = torch.zeros(2 * int(ns.sum()), dtype=torch.float64, device=dev)
    info = torch.zeros(len(ns), dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    gs.set_stream(stream.cuda_stream)
    step = lambda: gs.gschur_batched_device(ns, off, A, T, Z, W, info, complex_input=False, asynchronous=True)
    wall, kern_s = _timed(stream, step, args.steps, args.warmup)
    gs.set_stream(None)
    n64 = ns.astype(np.float64)
    out = _common(args, count * args.steps / wall, "matrices/s", 1e3 * wall / args.steps,
                  f"batched real Schur (T, Z, values): {count} Float64 matrices n ~ U{{16..128}}, 1/2 Ginibre, "
                  "1/4 Toeplitz (-1 sub, +1 diag and 3 super), 1/4 Q(D+N)Q^T; seeded Philox",
                  {"count": count, "n_range": [16, 128], "seed": args.seed})
    flops = float((25.0 * n64 ** 3).sum())
    out["roofline"] = {"bound": "mfma", "engine": "valu-fp64", "achieved": flops / kern_s / 1e12, "peak": 78.6,
                       "unit": "TFLOP/s", "frac": flops / kern_s / PEAK_FP64, "traffic": None,
                       "note": "25 n^3 model flops per matrix over the batch's device time"}
    out["check"] = _residuals(mats, ns, off, T.cpu().numpy(), Z.cpu().numpy())
    out["check"]["info_nonzero"] = int((info.cpu().numpy() != 0).sum())
    if not args.no_cpu:
        out["cpu_baseline"] = _cpu_batched(mats, False, min(args.cpu_sample, count))
    c = out["check"]
    out["ok"] = bool(c["info_nonzero"] == 0 and c["max_backward_over_n_eps"] <= 50.0
                     and c["max_orth_over_n_eps"] <= 50.0)
    print(json.dumps(out), flush=True)
    if not out["ok"]:
        sys.exit(1)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _cpu_batched(mats, cplx, m):
    """The oracle (C++ restatement of gschur!) over the first m matrices,
    OpenMP over matrices on up to 16 host cores, and single-threaded over
    the first m/16 of them (the reference is a single-thread algorithm)."""
    import genericschur as gs
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    dt_ = np.complex128 if cplx else np.float64
    sn, so, sf = gs.pack(mats[:m], dt_)
    cores = min(16, os.cpu_count() or 1)
    tc = time.perf_counter()
    pyoracle.schur_batched(sn, so, sf, cplx, nthreads=cores)
    dt = time.perf_counter() - tc
    m1 = max(1, m // 16)
    sn1, so1, sf1 = gs.pack(mats[:m1], dt_)
    tc = time.perf_counter()
    pyoracle.schur_batched(sn1, so1, sf1, cplx, nthreads=1)
    dt1 = time.perf_counter() - tc
    return {"value": m / dt, "unit": "matrices/s", "cores": cores, "kind": "port",
            "sample": f"first {m} matrices of the same workload; C++ restatement of gschur! (oracle/), OpenMP over "
                      f"matrices, {cores} threads, {dt:.1f}s",
            "one_thread_value": m1 / dt1, "one_thread_sample": f"first {m1} matrices, 1 thread, {dt1:.1f}s",
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "omp_num_threads": os.environ.get("OMP_NUM_THREADS")}


def large_object(args, with_cpu=True):
    """One large real matrix A = Q (Lambda + U) Q^T of order args.large_n on
    the large path (blocked Hessenberg, chunked QR), against the plain
    unblocked kernel timed at args.plain_n and scaled by n^3; backward and
    orthogonality residuals computed on the device."""
    import genericschur as gs
    from genericschur import gen, _lib
    import torch
    dev = torch.device("cuda:0")

    def solve(n, plain, timing):
        A = gen.workload4(seed=args.seed, n=n)
        Ad = torch.from_numpy(np.asfortranarray(A).ravel(order="F")).to(dev)
        T = torch.zeros_like(Ad)
        Z = torch.zeros_like(Ad)
        W = torch.zeros(2 * n, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = gs.gschur_device(n, Ad, T, Z, W, False, plain=plain, timing=timing)
        gs.synchronize()
        dt = time.perf_counter() - t0
        Am, Tm, Zm = (x.view(n, n).t() for x in (Ad, T, Z))
        eps = np.finfo(np.float64).eps
        bwd = (torch.linalg.norm(Am - Zm @ Tm @ Zm.t()) / torch.linalg.norm(Am)).item() / (n * eps)
        orth = torch.linalg.norm(Zm.t() @ Zm - torch.eye(n, dtype=torch.float64, device=dev)).item() / (n * eps)
        return dt, info, bwd, orth

    n = args.large_n
    log(f"large path n={n}")
    dt, info, bwd, orth = solve(n, False, True)
    recs = _lib.last_launches()
    st = _lib.last_stats()
    ph = {int(r["phase"]): r["us"] * 1e-6 for r in recs if r["kind"] == "phase"}
    g = [r for r in recs if r["kind"] == "gemm"]
    gf = sum(r["flops"] for r in g)
    gt = sum(r["us"] for r in g) * 1e-6
    log(f"plain kernel n={args.plain_n}")
    pdt, pinfo, pbwd, porth = solve(args.plain_n, True, False)
    plain_est = pdt * (n / args.plain_n) ** 3
    flops = 25.0 * n ** 3
    out = {"workload": f"one {n}x{n} Float64 matrix Q(Lambda+U)Q^T (Haar Q, log-uniform real and complex-pair "
                       "eigenvalues in [1e-3, 1e3], U strictly upper N(0,1)), T, Z and values",
           "n": n, "seconds": dt, "info": info, "sweeps": st[0] if st else None,
           "bulge_steps": st[1] if len(st) > 1 else None,
           "phase_seconds": {"copy_scale": ph.get(0), "hessenberg": ph.get(1), "q_formation": ph.get(2),
                             "qr": ph.get(3), "finish": ph.get(4)},
           "gemm_phase": {"tflops": gf / gt / 1e12 if gt else None,
                          "frac_fp64_peak": gf / gt / PEAK_FP64 if gt else None,
                          "seconds": gt, "gemms": len(g), "flops": gf},
           "whole_solve_tflops": flops / dt / 1e12, "whole_solve_frac": flops / dt / PEAK_FP64,
           "backward_over_n_eps": bwd, "orth_over_n_eps": orth,
           "plain": {"n": args.plain_n, "seconds": pdt, "info": pinfo, "backward_over_n_eps": pbwd,
                     "orth_over_n_eps": porth, "seconds_scaled_to_n": plain_est,
                     "note": f"plain unblocked kernel timed at n={args.plain_n}, scaled by (n/{args.plain_n})^3"},
           "speedup_over_plain": plain_est / dt,
           "ok": bool(info == 0 and bwd <= 50.0 and orth <= 50.0)}
    if with_cpu:
        # the oracle single-threaded (the reference's gschur! is scalar code)
        # at a smaller order of the same construction, scaled by n^3
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        nc = min(args.large_cpu_n, n)
        Ac = gen.workload4(seed=args.seed, n=nc)
        tc = time.perf_counter()
        pyoracle.gschur(Ac)
        cdt = time.perf_counter() - tc
        est = cdt * (n / nc) ** 3
        out["cpu_baseline"] = {"value": 1.0 / est, "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"oracle (C++ restatement of gschur!) on one {nc}x{nc} matrix of the same "
                                         f"construction, 1 thread, {cdt:.1f}s, scaled by ({n}/{nc})^3 = "
                                         f"{est:.0f}s per solve"}
    return out, gf, gt


def bench_workload4(args):
    """One large real matrix A = Q (Lambda + U) Q^T: the blocked large-matrix
    path against the plain unblocked kernel."""
    lg, gf, gt = large_object(args, with_cpu=not args.no_cpu)
    n = lg["n"]
    out = _common(args, 1.0 / lg["seconds"], "solves/s", lg["seconds"] * 1e3, lg["workload"],
                  {"n": n, "seed": args.seed})
    out["steps"] = 1
    out["warmup"] = 0
    out["large"] = lg
    if "cpu_baseline" in lg:
        out["cpu_baseline"] = lg.pop("cpu_baseline")
    out["roofline"] = {"bound": "mfma", "engine": "mfma-fp64", "achieved": gf / gt / 1e12 if gt else None,
                       "peak": 78.6, "unit": "TFLOP/s", "frac": gf / gt / PEAK_FP64 if gt else None,
                       "traffic": None,
                       "kernel": "gs::dgemm_kernel (GEMM phases: blocked Hessenberg and Q formation)",
                       "whole_solve_tflops": lg["whole_solve_tflops"]}
    out["ok"] = lg["ok"]
    print(json.dumps(out), flush=True)
    if not out["ok"]:
        sys.exit(1)


def single_object(args, with_cpu=True):
    """One n x n ComplexF64 Ginibre matrix (the reference's own CPU-runnable
    case, n = 256): gschur! on device-resident data, median of repeated
    solves, against the oracle single-threaded on the same matrix."""
    import genericschur as gs
    from genericschur import gen, _lib
    import torch
    dev = torch.device("cuda:0")
    n = args.single_n
    A = gen.workload1(seed=args.seed, n=n)
    a = np.asfortranarray(A).ravel(order="F").view(np.float64)
    Ad = torch.from_numpy(a).to(dev)
    T = torch.zeros_like(Ad)
    Z = torch.zeros_like(Ad)
    W = torch.zeros(2 * n, dtype=torch.float64, device=dev)
    A0 = Ad.clone()
    times = []
    for r in range(1 + args.single_reps):
        Ad.copy_(A0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        info = gs.gschur_device(n, Ad, T, Z, W, True)
        gs.synchronize()
        if r:
            times.append(time.perf_counter() - t0)
    st = _lib.last_stats()
    dt = float(np.median(times))
    Th = T.cpu().numpy().view(np.complex128).reshape((n, n), order="F")
    Zh = Z.cpu().numpy().view(np.complex128).reshape((n, n), order="F")
    eps = np.finfo(np.float64).eps
    bwd = np.linalg.norm(A - Zh @ Th @ Zh.conj().T) / np.linalg.norm(A) / (n * eps)
    orth = np.linalg.norm(Zh.conj().T @ Zh - np.eye(n)) / (n * eps)
    flops = 100.0 * n ** 3
    out = {"workload": f"one {n}x{n} ComplexF64 Ginibre matrix, Re, Im ~ N(0, 1/2), seeded Philox; T, Z, values",
           "n": n, "seconds": dt, "solves_per_s": 1.0 / dt, "reps": args.single_reps, "info": info,
           "sweeps": st[0] if st else None, "bulge_steps": st[1] if len(st) > 1 else None,
           "tflops": flops / dt / 1e12, "frac_fp64_peak": flops / dt / PEAK_FP64,
           "backward_over_n_eps": float(bwd), "orth_over_n_eps": float(orth),
           "ok": bool(info == 0 and bwd <= 50.0 and orth <= 50.0)}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        tc = time.perf_counter()
        pyoracle.gschur(A)
        cdt = time.perf_counter() - tc
        out["cpu_baseline"] = {"value": 1.0 / cdt, "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"oracle (C++ restatement of gschur!) on the same matrix, 1 thread, "
                                         f"{cdt:.2f}s", "speedup_gpu_over_cpu": cdt / dt}
    return out


def bench_workload5(args):
    """Eigenvalue loss L = sum w Re(lambda) + v Im(lambda)^2 over 2048 real
    Ginibre matrices n ~ U{32..64}: batched Schur forward + VJP backward."""
    import ctypes
    import genericschur as gs
    from genericschur import gen, _lib
    import torch
    count = 2048 if args.count == 8192 else args.count
    mats, ws, vs = gen.workload5(seed=args.seed, count=count)
    ns, off, flat = gs.pack(mats, np.float64)
    dev = torch.device("cuda:0")
    A = torch.from_numpy(flat).to(dev)
    T = torch.zeros_like(A)
    Z = torch.zeros_like(A)
    Abar = torch.zeros_like(A)
    nw = int(ns.sum())
    W = torch.zeros(2 * nw, dtype=torch.float64, device=dev)
    info = torch.zeros(len(ns), dtype=torch.int32, device=dev)
    vinfo = torch.zeros(len(ns), dtype=torch.int32, device=dev)
    wt = torch.from_numpy(np.concatenate(ws)).to(dev)
    vt = torch.from_numpy(np.concatenate(vs)).to(dev)
    g = torch.zeros(2 * nw, dtype=torch.float64, device=dev)
    ns_c = np.ascontiguousarray(ns, dtype=np.int32)
    off_c = np.ascontiguousarray(off, dtype=np.int64)
    stream = torch.cuda.Stream(device=dev)
    gs.set_stream(stream.cuda_stream)
    h = _lib.handle()
    a = lambda t: ctypes.c_void_p(t.data_ptr())

    def step():
        gs.gschur_batched_device(ns, off, A, T, Z, W, info, complex_input=False, asynchronous=True)
        with torch.cuda.stream(stream):
            # gbar = dL/dRe + i dL/dIm = w + 2 i v Im(lambda)
            g[0::2] = wt
            g[1::2] = 2.0 * vt * W[1::2]
        rc = _lib.lib().gs_eig_vjp_batched_d(h, len(ns), ctypes.c_void_p(ns_c.ctypes.data),
                                             ctypes.c_void_p(off_c.ctypes.data), a(T), a(Z), a(g), a(Abar), a(vinfo),
                                             _lib.GS_DEVICE_PTRS | _lib.GS_ASYNC)
        _lib.check(rc, h, "gs_eig_vjp_batched_d")

    wall, kern_s = _timed(stream, step, args.steps, args.warmup)
    gs.set_stream(None)
    n64 = ns.astype(np.float64)
    out = _common(args, count * args.steps / wall, "matrices/s (forward + backward)", 1e3 * wall / args.steps,
                  f"eigenvalue loss sum w Re(lambda) + v Im(lambda)^2 over {count} Float64 Ginibre matrices "
                  "n ~ U{32..64} (min eigenvalue separation >= 1e-3): batched Schur + VJP kernel",
                  {"count": count, "n_range": [32, 64], "seed": args.seed})
    flops = float(((25.0 + 2 * 8.0 / 3.0 + 8.0) * n64 ** 3).sum())
    out["roofline"] = {"bound": "mfma", "engine": "valu-fp64", "achieved": flops / kern_s / 1e12, "peak": 78.6,
                       "unit": "TFLOP/s", "frac": flops / kern_s / PEAK_FP64, "traffic": None,
                       "note": "(25 + 16/3 + 8) n^3 model flops per matrix (Schur, two eigenvector sides, "
                               "assembly) over the device time"}
    out["check"] = _residuals(mats, ns, off, T.cpu().numpy(), Z.cpu().numpy())
    out["check"]["info_nonzero"] = int((info.cpu().numpy() != 0).sum())
    out["check"]["vjp_info_nonzero"] = int((vinfo.cpu().numpy() != 0).sum())
    c = out["check"]
    out["ok"] = bool(c["info_nonzero"] == 0 and c["vjp_info_nonzero"] == 0 and c["max_backward_over_n_eps"] <= 50.0
                     and c["max_orth_over_n_eps"] <= 50.0)
    if not args.no_cpu:
        out["cpu_baseline"] = _cpu_batched(mats, False, count)
        out["cpu_baseline"]["note"] = "forward (Schur) only: the reference has no eigenvalue rrule"
    print(json.dumps(out), flush=True)
    if not out["ok"]:
        sys.exit(1)


def bench_workload7(args):
    """One large Hermitian eigensystem: geigen!(Hermitian(A)) with
    QRIteration, values ascending and vectors, of one ComplexF64 G + G^H of
    order --herm-big-n (--herm-real: Float64), device-resident; a step is one
    solve (at most 3 timed)."""
    import genericschur as gs
    from genericschur import _lib
    import torch
    n = args.herm_big_n
    cplx = not args.herm_real
    rng = np.random.default_rng(args.seed)
    G = rng.standard_normal((n, n)) + (1j * rng.standard_normal((n, n)) if cplx else 0)
    Ah = G + G.conj().T
    dt_ = np.complex128 if cplx else np.float64
    dev = torch.device("cuda:0")
    A = torch.from_numpy(np.asfortranarray(Ah, dtype=dt_).reshape(-1, order="F").view(np.float64)).to(dev)
    V = torch.zeros_like(A)
    W = torch.zeros(n, dtype=torch.float64, device=dev)
    info = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    gs.set_stream(stream.cuda_stream)
    h = _lib.handle()
    f = _lib.lib().gs_heev_z if cplx else _lib.lib().gs_heev_d
    a = lambda t: ctypes.c_void_p(t.data_ptr())
    flags = _lib.GS_WANTZ | _lib.GS_SORT | _lib.GS_DEVICE_PTRS

    def step():
        rc = f(h, 1, n, a(A), n, n * n, 0, a(W), a(V), n, n * n, a(info), flags)
        _lib.check(rc, h, "gs_heev")

    steps = max(1, min(args.steps, 3))
    wall, kern_s = _timed(stream, step, steps, min(args.warmup, 1))
    gs.set_stream(None)
    kind = "ComplexF64 Hermitian" if cplx else "Float64 symmetric"
    out = _common(args, steps / wall, "solves/s", 1e3 * wall / steps,
                  f"one Hermitian eigensystem geigen!(Hermitian(A), QRIteration()) (values ascending and vectors): "
                  f"one {kind} matrix G + G^H of order {n}, G Gaussian; seeded numpy",
                  {"n": n, "complex": cplx, "seed": args.seed})
    out["metric"] = "Hermitian eigensystem solves/s (one large matrix, QR iteration)"
    out["steps"] = steps
    Wh = W.cpu().numpy()
    Vh = V.cpu().numpy().view(dt_).reshape(n, n).T
    eps = np.finfo(np.float64).eps
    idx = np.linspace(0, n - 1, 32).astype(int)
    res = np.linalg.norm(Ah @ Vh[:, idx] - Vh[:, idx] * Wh[idx]) / np.linalg.norm(Ah) / (n * eps)
    orth = np.abs(Vh[:, idx].conj().T @ Vh - np.eye(n)[idx]).max() / (n * eps)
    out["check"] = {"columns": 32, "residual_over_n_eps": float(res), "orth_over_n_eps": float(orth),
                    "info": int(info.cpu().item()), "sorted": bool(np.all(np.diff(Wh) >= 0))}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        nc = min(512, n)
        Gc = rng.standard_normal((nc, nc)) + (1j * rng.standard_normal((nc, nc)) if cplx else 0)
        tc = time.perf_counter()
        pyoracle.heev(Gc + Gc.conj().T, "L")
        dtc = time.perf_counter() - tc
        est = dtc * (n / nc) ** 3
        out["cpu_baseline"] = {"value": 1.0 / est, "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"oracle geigen! (C++ restatement) on one {nc}x{nc} matrix of the same kind, "
                                         f"1 thread, {dtc:.2f}s, scaled by ({n}/{nc})^3 = {est:.1f}s per solve"}
    c = out["check"]
    out["ok"] = bool(c["info"] == 0 and c["sorted"] and c["residual_over_n_eps"] <= 50.0
                     and c["orth_over_n_eps"] <= 50.0)
    print(json.dumps(out), flush=True)
    if not out["ok"]:
        sys.exit(1)


def bench_workload8(args):
    """The eigenvectors of one large real Schur factorization: T, Z of the
    large-matrix workload's construction at order --vec-n from the large path,
    then geigvecs right and left (gs_trevc_d), device-resident; a step is the
    right and the left vectors (at most 3 timed)."""
    import genericschur as gs
    from genericschur import _lib, gen
    import torch
    n = args.vec_n
    dev = torch.device("cuda:0")
    A = gen.workload4(seed=args.seed, n=n)
    Ad = torch.from_numpy(np.asfortranarray(A).ravel(order="F")).to(dev)
    T = torch.zeros_like(Ad)
    Z = torch.zeros_like(Ad)
    W = torch.zeros(2 * n, dtype=torch.float64, device=dev)
    info = gs.gschur_device(n, Ad, T, Z, W, False)
    gs.synchronize()
    VR = torch.zeros(2 * n * n, dtype=torch.float64, device=dev)
    VL = torch.zeros_like(VR)
    stream = torch.cuda.Stream(device=dev)
    gs.set_stream(stream.cuda_stream)
    h = _lib.handle()
    f = _lib.lib().gs_trevc_d
    a = lambda t: ctypes.c_void_p(t.data_ptr())

    def step():
        for left, V in ((0, VR), (1, VL)):
            rc = f(h, n, a(T), n, a(Z), n, a(V), n, left, 1, _lib.GS_DEVICE_PTRS | _lib.GS_ASYNC)
            _lib.check(rc, h, "gs_trevc_d")

    steps = max(1, min(args.steps, 3))
    wall, kern_s = _timed(stream, step, steps, min(args.warmup, 1))
    gs.set_stream(None)
    out = _common(args, steps / wall, "solves/s", 1e3 * wall / steps,
                  f"geigvecs (right and left, normalised) of one real Schur factorization of order {n}: T, Z of "
                  f"Q (Lambda + U) Q^T from the large path", {"n": n, "seed": args.seed})
    out["metric"] = "eigenvector sets/s (right + left, one large real Schur factorization)"
    out["steps"] = steps
    # residual of a sample of right vectors against A
    eps = np.finfo(np.float64).eps
    w = W.cpu().numpy().view(np.complex128)
    Vr = VR.cpu().numpy().view(np.complex128).reshape(n, n).T
    idx = np.linspace(0, n - 1, 32).astype(int)
    res = np.linalg.norm(A @ Vr[:, idx] - Vr[:, idx] * w[idx]) / np.linalg.norm(A) / (n * eps)
    out["check"] = {"columns": 32, "right_residual_over_n_eps": float(res), "schur_info": int(info)}
    out["ok"] = bool(info == 0 and np.isfinite(res))
    print(json.dumps(out), flush=True)
    if not out["ok"]:
        sys.exit(1)


def bench_workload6(args):
    """Batched Hermitian eigensystems: geigen!(Hermitian(A)) with QRIteration
    (eigenvalues ascending and eigenvectors) of 4096 ComplexF64 Hermitian
    matrices of order 64 (--herm-real: Float64 symmetric), one launch."""
    import genericschur as gs
    from genericschur import _lib
    import torch
    count = 4096 if args.count == 8192 else args.count
    n = args.herm_n
    cplx = not args.herm_real
    rng = np.random.default_rng(args.seed)
    G = rng.standard_normal((count, n, n))
    if cplx:
        G = G + 1j * rng.standard_normal((count, n, n))
    mats = G + np.conj(np.transpose(G, (0, 2, 1)))
    dt_ = np.complex128 if cplx else np.float64
    flat = np.ascontiguousarray(np.transpose(mats, (0, 2, 1)), dtype=dt_)  # column-major per matrix
    dev = torch.device("cuda:0")
    A = torch.from_numpy(flat.reshape(-1).view(np.float64)).to(dev)
    V = torch.zeros_like(A)
    W = torch.zeros(count * n, dtype=torch.float64, device=dev)
    info = torch.zeros(count, dtype=torch.int32, device=dev)
    stream = torch.cuda.Stream(device=dev)
    gs.set_stream(stream.cuda_stream)
    h = _lib.handle()
    f = _lib.lib().gs_heev_z if cplx else _lib.lib().gs_heev_d
    a = lambda t: ctypes.c_void_p(t.data_ptr())
    flags = _lib.GS_WANTZ | _lib.GS_SORT | _lib.GS_DEVICE_PTRS | _lib.GS_ASYNC

    def step():
        rc = f(h, count, n, a(A), n, n * n, 0, a(W), a(V), n, n * n, a(info), flags)
        _lib.check(rc, h, "gs_heev")

    wall, kern_s = _timed(stream, step, args.steps, args.warmup)
    gs.set_stream(None)
    kind = "ComplexF64 Hermitian" if cplx else "Float64 symmetric"
    out = _common(args, count * args.steps / wall, "matrices/s", 1e3 * wall / args.steps,
                  f"batched Hermitian eigensystems geigen!(Hermitian(A), QRIteration()) (values ascending and "
                  f"vectors): {count} {kind} matrices G + G^H of order {n}, G Gaussian; seeded numpy",
                  {"count": count, "n": n, "complex": cplx, "seed": args.seed})
    out["metric"] = "Hermitian eigensystems/s (batched, QR iteration)"
    # model flops per matrix: tridiagonalization and Q formation (xHETRD +
    # xUNGTR: 16/3 n^3 each complex, 4/3 real) and about n^2 rotations of
    # 12 n (complex Z) / 6 n (real Z) flops
    fl = (16.0 / 3 + 16.0 / 3 + 12.0) if cplx else (4.0 / 3 + 4.0 / 3 + 6.0)
    flops = fl * n ** 3 * count
    out["roofline"] = {"bound": "mfma", "engine": "valu-fp64", "achieved": flops / kern_s / 1e12, "peak": 78.6,
                       "unit": "TFLOP/s", "frac": flops / kern_s / PEAK_FP64, "traffic": None,
                       "note": f"{fl:.2f} n^3 model flops per matrix over the device time of the launch"}
    Wh = W.cpu().numpy().reshape(count, n)
    Vh = V.cpu().numpy().view(dt_).reshape(count, n, n).transpose(0, 2, 1)
    eps = np.finfo(np.float64).eps
    rb = ro = 0.0
    for k in np.linspace(0, count - 1, min(64, count)).astype(int):
        Ak, Vk, wk = mats[k], Vh[k], Wh[k]
        rb = max(rb, np.linalg.norm(Ak @ Vk - Vk * wk) / np.linalg.norm(Ak) / (n * eps))
        ro = max(ro, np.linalg.norm(Vk.conj().T @ Vk - np.eye(n)) / (n * eps))
    out["check"] = {"sample": int(min(64, count)), "max_residual_over_n_eps": rb, "max_orth_over_n_eps": ro,
                    "info_nonzero": int((info.cpu().numpy() != 0).sum()), "bound_over_n_eps": 50.0}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        m = min(args.cpu_sample, count)
        cores = min(16, os.cpu_count() or 1)
        tc = time.perf_counter()
        pyoracle.heev_batched(mats[:m], nthreads=cores)
        dtc = time.perf_counter() - tc
        m1 = max(1, m // 16)
        tc = time.perf_counter()
        pyoracle.heev_batched(mats[:m1], nthreads=1)
        dt1 = time.perf_counter() - tc
        out["cpu_baseline"] = {"value": m / dtc, "unit": "matrices/s", "cores": cores, "kind": "port",
                               "sample": f"first {m} matrices; C++ restatement of geigen! (oracle/), OpenMP over "
                                         f"matrices, {cores} threads, {dtc:.1f}s",
                               "one_thread_value": m1 / dt1, "one_thread_sample": f"first {m1} matrices, 1 thread"}
    c = out["check"]
    out["ok"] = bool(c["info_nonzero"] == 0 and c["max_residual_over_n_eps"] <= 50.0
                     and c["max_orth_over_n_eps"] <= 50.0)
    print(json.dumps(out), flush=True)
    if not out["ok"]:
        sys.exit(1)


if __name__ == "__main__":
    main()
