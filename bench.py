#!/usr/bin/env python3
"""Benchmark of the GPU_BATCH reorder path on one MI355X.

Default workload (one JSON line on stdout): 3D 7-point stencil 128^3 (n = 2,097,152, nnz = 14,581,760),
seeded symmetric random relabelling, resident in HBM. A step is one Cuthill-McKee reorder of that matrix
(state-reset kernel + persistent reorder kernel, synchronised), from the pseudo-peripheral start node found
once on the GPU before timing (the reference also excludes the start search from its timed region,
src/device/rcm_gpu_batch.cu:47-99).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--batched COUNT] [--serving SECONDS]

Fields beyond the contract: kernel-only nnz/s, end-to-end time (upload + start search + reorder + bandwidth
+ download), the bandwidth of the result, and, when asked for, the batched and serving forms.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "nnz reordered/sec + % HBM roofline; batched matrices/sec, p50/p99 latency"
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
SEED_C2 = 2


def alg_bytes(n: int, nnz: int) -> int:
    """Compulsory traffic of one reorder: row_ptr and col_idx read once, the permutation written once and
    read once as the frontier: 4(n+1) + 4 nnz + 4n + 4n bytes."""
    return 4 * nnz + 12 * n + 4


def cpu_baseline(timeout_s: int = 240):
    """Reference CPU paths timed on this host, in a child process (the reference's CPU_BATCH is third-party
    code; a crash there must not end the benchmark)."""
    code = r"""
import json, os, sys, time
sys.path.insert(0, %r); sys.path.insert(0, os.path.join(%r, 'tests'))
import numpy as np, oracle_lib as O, parallelbatchrcm_amd as P
m = P.workloads.config(2, seed=%d)
rc, s, it, _ = O.peripheral(m.row_offsets, m.col_ids)
threads = max(1, min(16, os.cpu_count() or 1))
out = {"n": m.rows, "nnz": m.nnz, "threads": threads}
if O.ref_available():
    _, ser = O.ref_rcm(m.row_offsets, m.col_ids, s, threads=1, batchsize=1)
    _, bat = O.ref_rcm(m.row_offsets, m.col_ids, s, threads=threads, batchsize=512)
    out.update(kind="reference", serial_ms=ser, batch_ms=bat)
else:
    ts = []
    for _ in range(5):
        t = time.perf_counter(); O.cm(m.row_offsets, m.col_ids, s); ts.append((time.perf_counter() - t) * 1e3)
    out.update(kind="port", serial_ms=sum(ts) / len(ts), batch_ms=None)
try:
    out["cpu_model"] = [l.split(':', 1)[1].strip() for l in open('/proc/cpuinfo') if l.startswith('model name')][0]
except Exception:
    out["cpu_model"] = "unknown"
print("CPUBASE " + json.dumps(out))
""" % (ROOT, ROOT, SEED_C2)
    try:
        res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return None
    for line in res.stdout.splitlines():
        if line.startswith("CPUBASE "):
            return json.loads(line[8:])
    print(res.stderr[-2000:], file=sys.stderr)
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workers", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched", type=int, default=0, help="also run the batched form on COUNT matrices")
    ap.add_argument("--serving", type=float, default=0.0, help="also run the serving form for SECONDS")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "The batched form shards by matrix; see DESIGN.md.", file=sys.stderr)
        return 2

    import numpy as np

    import parallelbatchrcm_amd as P

    t0 = time.perf_counter()
    m = P.workloads.config(2, seed=SEED_C2)
    gen_s = time.perf_counter() - t0
    n, nnz = m.rows, m.nnz

    # end-to-end once: upload, GPU start search, reorder, bandwidth, download
    e2e0 = time.perf_counter()
    ctx = P.RcmContext(timeout_ms=30000)
    if args.workers:
        ctx.set_param("workers", args.workers)
    ctx.load(m)
    start = ctx.find_start()
    ctx.run(start, P.RCM_FLAG_REVERSE)
    bw = ctx.bandwidth()
    perm = ctx.perm()
    e2e_ms = (time.perf_counter() - e2e0) * 1e3
    tm = ctx.timing()
    assert P.checkPermutation(perm) == 0

    for _ in range(args.warmup):
        ctx.run(start)
    kern = []
    t1 = time.perf_counter()
    for _ in range(args.steps):
        ctx.run(start)
        kern.append(ctx.timing().run_kernel_ms)
    wall = time.perf_counter() - t1
    last = ctx.timing()
    ms_per_step = wall * 1e3 / args.steps
    kern_ms = float(np.mean(kern))
    balg = alg_bytes(n, nnz)
    achieved = balg / (kern_ms * 1e-3) / 1e9

    line = {
        "metric": METRIC,
        "value": nnz / (ms_per_step * 1e-3),
        "unit": "nnz/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded 3D 7-point stencil 128^3, symmetric random relabelling)",
        "config": {"workload": "3D 7-point stencil 128^3, n=2097152, nnz=14581760, stable CM, GPU start search",
                   "n": n, "nnz": nnz, "start_node": start, "workers": last.workers, "batches": last.batches},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_reorder (persistent)", "kernel_ms": kern_ms, "alg_bytes": balg},
        "kernel_nnz_per_s": nnz / (kern_ms * 1e-3),
        "kernel_ms_p50": float(np.percentile(kern, 50)),
        "kernel_ms_min": float(np.min(kern)),
        "init_kernel_ms": last.init_kernel_ms,
        "end_to_end_ms": e2e_ms,
        "start_search_ms": tm.start_search_ms,
        "start_search_iterations": tm.start_iterations,
        "load_ms": tm.load_ms,
        "bandwidth": bw,
        "generate_s": gen_s,
    }
    ctx.close()

    if not args.no_cpu:
        cb = cpu_baseline()
        if cb:
            best_ms = cb["batch_ms"] if cb.get("batch_ms") else cb["serial_ms"]
            cores = cb["threads"] if cb.get("batch_ms") else 1
            line["cpu_baseline"] = {
                "value": cb["nnz"] / (best_ms * 1e-3),
                "unit": "nnz/s",
                "cores": cores,
                "kind": cb["kind"],
                "sample": (f"same 128^3 matrix and start; reference RealCPU::rcm, 25 iterations mean: "
                           f"CPU_BATCH {cores} threads batch 512 = {cb.get('batch_ms')} ms, "
                           f"serial = {cb['serial_ms']} ms; host {cb.get('cpu_model')}"),
                "serial_nnz_per_s": cb["nnz"] / (cb["serial_ms"] * 1e-3),
            }
        else:
            line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
