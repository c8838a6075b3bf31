#!/usr/bin/env python3
"""Benchmark of the GPU_BATCH reorder path on one MI355X.

Default workload (one JSON line on stdout): 3D 7-point stencil 128^3 (n = 2,097,152, nnz = 14,581,760),
seeded symmetric random relabelling, resident in HBM. A step is one Cuthill-McKee reorder of that matrix
(state-reset kernel + persistent reorder kernel, synchronised), from the pseudo-peripheral start node found
once on the GPU before timing (the reference also excludes the start search from its timed region,
src/device/rcm_gpu_batch.cu:47-99).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--batched COUNT] [--serving REQUESTS]

Fields beyond the contract: kernel-only nnz/s, end-to-end time (upload + start search + reorder + bandwidth
+ download), the bandwidth of the result, the batched form (4096-matrix mix, n log-uniform in [1k, 200k],
stencils and random geometric graphs: matrices/s of one persistent launch against one-at-a-time plain GPU
calls on a seeded sample of the same matrices) and the serving form (a seeded Poisson stream of requests,
95% that mix and 5% R-MAT scale 18-20, at several arrival rates: p50/p99 latency and throughput).
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


CPU_SHARE = 16  # host cores one GPU's jobs may use on the GPU host


def cpu_baseline(timeout_s: int = 400):
    """Reference CPU paths timed on this host, in a child process (the reference's CPU_BATCH is third-party
    code; a crash there must not end the benchmark): on the 128^3 stencil the serial rcm_seq<true> and CPU_BATCH
    (batch size 512) at every thread count t = 1 .. the host share, 25 iterations each (src/main.cpp:456-464,
    src/rcm_cpu.cpp:113-191); on the 100 x 100 2D stencil (the reference's CPU-path configuration) the serial
    path and CPU_BATCH at the best t."""
    code = r"""
import json, os, sys, time
sys.path.insert(0, %r); sys.path.insert(0, os.path.join(%r, 'tests'))
import numpy as np, oracle_lib as O, parallelbatchrcm_amd as P
m = P.workloads.config(2, seed=%d)
rc, s, it, _ = O.peripheral(m.row_offsets, m.col_ids)
try:
    affinity = len(os.sched_getaffinity(0))
except Exception:
    affinity = os.cpu_count() or 1
share = max(1, min(%d, affinity))
out = {"n": m.rows, "nnz": m.nnz, "nproc": os.cpu_count(), "affinity": affinity, "share": share}
if O.ref_available():
    _, ser = O.ref_rcm(m.row_offsets, m.col_ids, s, threads=1, batchsize=1)
    sweep = {}
    for t in range(1, share + 1):
        _, sweep[t] = O.ref_rcm(m.row_offsets, m.col_ids, s, threads=t, batchsize=512)
        print("CPUSWEEP", t, sweep[t], flush=True)
    best = min(sweep, key=sweep.get)
    c1 = P.workloads.config(1, seed=1)
    rc1, s1, _, _ = O.peripheral(c1.row_offsets, c1.col_ids)
    _, c1_ser = O.ref_rcm(c1.row_offsets, c1.col_ids, s1, threads=1, batchsize=1)
    _, c1_bat = O.ref_rcm(c1.row_offsets, c1.col_ids, s1, threads=best, batchsize=512)
    out.update(kind="reference", serial_ms=ser, batch_ms=sweep[best], best_threads=best,
               batch_sweep_ms={str(t): v for t, v in sweep.items()},
               c1={"n": c1.rows, "nnz": c1.nnz, "serial_ms": c1_ser, "batch_ms": c1_bat, "threads": best})
else:
    ts = []
    for _ in range(5):
        t = time.perf_counter(); O.cm(m.row_offsets, m.col_ids, s); ts.append((time.perf_counter() - t) * 1e3)
    out.update(kind="port", serial_ms=sum(ts) / len(ts), batch_ms=None, best_threads=1)
try:
    out["cpu_model"] = [l.split(':', 1)[1].strip() for l in open('/proc/cpuinfo') if l.startswith('model name')][0]
except Exception:
    out["cpu_model"] = "unknown"
print("CPUBASE " + json.dumps(out))
""" % (ROOT, ROOT, SEED_C2, CPU_SHARE)
    try:
        res = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=timeout_s)
    except subprocess.TimeoutExpired:
        return None
    for line in res.stdout.splitlines():
        if line.startswith("CPUBASE "):
            return json.loads(line[8:])
    print(res.stderr[-2000:], file=sys.stderr)
    return None


def bench_batched(count: int, seed: int, steps: int, warmup: int, sample: int, autotune: bool = True):
    """Batched form on the 4096-matrix mix: every matrix reordered by one persistent launch (device-resident,
    hipGraph replay), against one-at-a-time plain GPU calls (rcm_ctx_run per matrix, inputs resident) on a
    seeded sample of the same matrices."""
    import numpy as np

    import parallelbatchrcm_amd as P

    t0 = time.perf_counter()
    bs = P.workloads.batch_mix(count, seed=seed)
    gen_s = time.perf_counter() - t0
    out = {"matrices": count, "sum_n": int(bs.n.astype(np.int64).sum()), "sum_nnz": int(bs.nnz.sum()),
           "generate_s": gen_s}
    e2e0 = time.perf_counter()
    with P.RcmBatch(timeout_ms=120000, graph=True) as b:
        b.load(bs)
        starts, iters, st = b.order()  # start search and order in one call
        if not (st == 0).all():
            raise RuntimeError(f"batched search and order failed for {int((st != 0).sum())} matrices")
        bw = b.bandwidth()
        perms = b.perms()
        e2e_s = time.perf_counter() - e2e0
        tm = b.timing()
        out.update(load_ms=tm.load_ms, order_ms=tm.order_ms, order_launches=int(tm.order_passes),
                   end_to_end_s=e2e_s, workers=tm.workers,
                   sweeps_histogram={str(int(v)): int(c) for v, c in zip(*np.unique(iters, return_counts=True))})
        # the two-call path's search alone (find_start; the run is timed below), for comparison
        s2, _, st2 = b.find_start()
        if not (np.array_equal(s2, starts) and (st2 == 0).all()):
            raise RuntimeError("batched start search disagrees with the fused search")
        tm = b.timing()
        out.update(start_search_ms=tm.start_search_ms, start_sweeps=tm.start_sweeps)
        if autotune:
            t0 = time.perf_counter()
            r = b.autotune(starts)
            out["autotune"] = {"variant": r.variant, "threads": r.threads, "workers": r.workers,
                               "candidates": r.candidates, "best_ms": r.best_ms, "default_ms": r.default_ms,
                               "helper_nodes": r.helper_nodes, "tuning_s": time.perf_counter() - t0}
        for _ in range(warmup):
            b.run(starts)
        walls, devs = [], []
        for _ in range(steps):
            w0 = time.perf_counter()
            b.run(starts)
            walls.append(time.perf_counter() - w0)
            devs.append(b.timing().run_ms)
        tm = b.timing()
        out.update(run_ms_mean=float(np.mean(devs)), run_wall_ms_mean=float(np.mean(walls)) * 1e3,
                   matrices_per_s=count / float(np.mean(walls)), nnz_per_s=out["sum_nnz"] / float(np.mean(walls)),
                   batches=int(tm.batches), graph_replays=int(tm.graph_replays),
                   end_to_end_matrices_per_s=count / e2e_s)
    # every order of the first run checked on the GPU (outside the timed regions)
    out["verified"] = verify_all([bs.matrix(k) for k in range(count)], starts, perms)
    # one-at-a-time plain calls on a seeded sample (inputs resident, start known: the same work per matrix)
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(count, size=min(sample, count), replace=False))
    single = 0.0
    with P.RcmContext(timeout_ms=60000) as c:
        c.load(bs.matrix(int(idx[0])))
        c.run(int(starts[idx[0]]))  # warm the kernel
        for k in idx:
            c.load(bs.matrix(int(k)))
            w0 = time.perf_counter()
            c.run(int(starts[k]))
            single += time.perf_counter() - w0
    out["one_at_a_time_sample"] = int(idx.shape[0])
    out["one_at_a_time_matrices_per_s"] = idx.shape[0] / single
    out["speedup_vs_one_at_a_time"] = out["matrices_per_s"] / out["one_at_a_time_matrices_per_s"]
    out["bandwidth_sum"] = int(bw.sum())
    return out


def bench_serving(requests: int, fracs, seed: int, lanes: int, cold_sample: int, workers: int = 0):
    """Serving form: a seeded stream of reorder requests (95% the batched mix, 5% R-MAT scale 18-20,
    a,b,c = .57,.19,.19, symmetrised, diagonal, connected) through rcm_server. The same requests are first served
    one at a time on equal terms: each alone, with upload, GPU start search, reorder and download (a warm
    context; the plain one-shot call rcm_gpu_batch, which builds its context per call as GPU_BATCH::rcm does,
    on a sample). Capacity is the server's throughput with every request submitted at once; the latency runs
    use Poisson arrivals at fractions of that capacity (latency: arrival -> result on the host). Every result of
    the capacity run is checked on the GPU afterwards."""
    import threading

    import numpy as np

    import parallelbatchrcm_amd as P

    rng = np.random.default_rng(seed)
    nmix = requests - max(1, requests // 20)
    bs = P.workloads.batch_mix(nmix, seed=seed)
    pool = [bs.matrix(k) for k in range(nmix)]
    rmats = [P.connect_components(P.workloads.rmat(sc, 16, 0.57, 0.19, 0.19, seed=seed + sc))[0] for sc in (18, 19, 20)]
    reqs = pool + [rmats[i % 3] for i in range(requests - nmix)]
    order = rng.permutation(len(reqs))
    reqs = [reqs[i] for i in order]
    out = {"requests": len(reqs), "rmat_requests": requests - nmix, "lanes": lanes,
           "sum_nnz": int(sum(m.nnz for m in reqs)), "runs": []}

    # one at a time, each request alone on the GPU (warm context)
    with P.RcmContext(timeout_ms=60000) as c:
        c.load(reqs[0])
        c.order()
        t0 = time.perf_counter()
        for m in reqs:
            c.load(m)
            c.order()  # start search and order in one call, as the server does per group
            c.perm()
        warm = time.perf_counter() - t0
    out["one_at_a_time_per_s"] = len(reqs) / warm
    sample = reqs[: min(cold_sample, len(reqs))]
    t0 = time.perf_counter()
    for m in sample:
        P.rcm_gpu_batch(m, -1, P.RCM_FLAG_STABLE)
    out["one_shot_call_per_s"] = len(sample) / (time.perf_counter() - t0)
    out["one_shot_call_sample"] = len(sample)
    print(f"[bench] serving baseline: one at a time {out['one_at_a_time_per_s']:.1f} requests/s (warm context), "
          f"one-shot calls {out['one_shot_call_per_s']:.1f}/s", file=sys.stderr, flush=True)

    def serve(gaps, borrow=True):
        lat, dev, que, res = [], [], [], [None] * len(reqs)
        with P.RcmServer(lanes=lanes, workers=workers) as srv:
            ids = [None] * len(reqs)

            def collect():
                for i in range(len(reqs)):
                    while ids[i] is None:
                        time.sleep(0.0002)
                    rc, perm, js = srv.wait(ids[i], 120000)
                    if rc != 0 or perm.shape[0] != reqs[i].rows:
                        raise RuntimeError(f"request {i} failed: {rc}")
                    lat.append(js.latency_ms)
                    dev.append(js.device_ms)
                    que.append(js.queue_ms)
                    res[i] = (js.start_node, perm)

            th = threading.Thread(target=collect)
            th.start()
            t0 = time.perf_counter()
            due = t0
            for i, m in enumerate(reqs):
                if gaps is not None:
                    due += gaps[i]
                    dt = due - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                ids[i] = srv.submit(m, -1, P.RCM_FLAG_STABLE, borrow=borrow)
            submitted = time.perf_counter() - t0
            th.join()
            wall = time.perf_counter() - t0
            st = srv.stats()
        return wall, lat, dev, que, st, res, submitted

    # the requests' arrays outlive their requests here, so the server reads them in place (RCM_FLAG_BORROW), as
    # the one-at-a-time loop above uploads from them; the copying submit is measured beside it
    wall_copy, *_, sub_copy = serve(None, borrow=False)
    out["capacity_copying_submit_per_s"] = len(reqs) / wall_copy
    out["capacity_copying_submit_s"] = sub_copy
    wall, lat, dev, que, st, res, sub = serve(None)
    cap = len(reqs) / wall
    out["capacity_per_s"] = cap
    out["capacity_submit_s"] = sub
    out["pinned_lanes"] = int(st.pinned_lanes)
    out["capacity_p50_ms"] = float(np.percentile(lat, 50))
    out["capacity_p99_ms"] = float(np.percentile(lat, 99))
    out["capacity_groups"] = int(st.loop_iterations)
    g = max(1, int(st.loop_iterations))
    out["capacity_group_breakdown_ms"] = {  # per group, mean: where a group's host wall time goes
        "host_copy": st.group_host_copy_ms / g, "load_h2d_validate": st.group_load_ms / g,
        "search_and_order": st.group_order_ms / g, "bandwidth_d2h": st.group_readback_ms / g,
        "deliver": st.group_deliver_ms / g, "launches": st.group_launches / g}
    out["speedup_vs_one_at_a_time"] = cap / out["one_at_a_time_per_s"]
    print(f"[bench] serving capacity {cap:.1f} requests/s ({out['speedup_vs_one_at_a_time']:.1f}x one at a time; "
          f"submitting took {sub * 1e3:.0f} ms); with a copying submit {out['capacity_copying_submit_per_s']:.1f}/s "
          f"(submitting took {sub_copy * 1e3:.0f} ms); lanes with pinned staging {st.pinned_lanes} of {lanes}",
          file=sys.stderr, flush=True)
    out["verified"] = verify_all(reqs, [r[0] for r in res], [r[1] for r in res])
    for f in fracs:
        rate = f * cap
        gaps = rng.exponential(1.0 / rate, size=len(reqs))
        wall, lat, dev, que, st, _, _ = serve(gaps)
        out["runs"].append({
            "fraction_of_capacity": f, "arrival_rate_per_s": rate, "throughput_per_s": len(reqs) / wall,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mean_ms": float(np.mean(lat)), "p99_queue_ms": float(np.percentile(que, 99)),
            "p50_device_ms": float(np.percentile(dev, 50)), "max_inflight": int(st.max_inflight),
            "groups": int(st.loop_iterations),
            "group_order_ms_mean": st.group_order_ms / max(1, int(st.loop_iterations))})
        print(f"[bench] serving at {f:.2f} of capacity ({rate:.0f}/s): p50 {out['runs'][-1]['p50_ms']:.1f} ms "
              f"p99 {out['runs'][-1]['p99_ms']:.1f} ms", file=sys.stderr, flush=True)
    return out


def permute_bytes(n: int, nnz: int) -> int:
    """Compulsory HBM bytes of B = P A P^T with float64 values: read A (row_ptr, col_idx, values) and perm,
    write B."""
    return 2 * (4 * (n + 1) + 12 * nnz) + 4 * n


def canonicalize_bytes(n: int, nnz_in: int, nnz_out: int) -> int:
    """Compulsory HBM bytes of COO -> CSR + prepareRCM with float64 values: read (row, col, value) triples,
    write row_ptr, col_idx and values."""
    return 16 * nnz_in + 4 * (n + 1) + 12 * nnz_out


def bench_pipeline(m, perm, steps: int, seed: int):
    """The steps either side of the reorder on the 128^3 stencil: canonicalisation of its lower triangle as a
    shuffled COO with float64 values (a MatrixMarket 'symmetric' file's content), and B = P A P^T of the
    canonical matrix with the RCM permutation. Device time from HIP events around each call; the host C++
    convert + prepareRCM on the same COO beside it."""
    import numpy as np

    import parallelbatchrcm_amd as P

    rng = np.random.default_rng(seed)
    n = m.rows
    rows = np.repeat(np.arange(n, dtype=np.int32), np.diff(m.row_offsets))
    lower = rows >= m.col_ids
    r = rows[lower]
    c = m.col_ids[lower]
    order = rng.permutation(r.size)
    r, c = np.ascontiguousarray(r[order]), np.ascontiguousarray(c[order])
    v = rng.standard_normal(r.size)
    out = {"workload": "128^3 7-point stencil: lower triangle as a shuffled COO with float64 values "
                       f"({r.size} entries) -> canonical CSR ({m.nnz} entries); then P A P^T with the RCM order"}
    with P.MatrixOps() as ops:
        canon, links = ops.canonicalize(n, n, r, c, v, True)  # warm-up
        cms = []
        for _ in range(steps):
            ops.canonicalize(n, n, r, c, v, True)
            cms.append(ops.last_ms)
        b = ops.permute(canon, perm)  # warm-up
        pms = []
        for _ in range(steps):
            ops.permute(canon, perm)
            pms.append(ops.last_ms)
    t0 = time.perf_counter()
    host, hlinks = P.prepare_rcm(P.convert_coo(n, n, r, c, v), True)
    host_ms = (time.perf_counter() - t0) * 1e3
    same = (hlinks == links and np.array_equal(host.col_ids, canon.col_ids) and np.array_equal(host.data, canon.data))
    cm = float(np.mean(cms))
    pm = float(np.mean(pms))
    cb = canonicalize_bytes(n, r.size, canon.nnz)
    pb = permute_bytes(n, canon.nnz)
    out["canonicalize"] = {"device_ms": cm, "device_ms_min": float(np.min(cms)), "alg_bytes": cb,
                           "achieved_GBs": cb / (cm * 1e-3) / 1e9, "frac": cb / (cm * 1e-3) / 1e9 / HBM_PEAK_GBS,
                           "entries_per_s": r.size / (cm * 1e-3), "links": links,
                           "host_cpp_ms": host_ms, "equal_to_host": bool(same)}
    out["permute"] = {"device_ms": pm, "device_ms_min": float(np.min(pms)), "alg_bytes": pb,
                      "achieved_GBs": pb / (pm * 1e-3) / 1e9, "frac": pb / (pm * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "nnz_per_s": canon.nnz / (pm * 1e-3)}
    print(f"[bench] pipeline: canonicalize {cm:.2f} ms (host C++ {host_ms:.0f} ms), permute {pm:.2f} ms",
          file=sys.stderr, flush=True)
    return out


def bench_variants_rmat(scale: int, steps: int, seed: int):
    """The engine variants on a power-law graph: R-MAT (a, b, c) = (.57, .19, .19), edge factor 16, symmetrised,
    repaired; kernel milliseconds of the batch engine's default (0) and widest (4) variants and of the
    level-synchronous engine (5), and what the autotuner picks. Every variant gives the same order."""
    import numpy as np

    import parallelbatchrcm_amd as P

    m = P.workloads.rmat(scale, seed=seed)
    m, links = P.connect_components(m)
    out = {"workload": f"R-MAT scale {scale} (n={m.rows}, nnz={m.nnz}, {links} repair links)"}
    with P.RcmContext(timeout_ms=60000) as ctx:
        ctx.load(m)
        s = ctx.find_start()
        ref = None
        for v in (0, 4, 5):
            ctx.set_param("variant", v)
            ts = []
            for _ in range(steps + 1):
                ctx.run(s)
                ts.append(ctx.timing().run_kernel_ms)
            out[f"variant{v}_ms"] = float(np.median(ts[1:]))
            out[f"variant{v}_steps"] = ctx.timing().batches  # batches (0-4) or BFS levels (5)
            # the order checked on the device against the characterisation of the stable CM order (after timing)
            out[f"variant{v}_verified"] = ctx.verify(s) == 0
        ctx.set_param("variant", 0)
        tune = ctx.autotune(s, repeats=2)
        out["autotune_variant"] = tune.variant
        out["autotune_ms"] = tune.best_ms
    print(f"[bench] R-MAT {scale}: variant 0 {out['variant0_ms']:.2f} ms, 4 {out['variant4_ms']:.2f} ms, "
          f"level-synchronous {out['variant5_ms']:.2f} ms; autotune picks {tune.variant}", file=sys.stderr, flush=True)
    if not all(out[f"variant{v}_verified"] for v in (0, 4, 5)):
        raise RuntimeError(f"R-MAT {scale}: an engine variant's order failed the device check")
    return out


def verify_all(mats, starts, perms, flags=0):
    """Checks every order on the GPU (rcm_ctx_verify: the characterisation of the stable CM order, O(n + nnz)),
    outside any timed region. Returns the number of orders checked; raises on the first that fails."""
    import parallelbatchrcm_amd as P

    with P.RcmContext() as c:
        for k, (m, s, p) in enumerate(zip(mats, starts, perms)):
            c.load(m)
            r = c.verify(int(s), p, flags)
            if r != 0:
                raise RuntimeError(f"order {k} (n={m.rows}) is not the stable CM order from {s}: position {r - 1}")
    return len(mats)


def bench_single(which: int, scale: int, seed: int, steps: int, name: str):
    """A single-matrix leg: kernel milliseconds (autotuned engine), end-to-end time and the device check."""
    import numpy as np

    import parallelbatchrcm_amd as P

    m = P.workloads.config(which, scale=scale, seed=seed)
    with P.RcmContext(timeout_ms=30000) as ctx:
        w0 = time.perf_counter()
        ctx.load(m)
        s, _ = ctx.order(flags=P.RCM_FLAG_STABLE | P.RCM_FLAG_REVERSE)
        bw = ctx.bandwidth()
        ctx.perm()
        e2e = (time.perf_counter() - w0) * 1e3
        tm = ctx.timing()
        ctx.find_start()  # the two-call path's search alone, for comparison
        tm_search = ctx.timing().start_search_ms
        tune = ctx.autotune(s, repeats=3)
        ts = []
        for _ in range(steps):
            ctx.run(s)
            ts.append(ctx.timing().run_kernel_ms)
        ok = ctx.verify(s) == 0
    k = float(np.mean(ts))
    out = {"workload": name, "n": m.rows, "nnz": m.nnz, "start_node": s, "kernel_ms": k,
           "kernel_nnz_per_s": m.nnz / (k * 1e-3), "achieved_GBs": alg_bytes(m.rows, m.nnz) / (k * 1e-3) / 1e9,
           "end_to_end_warm_ms": e2e, "order_ms": tm.order_ms, "order_passes": tm.order_passes,
           "start_search_ms": tm_search, "start_sweeps": tm.start_iterations, "bandwidth": bw,
           "variant": tune.variant, "workers": tune.workers, "verified": ok}
    print(f"[bench] {name}: kernel {k:.3f} ms, end to end {e2e:.1f} ms, verified {ok}", file=sys.stderr, flush=True)
    if not ok:
        raise RuntimeError(f"{name}: the order failed the device check")
    return out


COPY_BYTES = 256 << 20  # calibration copy for the HBM traffic counters


def pmc_probe(runs: int, claim_filter: int = -1, variant: int = -1, workers: int = 0, legs: str = "single") -> int:
    """Workload of the rocprofv3 --pmc passes: the 128^3 stencil reorder `runs` times, then `runs` elementwise
    kernels that read and write COPY_BYTES each (a plain copy of known traffic) to calibrate the counters. The
    engine configuration is given explicitly (variant, workgroups), set to what the timed leg's autotuner picks:
    tuning under the counters would time perturbed runs. `legs` may add "batched" (the 4096-matrix mix, `runs`
    launches of k_reorder_many with the configuration the batched leg's autotuner picks: variant 4, 256
    workgroups, one helper per 131072 nodes) and "pipeline" (`runs` canonicalisations of the 128^3 stencil's
    lower triangle, the pipeline leg's input)."""
    import numpy as np
    import torch

    import parallelbatchrcm_amd as P

    legs = set(legs.split(","))
    m = P.workloads.config(2, seed=SEED_C2)
    if "single" in legs:
        with P.RcmContext(timeout_ms=30000) as ctx:
            if claim_filter >= 0:
                ctx.set_param("claim_filter", claim_filter)
            if variant >= 0:
                ctx.set_param("variant", variant)
            if workers > 0:
                ctx.set_param("workers", workers)
            ctx.load(m)
            s = ctx.find_start()
            print(f"pmc probe: variant {variant}, {workers} workgroups (-1 / 0: library default)", flush=True)
            for _ in range(runs):
                ctx.run(s)
    if "batched" in legs:
        bs = P.workloads.batch_mix(4096, seed=4)
        with P.RcmBatch(timeout_ms=120000, graph=False) as b:
            b.set_param("variant", 4)
            b.set_param("workers", 256)
            b.set_param("helper_nodes", 131072)
            b.load(bs)
            starts, _, st = b.find_start()
            print("pmc probe: batched mix of 4096, variant 4, 256 workgroups, helper nodes 131072", flush=True)
            for _ in range(runs):
                if not (b.run(starts) == 0).all():
                    raise RuntimeError("batched run failed")
    if "pipeline" in legs:
        rng = np.random.default_rng(SEED_C2)
        rows = np.repeat(np.arange(m.rows, dtype=np.int32), np.diff(m.row_offsets))
        lower = rows >= m.col_ids
        order = rng.permutation(int(lower.sum()))
        r = np.ascontiguousarray(rows[lower][order])
        c = np.ascontiguousarray(m.col_ids[lower][order])
        v = rng.standard_normal(r.size)
        print(f"pmc probe: canonicalisation of {r.size} triples", flush=True)
        with P.MatrixOps() as ops:
            for _ in range(runs):
                ops.canonicalize(m.rows, m.rows, r, c, v, True)
    a = torch.empty(COPY_BYTES // 4, dtype=torch.int32, device="cuda")
    a.fill_(1)
    b = torch.empty_like(a)
    for _ in range(runs):
        torch.add(a, 1, out=b)  # an elementwise kernel: reads and writes COPY_BYTES
    torch.cuda.synchronize()
    print(f"pmc probe done: {runs} reorders, {runs} copies of {COPY_BYTES} bytes", flush=True)
    return 0


def measured_traffic():
    """HBM bytes per k_reorder launch on the 128^3 stencil from the committed rocprofv3 --pmc summary
    (profiles/k_reorder_stencil3d_128_traffic.md, line 'traffic_bytes_per_launch: <int>'), or None."""
    path = os.path.join(ROOT, "profiles", "k_reorder_stencil3d_128_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workers", type=int, default=0)
    ap.add_argument("--variant", type=int, default=-1, help="--pmc-probe only: engine variant (-1: library default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--pmc-probe", type=int, default=0, help="only the workload of the rocprofv3 --pmc passes")
    ap.add_argument("--pmc-legs", type=str, default="single",
                    help="--pmc-probe only: comma list of single, batched, pipeline")
    ap.add_argument("--claim-filter", type=int, default=-1, help="claim pre-read filter 0/1/2 (-1: library default)")
    ap.add_argument("--no-autotune", dest="autotune", action="store_false",
                    help="time the default configuration instead of the autotuned one")
    ap.add_argument("--batched", type=int, default=4096,
                    help="batched form on COUNT matrices of the mix (0 = skip)")
    ap.add_argument("--batched-steps", type=int, default=5)
    ap.add_argument("--one-at-a-time-sample", type=int, default=256)
    ap.add_argument("--serving", type=int, default=400, help="serving form: requests per run (0 = skip)")
    ap.add_argument("--serving-fracs", type=str, default="0.25,0.5,0.75",
                    help="Poisson arrival rates as fractions of the measured capacity")
    ap.add_argument("--serving-lanes", type=int, default=4)
    ap.add_argument("--serving-workers", type=int, default=0,
                    help="persistent workgroups per serving lane (0: the library default)")
    ap.add_argument("--serving-cold-sample", type=int, default=48)
    ap.add_argument("--c3", type=int, default=1, help="the 1M-point random geometric graph leg (0 = skip)")
    ap.add_argument("--rmat", type=int, default=20, help="R-MAT scale of the engine-variant leg (0 = skip)")
    ap.add_argument("--pipeline", type=int, default=5,
                    help="steps of the canonicalise and permute legs on the 128^3 stencil (0 = skip)")
    args = ap.parse_args()
    if args.pmc_probe:
        return pmc_probe(args.pmc_probe, args.claim_filter, args.variant, args.workers, args.pmc_legs)
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: there is no multi-GPU path (asked for --gpus {args.gpus}). A set of "
              "matrices would shard by matrix, one process per GPU, each running the batched form on its share.",
              file=sys.stderr)
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
    if args.claim_filter >= 0:
        ctx.set_param("claim_filter", args.claim_filter)
    ctx.load(m)
    start, _ = ctx.order(flags=P.RCM_FLAG_STABLE | P.RCM_FLAG_REVERSE)  # start search and order in one call
    bw = ctx.bandwidth()
    perm = ctx.perm()
    e2e_ms = (time.perf_counter() - e2e0) * 1e3  # includes creating the context (the process's first HIP calls)
    # the same steps again on the existing context
    w0 = time.perf_counter()
    ctx.load(m)
    start, _ = ctx.order(flags=P.RCM_FLAG_STABLE | P.RCM_FLAG_REVERSE)
    bw = ctx.bandwidth()
    perm = ctx.perm()
    e2e_warm_ms = (time.perf_counter() - w0) * 1e3
    tm = ctx.timing()
    if ctx.find_start() != start:  # the two-call path's search alone, for comparison
        raise RuntimeError("start search disagrees with the fused search")
    search_ms = ctx.timing().start_search_ms
    assert P.checkPermutation(perm) == 0

    tune = None
    if args.autotune:
        t0 = time.perf_counter()
        tune = ctx.autotune(start, repeats=3)
        tune_s = time.perf_counter() - t0
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
    # a plain 1 GiB device copy, after the timed region: the better of the library's copy kernel and the runtime's
    copy_gbs = max(P.copy_bandwidth(1 << 30, 10)[0], P.copy_bandwidth(1 << 30, 10, runtime=True)[0])

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
                   "n": n, "nnz": nnz, "start_node": start, "workers": last.workers, "batches": last.batches,
                   "variant": tune.variant if tune is not None else 0},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "k_reorder (persistent)", "kernel_ms": kern_ms, "alg_bytes": balg,
                     "traffic_source": "profiles/k_reorder_stencil3d_128_traffic.md (the autotuned configuration)",
                     "measured_copy_GBs": copy_gbs, "frac_of_measured_copy": achieved / copy_gbs},
        "kernel_nnz_per_s": nnz / (kern_ms * 1e-3),
        "kernel_ms_p50": float(np.percentile(kern, 50)),
        "kernel_ms_min": float(np.min(kern)),
        "init_kernel_ms": last.init_kernel_ms,
        "end_to_end_ms": e2e_ms,
        "end_to_end_warm_ms": e2e_warm_ms,
        "order_ms": tm.order_ms,
        "order_passes": tm.order_passes,
        "start_search_ms": search_ms,
        "start_search_iterations": tm.start_iterations,
        "load_ms": tm.load_ms,
        "bandwidth": bw,
        "generate_s": gen_s,
    }
    if tune is not None:
        line["autotune"] = {"variant": tune.variant, "threads": tune.threads, "workers": tune.workers,
                            "candidates": tune.candidates, "best_ms": tune.best_ms, "default_ms": tune.default_ms,
                            "tuning_s": tune_s}
    ctx.close()

    if args.pipeline > 0:
        line["pipeline"] = bench_pipeline(m, perm, args.pipeline, 9)

    if args.c3:
        line["c3"] = bench_single(3, 0, 3, args.steps, "random geometric graph, 1M points in the unit square, mean "
                                  "degree ~12, relabelled and repaired")

    if args.rmat > 0:
        line["variants_rmat"] = bench_variants_rmat(args.rmat, 3, args.rmat)

    if args.batched > 0:
        print(f"[bench] batched form: {args.batched} matrices", file=sys.stderr, flush=True)
        line["batched"] = bench_batched(args.batched, 4, args.batched_steps, 1, args.one_at_a_time_sample,
                                        args.autotune)

    if args.serving > 0:
        fracs = [float(x) for x in args.serving_fracs.split(",") if x]
        print(f"[bench] serving form: {args.serving} requests", file=sys.stderr, flush=True)
        line["serving"] = bench_serving(args.serving, fracs, 5, args.serving_lanes, args.serving_cold_sample,
                                        args.serving_workers)

    if not args.no_cpu:
        print("[bench] CPU baseline (reference CPU paths on this host)", file=sys.stderr, flush=True)
        cb = cpu_baseline()
        if cb:
            best_ms = cb["batch_ms"] if cb.get("batch_ms") else cb["serial_ms"]
            cores = cb["best_threads"] if cb.get("batch_ms") else 1
            line["cpu_baseline"] = {
                "value": cb["nnz"] / (best_ms * 1e-3),
                "unit": "nnz/s",
                "cores": cores,
                "kind": cb["kind"],
                "sample": (f"same 128^3 matrix and start; reference RealCPU::rcm, 25 iterations mean each: "
                           f"CPU_BATCH batch 512 at t = 1..{cb['share']} threads (host nproc {cb['nproc']}, "
                           f"affinity {cb['affinity']}, share {cb['share']}), best t = {cores}: {best_ms:.2f} ms; "
                           f"serial = {cb['serial_ms']:.2f} ms; host {cb.get('cpu_model')}"),
                "serial_nnz_per_s": cb["nnz"] / (cb["serial_ms"] * 1e-3),
                "nproc": cb["nproc"],
                "batch_sweep_ms": cb.get("batch_sweep_ms"),
                "c1_cpu": cb.get("c1"),
                "gpu_over_cpu_batch": (cb["nnz"] / (best_ms * 1e-3)) and line["kernel_nnz_per_s"] / (
                    cb["nnz"] / (best_ms * 1e-3)),
            }
        else:
            line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
