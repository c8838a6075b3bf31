"""GPU: the serving workload's power-law members at their real sizes, bit-exact against the oracle.

R-MAT graphs of scale 18, 19 and 20 (edge factor 16, (a, b, c) = (.57, .19, .19), symmetrised, duplicates
removed, diagonal added, components linked by the reference's repair) go

* through the single-matrix context: the start search, every engine variant (0-4 and the level-synchronous 5),
  the autotuned choice, reversed output and the bandwidth;
* through the server, in a seeded Poisson stream of the serving mix (the batched mix of stencils and random
  geometric graphs, n log-uniform in [1k, 200k], with the R-MAT requests among them).

The oracle gives the start (SimplePeripheral::findPeripheral, src/simple_peripheral.cpp:92-141), the order
(RealCPU::rcm_seq<true>, src/rcm_cpu.cpp:57-104) and the bandwidth (computeBandwidth, src/main.cpp:143-184).
"""
from __future__ import annotations

import time

import numpy as np
import pytest

import oracle_lib as O
import parallelbatchrcm_amd as P

pytestmark = pytest.mark.gpu

_CACHE = {}


def _say(*a):
    print("[case]", *a, flush=True)


def _rmat(scale: int):
    """(matrix, start, sweeps, CM order, bandwidth) from the oracle; generated once per session."""
    if scale not in _CACHE:
        t0 = time.perf_counter()
        m = P.connect_components(P.workloads.rmat(scale, 16, 0.57, 0.19, 0.19, seed=50 + scale))[0]
        rc, s, it, _ = O.peripheral(m.row_offsets, m.col_ids)
        assert rc == 0
        want = O.cm(m.row_offsets, m.col_ids, s)
        _CACHE[scale] = (m, s, it, want, O.bandwidth(m.row_offsets, m.col_ids, want))
        _say(f"R-MAT {scale}: n={m.rows} nnz={m.nnz} start={s} ({time.perf_counter() - t0:.1f} s to build)")
    return _CACHE[scale]


@pytest.mark.parametrize("scale", [18, 20])
def test_rmat_context_all_variants_match_oracle(scale):
    m, s, it, want, bw = _rmat(scale)
    with P.RcmContext(timeout_ms=60000) as c:
        _say("R-MAT", scale, "start search")
        c.load(m)
        assert c.find_start() == s
        assert c.timing().start_iterations == it
        _say("R-MAT", scale, "default variant, reversed")
        c.run(s, P.RCM_FLAG_STABLE | P.RCM_FLAG_REVERSE)
        assert np.array_equal(c.perm(), want[::-1])
        assert c.bandwidth() == bw
        for v in range(6):
            _say("R-MAT", scale, "variant", v)
            c.set_param("variant", v)
            c.run(s)
            assert np.array_equal(c.perm(), want), v
        c.set_param("variant", 0)
        _say("R-MAT", scale, "autotune")
        tune = c.autotune(s, repeats=1)
        assert c.get_param("variant") == tune.variant
        c.run(s)
        assert np.array_equal(c.perm(), want)


def test_rmat_one_shot_matches_oracle():
    m, s, _, want, bw = _rmat(19)
    _say("R-MAT 19 one shot")
    perm, got_bw, used = P.rcm_gpu_batch(m, -1, P.RCM_FLAG_STABLE | P.RCM_FLAG_BANDWIDTH)
    assert used == s
    assert np.array_equal(perm, want)
    assert got_bw == bw


def test_server_poisson_stream_with_rmat_matches_oracle():
    """A seeded Poisson stream (200 requests/s) of 24 mix requests and 4 R-MAT requests (scales 18, 19, 20, 18),
    submitted while earlier ones run; every result equals the oracle's start, order and bandwidth."""
    bs = P.workloads.batch_mix(24, seed=61)
    reqs = [(bs.matrix(k), None) for k in range(len(bs))]
    reqs += [(None, sc) for sc in (18, 19, 20, 18)]
    rng = np.random.default_rng(62)
    reqs = [reqs[i] for i in rng.permutation(len(reqs))]
    mats = [m if m is not None else _rmat(sc)[0] for m, sc in reqs]
    gaps = rng.exponential(1.0 / 200.0, size=len(mats))
    _say("server Poisson stream", len(mats), "requests")
    with P.RcmServer(lanes=4, job_timeout_ms=60000) as srv:
        ids = []
        due = time.perf_counter()
        for m, g in zip(mats, gaps):
            due += g
            dt = due - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(m, -1, P.RCM_FLAG_STABLE | P.RCM_FLAG_BANDWIDTH))
        out = [srv.wait(i, 120000) for i in ids]
        st = srv.stats()
    assert st.completed == len(mats) and st.failed == 0
    assert st.max_inflight > 1  # requests overlapped on the device
    for k, ((m0, sc), m, (rc, perm, js)) in enumerate(zip(reqs, mats, out)):
        assert rc == 0, k
        if sc is not None:
            _, s, _, want, bw = _rmat(sc)
        else:
            rc0, s, _, _ = O.peripheral(m.row_offsets, m.col_ids)
            assert rc0 == 0
            want = O.cm(m.row_offsets, m.col_ids, s)
            bw = O.bandwidth(m.row_offsets, m.col_ids, want)
        assert js.start_node == s, k
        assert np.array_equal(perm, want), k
        assert js.bandwidth == bw, k
