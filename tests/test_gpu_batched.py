"""GPU: many matrices in one persistent launch (rcm_batch / rcm_gpu_batch_many) against the oracle, bit-exact,
and against single-matrix runs of the same matrices."""
from __future__ import annotations

import numpy as np
import pytest

import graphs as G
import oracle_lib as O
import parallelbatchrcm_amd as P

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _oracle(m: P.CSR, start=None):
    if start is None:
        rc, start, _, _ = O.peripheral(m.row_offsets, m.col_ids)
        assert rc == 0
    perm = O.cm(m.row_offsets, m.col_ids, start)
    return start, perm, O.bandwidth(m.row_offsets, m.col_ids, perm)


def _small_set():
    mats = [G.path(17), G.star(40), G.complete(9), G.grid2d(7, 5), G.random_connected(500, 1500, seed=4),
            G.with_duplicates(*G.random_connected(300, 600, seed=5), seed=6), G.path(1), G.path(2)]
    mats = [P.CSR(*m) for m in mats]
    mats.append(P.workloads.config(2, scale=12, seed=2))
    mats.append(P.workloads.config(3, scale=5000, seed=3))
    return P.BatchSet.from_matrices(mats)


def test_batch_small_set_matches_oracle():
    bs = _small_set()
    _say("batch small set", len(bs))
    with P.RcmBatch(timeout_ms=20000) as b:
        b.load(bs)
        starts, _, st = b.find_start()
        assert (st == 0).all()
        st = b.run(starts, P.RCM_FLAG_STABLE)
        assert (st == 0).all()
        bw = b.bandwidth()
        perms = b.perms()
    for k in range(len(bs)):
        s, want, wbw = _oracle(bs.matrix(k))
        assert starts[k] == s, k
        assert np.array_equal(perms[k], want), k
        assert bw[k] == wbw, k


@pytest.mark.parametrize("workers", [1, 5, 64, 0])
def test_batch_mix_matches_oracle(workers):
    bs = P.workloads.batch_mix(48, n_lo=500, n_hi=20000, seed=11)
    _say("batch mix 48, workers", workers)
    with P.RcmBatch(timeout_ms=30000, workers=workers) as b:
        b.load(bs)
        starts, _, st = b.find_start()
        assert (st == 0).all()
        assert (b.run(starts) == 0).all()
        bw = b.bandwidth()
        perms = b.perms()
    for k in range(len(bs)):
        s, want, wbw = _oracle(bs.matrix(k))
        assert starts[k] == s, k
        assert np.array_equal(perms[k], want), k
        assert bw[k] == wbw, k


def test_batch_equals_single_runs_and_graph_replays():
    bs = P.workloads.batch_mix(24, n_lo=1000, n_hi=30000, seed=12)
    _say("batch vs single, graph replays")
    alone = []
    with P.RcmContext(timeout_ms=20000) as c:
        for k in range(len(bs)):
            c.load(bs.matrix(k))
            s = c.find_start()
            c.run(s, P.RCM_FLAG_REVERSE)
            alone.append((s, c.perm()))
    starts = np.array([s for s, _ in alone], np.int32)
    with P.RcmBatch(timeout_ms=20000, graph=True) as b:
        b.load(bs)
        for rep in range(3):  # first run captures the graph, later runs replay it
            assert (b.run(starts, P.RCM_FLAG_REVERSE) == 0).all()
            perms = b.perms()
            for k in range(len(bs)):
                assert np.array_equal(perms[k], alone[k][1]), (rep, k)
            assert (b.run(starts, P.RCM_FLAG_STABLE) == 0).all()  # the forward graph between replays
            assert np.array_equal(b.perms()[0], alone[0][1][::-1])
        assert b.timing().graph_replays >= 4


def test_batch_disconnected_member_is_reported():
    rp, ci = G.from_edges(10, [(0, 1), (1, 2), (5, 6)])
    mats = [P.CSR(*G.path(9)), P.CSR(rp, ci), P.CSR(*G.star(12))]
    bs = P.BatchSet.from_matrices(mats)
    _say("batch with a disconnected member")
    perm, po, bw, st, rc = P.rcm_gpu_batch_many(bs, [0, 0, 0], P.RCM_FLAG_STABLE | P.RCM_FLAG_BANDWIDTH)
    assert st.tolist() == [0, P.RCM_ENOTCONNECTED, 0]
    assert rc == P.RCM_ENOTCONNECTED
    for k in (0, 2):
        m = mats[k]
        want = O.cm(m.row_offsets, m.col_ids, 0)
        assert np.array_equal(perm[po[k]: po[k] + m.rows], want)
        assert bw[k] == O.bandwidth(m.row_offsets, m.col_ids, want)
    assert bw[1] == -1


def test_gpu_batch_many_one_shot_auto_and_given_starts():
    bs = P.workloads.batch_mix(16, n_lo=1000, n_hi=10000, seed=13)
    _say("rcm_gpu_batch_many one shot")
    given = np.full(len(bs), -1, np.int32)
    given[::2] = 0
    perm, po, bw, st, rc = P.rcm_gpu_batch_many(bs, given, P.RCM_FLAG_REVERSE | P.RCM_FLAG_BANDWIDTH)
    assert rc == 0 and (st == 0).all()
    for k in range(len(bs)):
        m = bs.matrix(k)
        s, want, wbw = _oracle(m, 0 if given[k] == 0 else None)
        assert np.array_equal(perm[po[k]: po[k] + m.rows], want[::-1]), k
        assert bw[k] == wbw


@pytest.mark.slow
def test_batch_full_mix_4096():
    """The 4096-matrix mix (n log-uniform in [1k, 200k]): every order is a permutation satisfying the CM
    characterisation from its start; a seeded sample of 64 matrices is checked bit-exact with the oracle."""
    bs = P.workloads.batch_mix(4096, seed=4)
    _say("batch 4096 mix")
    with P.RcmBatch(timeout_ms=60000) as b:
        b.load(bs)
        starts, _, st = b.find_start()
        assert (st == 0).all()
        assert (b.run(starts) == 0).all()
        bw = b.bandwidth()
        perms = b.perms()
    rng = np.random.default_rng(4)
    sample = set(rng.choice(len(bs), 64, replace=False).tolist())
    for k in range(len(bs)):
        m = bs.matrix(k)
        assert O.check_perm(perms[k]) == 0, k
        assert O.verify_order(m.row_offsets, m.col_ids, int(starts[k]), perms[k]) == 0, k
        if k in sample:
            s, want, wbw = _oracle(m)
            assert starts[k] == s and np.array_equal(perms[k], want) and bw[k] == wbw, k


def test_batch_with_hub_members():
    """R-MAT and hub graphs (rows beyond the batch slots) inside a many-matrix launch."""
    mats = [P.CSR(*G.hub_graph(5000, 9000, seed=1)), P.CSR(*G.path(40)),
            P.CSR(*G.hub_graph(2600, 3000, seed=2, hubs=2))]
    m = P.workloads.rmat(14, seed=5)
    mats.append(P.connect_components(m)[0])
    bs = P.BatchSet.from_matrices(mats)
    _say("batch with hub members")
    with P.RcmBatch(timeout_ms=20000) as b:
        b.load(bs)
        starts, _, st = b.find_start()
        assert (st == 0).all()
        assert (b.run(starts) == 0).all()
        bw = b.bandwidth()
        perms = b.perms()
    for k in range(len(bs)):
        s, want, wbw = _oracle(bs.matrix(k))
        assert starts[k] == s and np.array_equal(perms[k], want) and bw[k] == wbw, k


@pytest.mark.parametrize("variant", [1, 2, 3, 4])
def test_batch_variants_match_oracle(variant):
    bs = P.workloads.batch_mix(20, n_lo=500, n_hi=20000, seed=31)
    _say("batch variant", variant)
    with P.RcmBatch(timeout_ms=30000) as b:
        b.set_param("variant", variant)
        b.load(bs)
        starts, _, st = b.find_start()
        assert (st == 0).all()
        assert (b.run(starts) == 0).all()
        perms = b.perms()
    for k in range(len(bs)):
        s, want, _ = _oracle(bs.matrix(k))
        assert starts[k] == s and np.array_equal(perms[k], want), k


def test_batch_autotune_keeps_results():
    bs = P.workloads.batch_mix(32, n_lo=1000, n_hi=30000, seed=32)
    _say("batch autotune")
    with P.RcmBatch(timeout_ms=30000) as b:
        b.load(bs)
        starts, _, _ = b.find_start()
        r = b.autotune(starts)
        assert 0 <= r.variant < 5 and r.candidates >= 10 and 0 < r.best_ms <= r.default_ms
        assert (b.run(starts) == 0).all()
        perms = b.perms()
    for k in range(len(bs)):
        m = bs.matrix(k)
        assert np.array_equal(perms[k], O.cm(m.row_offsets, m.col_ids, int(starts[k]))), k


def test_batch_autotune_with_disconnected_member():
    """A member that fails (disconnected) does not void the timings: every candidate is timed over the whole
    set, the best time is positive and no worse than the default's, and the connected members stay exact."""
    rp, ci = G.from_edges(12, [(0, 1), (1, 2), (2, 3), (6, 7), (7, 8)])
    mats = [P.CSR(rp, ci)] + [P.workloads.config(3, scale=4000 + 500 * k, seed=40 + k) for k in range(4)]
    bs = P.BatchSet.from_matrices(mats)
    _say("batch autotune with a disconnected member")
    with P.RcmBatch(timeout_ms=30000) as b:
        b.load(bs)
        starts = [0] + [O.peripheral(m.row_offsets, m.col_ids)[1] for m in mats[1:]]
        r = b.autotune(starts)
        assert r.candidates >= 10 and 0 < r.best_ms <= r.default_ms
        assert r.helper_nodes in (8192, 32768, 131072)
        st = b.run(starts)
        assert st[0] == P.RCM_ENOTCONNECTED and (st[1:] == 0).all()
        perms = b.perms()
    for k in range(1, len(mats)):
        m = mats[k]
        assert np.array_equal(perms[k], O.cm(m.row_offsets, m.col_ids, int(starts[k]))), k


@pytest.mark.skip(reason="the one-workgroup-per-matrix engine (tests/emu/engine_wg.hpp) faulted the GPU on its "
                         "first run with no cause shown; it is not built into the library (wg_max_n > 0 is "
                         "rejected) and is tested only in the CPU emulator")
@pytest.mark.parametrize("wg_max_n", [0, 1 << 18, 1 << 30])
def test_batch_one_workgroup_engine_matches_oracle(wg_max_n):
    """Matrices with short rows go to one workgroup each (RCM_PARAM_WG_MAX_N), the rest (a hub graph, an R-MAT)
    to the batch engine; 0 sends everything to the batch engine. Starts, orders and bandwidths are the oracle's
    either way, reversed orders included."""
    mats = [P.CSR(*G.path(17)), P.CSR(*G.hub_graph(3000, 4000, seed=9)), P.CSR(*G.grid2d(40, 30)),
            P.CSR(*G.with_duplicates(*G.random_connected(400, 900, seed=15), seed=16)),
            P.workloads.config(1, seed=2), P.workloads.config(2, scale=20, seed=3),
            P.workloads.config(3, scale=20000, seed=4), P.CSR(*G.path(1)),
            P.connect_components(P.workloads.rmat(12, seed=5))[0]]
    bs = P.BatchSet.from_matrices(mats)
    _say("one-workgroup engine, wg_max_n", wg_max_n)
    with P.RcmBatch(timeout_ms=30000) as b:
        b.set_param("wg_max_n", wg_max_n)
        b.load(bs)
        starts, _, st = b.find_start()
        assert (st == 0).all()
        assert (b.run(starts, P.RCM_FLAG_STABLE | P.RCM_FLAG_REVERSE) == 0).all()
        bw = b.bandwidth()
        perms = b.perms()
    for k in range(len(bs)):
        s, want, wbw = _oracle(bs.matrix(k))
        assert starts[k] == s, k
        assert np.array_equal(perms[k], want[::-1]), k
        assert bw[k] == wbw, k


def test_batch_directed_graphs_with_empty_rows():
    """Directed graphs with many empty rows beside symmetric ones, in one many-matrix launch, from node 0."""
    mats = [P.CSR(*G.directed_with_empty_rows(3000, 3000, 40, 1)), P.CSR(*G.grid2d(7, 5)),
            P.CSR(*G.directed_with_empty_rows(2000, 1500, 24, 2))]
    bs = P.BatchSet.from_matrices(mats)
    _say("batch directed", len(bs))
    with P.RcmBatch(timeout_ms=20000) as b:
        b.load(bs)
        st = b.run(np.zeros(len(bs), np.int32), P.RCM_FLAG_STABLE)
        assert (st == 0).all()
        perms = b.perms()
    for k, m in enumerate(mats):
        assert np.array_equal(perms[k], O.cm(m.row_offsets, m.col_ids, 0)), k


def test_batch_helper_cap_keeps_results():
    """RCM_PARAM_HELPER_NODES changes how many workgroups help one matrix, never the orders."""
    bs = P.workloads.batch_mix(24, n_lo=500, n_hi=20000, seed=13)
    _say("batch helper caps", len(bs))
    with P.RcmBatch(timeout_ms=30000) as b:
        b.load(bs)
        starts, _, st = b.find_start()
        assert (st == 0).all()
        got = {}
        for h in (1, 8192, 1 << 30):
            b.set_param("helper_nodes", h)
            assert (b.run(starts, P.RCM_FLAG_STABLE) == 0).all()
            got[h] = b.perms()
        with pytest.raises(Exception):
            b.set_param("helper_nodes", 0)
        b.set_param("wg_max_n", 0)  # reserved knob: only 0 (off) is accepted
        with pytest.raises(P.RcmError):
            b.set_param("wg_max_n", 1)
    for k in range(len(bs)):
        want = O.cm(bs.matrix(k).row_offsets, bs.matrix(k).col_ids, int(starts[k]))
        for h, perms in got.items():
            assert np.array_equal(perms[k], want), (k, h)


def test_large_upload_through_the_pinned_pipeline():
    """A set whose CSRs total more than 32 MiB is uploaded by 4 host threads through pinned double buffers (pieces
    of 16 MiB that straddle the two arrays and the matrices); then a single matrix of that size through rcm_ctx. Every
    order and bandwidth against the oracle, from the oracle's own start."""
    mats = [P.workloads.config(2, scale=80, seed=41), P.workloads.config(3, scale=300000, seed=42),
            P.workloads.config(2, scale=72, seed=43)]
    total = sum(4 * (m.rows + 1) + 4 * m.nnz for m in mats)
    assert total > 32 << 20, total
    want = []
    for m in mats:
        rc, s, _, _ = O.peripheral(m.row_offsets, m.col_ids)
        assert rc == 0
        want.append((s, O.cm(m.row_offsets, m.col_ids, s)))
    bs = P.BatchSet.from_matrices(mats)
    _say("large upload batch", len(bs), total)
    with P.RcmBatch(timeout_ms=30000) as b:
        b.load(bs)
        starts, _, st = b.order()
        assert (st == 0).all()
        perms = b.perms()
        bw = b.bandwidth()
    for k, m in enumerate(mats):
        assert starts[k] == want[k][0] and np.array_equal(perms[k], want[k][1]), k
        assert bw[k] == O.bandwidth(m.row_offsets, m.col_ids, want[k][1]), k
    big = P.workloads.config(2, scale=112, seed=44)  # 4 * (n + 1) + 4 * nnz > 32 MiB alone
    assert 4 * (big.rows + 1) + 4 * big.nnz > 32 << 20
    _say("large upload ctx")
    with P.RcmContext(timeout_ms=30000) as c:
        c.load(big)
        s, _ = c.order()
        assert np.array_equal(c.perm(), O.cm(big.row_offsets, big.col_ids, s))
        assert s == O.peripheral(big.row_offsets, big.col_ids)[1]
