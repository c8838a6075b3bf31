"""GPU parity: the HIP engine (through the C ABI) against the oracle, bit-exact.

Every case prints its name before it touches the GPU so a failing run names it.
"""
from __future__ import annotations

import numpy as np
import pytest

import graphs as G
import oracle_lib as O
import parallelbatchrcm_amd as P

pytestmark = pytest.mark.gpu


def _csr(rp, ci):
    return P.CSR(rp, ci)


def _say(*a):
    print("[case]", *a, flush=True)


SMALL = {
    "path17": G.path(17),
    "path1": G.path(1),
    "path2_nodiag": G.path(2, diag=False),
    "star40": G.star(40),
    "complete9": G.complete(9),
    "grid7x5": G.grid2d(7, 5),
    "rand60": G.random_connected(60, 90, seed=3),
    "rand500": G.random_connected(500, 1500, seed=4),
    "dups300": G.with_duplicates(*G.random_connected(300, 600, seed=5), seed=6),
    "unsorted300": G.shuffle_rows(*G.random_connected(300, 900, seed=7), seed=8),
}


@pytest.mark.parametrize("name", list(SMALL))
def test_small_graphs_all_starts(ctx, name):
    rp, ci = SMALL[name]
    n = rp.shape[0] - 1
    _say(name, n)
    ctx.load(_csr(rp, ci))
    starts = range(n) if n <= 64 else np.random.default_rng(1).choice(n, 8, replace=False)
    for s in starts:
        ref = O.cm(rp, ci, int(s))
        assert ref.shape[0] == n
        ctx.run(int(s))
        got = ctx.perm()
        np.testing.assert_array_equal(got, ref, err_msg=f"{name} start {s}")


@pytest.mark.parametrize("which,scale", [(1, 0), (2, 12), (2, 24), (3, 5000), (3, 40000)])
def test_configs_scaled(ctx, which, scale):
    m = P.workloads.config(which, scale=scale, seed=100 + which)
    _say("config", which, scale, m.rows, m.nnz)
    rc, start, _, _ = O.peripheral(m.row_offsets, m.col_ids)
    assert rc == 0
    ref = O.cm(m.row_offsets, m.col_ids, start)
    ctx.load(m)
    ctx.run(start)
    got = ctx.perm()
    np.testing.assert_array_equal(got, ref)
    assert ctx.bandwidth() == O.bandwidth(m.row_offsets, m.col_ids, ref)


@pytest.mark.parametrize("which,scale", [(1, 0), (2, 20), (3, 20000)])
def test_start_search_matches_oracle(ctx, which, scale):
    m = P.workloads.config(which, scale=scale, seed=200 + which)
    _say("peripheral", which, scale)
    rc, start, iters, _ = O.peripheral(m.row_offsets, m.col_ids)
    assert rc == 0
    ctx.load(m)
    assert ctx.find_start() == start
    assert ctx.timing().start_iterations == iters


def test_reverse_and_one_shot():
    m = P.workloads.config(1, seed=300)
    _say("one-shot reverse")
    rc, start, _, _ = O.peripheral(m.row_offsets, m.col_ids)
    ref = O.cm(m.row_offsets, m.col_ids, start)
    perm, bw, used = P.rcm_gpu_batch(m, -1, P.RCM_FLAG_REVERSE | P.RCM_FLAG_BANDWIDTH)
    assert used == start
    np.testing.assert_array_equal(perm, ref[::-1])
    assert bw == O.bandwidth(m.row_offsets, m.col_ids, ref) == O.bandwidth(m.row_offsets, m.col_ids, ref[::-1])


def test_disconnected_is_an_error(ctx):
    rp, ci = G.from_edges(10, [(0, 1), (1, 2), (5, 6)])
    _say("disconnected")
    ctx.load(_csr(rp, ci))
    assert ctx.run_status(0) == P.RCM_ENOTCONNECTED


def test_malformed_csr_rejected(ctx):
    _say("malformed")
    rp = np.array([0, 2, 3], np.int32)
    ci = np.array([0, 5, 1], np.int32)  # column 5 out of range
    with pytest.raises(P.RcmError) as e:
        ctx.load(_csr(rp, ci))
    assert e.value.status == P.RCM_EINVAL


@pytest.mark.parametrize("workers", [1, 3, 16, 64])
def test_worker_counts_and_strict_mode_agree(ctx, workers):
    m = P.workloads.config(2, scale=20, seed=400)
    _say("workers", workers)
    rc, start, _, _ = O.peripheral(m.row_offsets, m.col_ids)
    ref = O.cm(m.row_offsets, m.col_ids, start)
    ctx.load(m)
    try:
        for strict in (0, 1):
            ctx.set_param("workers", workers)
            ctx.set_param("strict", strict)
            ctx.run(start)
            np.testing.assert_array_equal(ctx.perm(), ref, err_msg=f"workers={workers} strict={strict}")
    finally:
        ctx.set_param("workers", 0)
        ctx.set_param("strict", 0)


@pytest.mark.parametrize("sort", [1, 2])
def test_sort_modes_agree(ctx, sort):
    m = P.workloads.config(3, scale=30000, seed=500)
    _say("sort", sort)
    rc, start, _, _ = O.peripheral(m.row_offsets, m.col_ids)
    ref = O.cm(m.row_offsets, m.col_ids, start)
    ctx.load(m)
    try:
        ctx.set_param("sort", sort)
        ctx.run(start)
        np.testing.assert_array_equal(ctx.perm(), ref)
    finally:
        ctx.set_param("sort", 0)


@pytest.mark.slow
@pytest.mark.parametrize("which", [2, 3])
def test_full_size_configs(ctx, which):
    m = P.workloads.config(which, seed=which)
    _say("full config", which, m.rows, m.nnz)
    rc, start, _, _ = O.peripheral(m.row_offsets, m.col_ids)
    assert rc == 0
    ref = O.cm(m.row_offsets, m.col_ids, start)
    ctx.load(m)
    for rep in range(3):
        ctx.run(start)
        np.testing.assert_array_equal(ctx.perm(), ref, err_msg=f"rep {rep}")
    assert ctx.bandwidth() == O.bandwidth(m.row_offsets, m.col_ids, ref)
    assert ctx.find_start() == start


def test_context_reuse_across_sizes(ctx):
    """Batch records written by a run on a larger matrix must not leak into a later, smaller one."""
    big = P.workloads.config(3, scale=60000, seed=600)
    tiny_rp, tiny_ci = G.from_edges(6, [(0, 1), (1, 2), (3, 4)])
    mid = P.workloads.config(2, scale=22, seed=601)
    _say("reuse big -> tiny -> mid")
    rc, sb, _, _ = O.peripheral(big.row_offsets, big.col_ids)
    ctx.load(big)
    ctx.set_param("workers", 128)
    try:
        ctx.run(sb)
        np.testing.assert_array_equal(ctx.perm(), O.cm(big.row_offsets, big.col_ids, sb))
        ctx.load(_csr(tiny_rp, tiny_ci))
        assert ctx.run_status(0) == P.RCM_ENOTCONNECTED
        rc, sm, _, _ = O.peripheral(mid.row_offsets, mid.col_ids)
        ref = O.cm(mid.row_offsets, mid.col_ids, sm)
        ctx.load(mid)
        for _ in range(3):
            ctx.run(sm)
            np.testing.assert_array_equal(ctx.perm(), ref)
    finally:
        ctx.set_param("workers", 0)


def _rmat(scale, seed):
    m = P.workloads.rmat(scale, seed=seed)
    m, _ = P.connect_components(m)
    return m


LONG = {
    "star3000": lambda: P.CSR(*G.star(3000)),
    "hub5000": lambda: P.CSR(*G.hub_graph(5000, 9000, seed=1)),
    "twohubs2600": lambda: P.CSR(*G.hub_graph(2600, 3000, seed=2, hubs=2)),
    "hubdups": lambda: P.CSR(*G.with_duplicates(*G.hub_graph(2500, 2000, seed=3), seed=4)),
    "rmat14": lambda: _rmat(14, 5),
    "rmat16": lambda: _rmat(16, 6),
}


@pytest.mark.parametrize("name", list(LONG))
def test_long_rows_match_oracle(ctx, name):
    """Rows longer than a batch's 2048 adjacency slots (hubs): the long-row path, bit-exact."""
    m = LONG[name]()
    deg = np.diff(m.row_offsets)
    assert deg.max() > 2048
    rc, ps, _, _ = O.peripheral(m.row_offsets, m.col_ids)
    assert rc == 0
    for start in sorted({int(np.argmax(deg)), int(np.argmin(deg)), ps}):
        _say("long rows", name, "start", start)
        ctx.load(m)
        ctx.run(start, P.RCM_FLAG_STABLE)
        assert np.array_equal(ctx.perm(), O.cm(m.row_offsets, m.col_ids, start)), (name, start)
    _say("long rows start search", name)
    assert ctx.find_start() == ps
    ctx.run(ps)
    want = O.cm(m.row_offsets, m.col_ids, ps)
    assert ctx.bandwidth() == O.bandwidth(m.row_offsets, m.col_ids, want)


@pytest.mark.parametrize("flt", [1, 2])
def test_claim_filter_modes_agree(ctx, flt):
    """The claim pre-read filter skips atomics on already-claimed nodes; the order must not change."""
    cases = [P.workloads.config(3, scale=20000, seed=9), P.CSR(*G.hub_graph(5000, 9000, seed=1))]
    try:
        ctx.set_param("claim_filter", flt)
        for m in cases:
            rc, s, _, _ = O.peripheral(m.row_offsets, m.col_ids)
            _say("claim filter", flt, m.rows)
            ctx.load(m)
            ctx.run(s)
            assert np.array_equal(ctx.perm(), O.cm(m.row_offsets, m.col_ids, s))
    finally:
        ctx.set_param("claim_filter", 0)


@pytest.mark.parametrize("variant", [1, 2, 3, 4, 5])
def test_engine_variants_match_oracle(ctx, variant):
    """Every engine variant (block size, batch width, radix digit, child-batch degree quantum) gives the same
    order; the star's centre (degree 799) and the hubs exceed variant 3's degree quantum of 512."""
    cases = [P.CSR(*G.random_connected(500, 1500, seed=4)), P.workloads.config(2, scale=24, seed=2),
             P.CSR(*G.hub_graph(5000, 9000, seed=1)), _rmat(14, 5), P.workloads.config(3, scale=40000, seed=3),
             P.CSR(*G.star(800))]
    keep = ctx.get_param("variant")
    try:
        ctx.set_param("variant", variant)
        for m in cases:
            rc, s, _, _ = O.peripheral(m.row_offsets, m.col_ids)
            _say("variant", variant, m.rows)
            ctx.load(m)
            assert ctx.find_start() == s
            ctx.run(s)
            assert np.array_equal(ctx.perm(), O.cm(m.row_offsets, m.col_ids, s))
    finally:
        ctx.set_param("variant", keep)


def test_default_variant_is_the_tuned_one():
    """A fresh context runs variant 1 (512 threads over 2048 slots), the one its autotuner picks for the 128^3
    stencil, the geometric graph and R-MAT on MI355X."""
    with P.RcmContext() as c:
        assert c.get_param("variant") == 1


def test_autotune_keeps_results(ctx):
    m = P.workloads.config(3, scale=40000, seed=3)
    rc, s, _, _ = O.peripheral(m.row_offsets, m.col_ids)
    _say("autotune ctx", m.rows)
    ctx.load(m)
    keep = ctx.get_param("variant")
    try:
        r = ctx.autotune(s, repeats=2)
        assert 0 <= r.variant <= 5 and r.candidates >= 11
        assert r.workers >= 1 if r.variant < 5 else r.workers == 0
        assert 0 < r.best_ms <= r.default_ms
        ctx.run(s)
        assert np.array_equal(ctx.perm(), O.cm(m.row_offsets, m.col_ids, s))
    finally:
        ctx.set_param("variant", keep)
        ctx.set_param("workers", 0)


def test_level_synchronous_variant(ctx):
    """Variant 5, the level-synchronous engine: small graphs from several starts (duplicates, unsorted rows),
    hub graphs, R-MAT, the start search, reverse and a disconnected graph."""
    keep = ctx.get_param("variant")
    try:
        ctx.set_param("variant", 5)
        for name, (rp, ci) in SMALL.items():
            n = rp.shape[0] - 1
            ctx.load(_csr(rp, ci))
            for st in sorted({0, n // 2, n - 1}):
                _say("level variant", name, st)
                ctx.run(st)
                np.testing.assert_array_equal(ctx.perm(), O.cm(rp, ci, st), err_msg=f"{name} {st}")
        for name in ("hub5000", "hubdups", "rmat14", "rmat16"):
            m = LONG[name]()
            rc, ps, _, _ = O.peripheral(m.row_offsets, m.col_ids)
            _say("level variant", name)
            ctx.load(m)
            assert ctx.find_start() == ps
            ctx.run(ps, P.RCM_FLAG_REVERSE)
            want = O.cm(m.row_offsets, m.col_ids, ps)
            np.testing.assert_array_equal(ctx.perm(), want[::-1])
            assert ctx.bandwidth() == O.bandwidth(m.row_offsets, m.col_ids, want)
        rp, ci = G.path(10)
        rp2 = np.concatenate([rp, rp[1:] + rp[-1]])
        ci2 = np.concatenate([ci, ci + 10])
        _say("level variant disconnected")
        ctx.load(_csr(rp2, ci2))
        assert ctx.run_status(0) == P.RCM_ENOTCONNECTED
    finally:
        ctx.set_param("variant", keep)


@pytest.mark.parametrize("name,args", [("directed3000", (3000, 3000, 40, 1)), ("directed2000", (2000, 1500, 24, 2))])
def test_directed_graph_with_empty_rows(ctx, name, args):
    """Rows are out-neighbour lists and half the nodes have empty rows: batches mix parents without adjacency
    entries and parents with up to 40, through both sort paths."""
    rp, ci = G.directed_with_empty_rows(*args)
    _say(name, rp.shape[0] - 1)
    ctx.load(_csr(rp, ci))
    ref = O.cm(rp, ci, 0)
    try:
        for sort in (0, 1, 2):
            ctx.set_param("sort", sort)
            ctx.run(0)
            np.testing.assert_array_equal(ctx.perm(), ref, err_msg=f"{name} sort {sort}")
    finally:
        ctx.set_param("sort", 0)
