"""GPU: start search and order in one call (rcm_ctx_order, rcm_batch_order, and the one-shot entry points that use
them) against the oracle, bit-exact: the start and sweep count of SimplePeripheral::findPeripheral
(src/simple_peripheral.cpp:92-141) and the stable CM order of RealCPU::rcm_seq<true> (src/rcm_cpu.cpp:57-104)
from that start, for every speculation setting. The number of engine passes is checked against the schedule the
setting implies, so a speculated sweep that is wrongly taken as confirming (or wrongly rejected) fails."""
from __future__ import annotations

import numpy as np
import pytest

import graphs as G
import oracle_lib as O
import parallelbatchrcm_amd as P

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _expected_passes(iters: int, spec: int) -> int:
    """Engine passes of the fused call for a search that stops at sweep `iters`: every sweep from `spec` on is
    first tried as a CM run, and a failed try is followed by the plain sweep from the same start; the search
    ends with a CM run unless the stopping sweep was a tried one."""
    if iters <= 1:  # a single node: one CM run
        return 1
    if spec <= 0 or iters < spec:
        return iters + 1
    tried = iters - max(spec, 1) + 1  # sweeps spec..iters are tried; all but the last fail
    return (iters - tried) + 2 * (tried - 1) + 1


def _graphs():
    out = {
        "path17": P.CSR(*G.path(17)),
        "path2": P.CSR(*G.path(2)),
        "path1": P.CSR(*G.path(1)),
        "star40": P.CSR(*G.star(40)),
        "grid30x7": P.CSR(*G.grid2d(30, 7)),
        "random2000": P.CSR(*G.random_connected(2000, 5000, seed=9)),
        "duplicates": P.CSR(*G.with_duplicates(*G.random_connected(700, 1400, seed=3), seed=4)),
        "hub5000": P.CSR(*G.hub_graph(5000, 9000, seed=1)),
        "c1": P.workloads.config(1, seed=1),
        "c2_small": P.workloads.config(2, scale=24, seed=2),
        "c3_small": P.workloads.config(3, scale=40000, seed=3),
    }
    out["rmat14"] = P.connect_components(P.workloads.rmat(14, seed=14))[0]
    return out


GRAPHS = _graphs()


@pytest.fixture(scope="module")
def ctx():
    with P.RcmContext(timeout_ms=30000) as c:
        yield c


@pytest.mark.parametrize("spec", [0, 2, 3, 4])
@pytest.mark.parametrize("name", list(GRAPHS))
def test_ctx_order_matches_oracle(ctx, name, spec):
    m = GRAPHS[name]
    rc, start, iters, _ = O.peripheral(m.row_offsets, m.col_ids)
    assert rc == 0
    want = O.cm(m.row_offsets, m.col_ids, start)
    _say("ctx order", name, "speculate", spec)
    ctx.set_param("speculate", spec)
    ctx.load(m)
    s, it = ctx.order()
    assert (s, it) == (start, iters)
    assert np.array_equal(ctx.perm(), want)
    t = ctx.timing()
    assert t.start_iterations == iters
    assert t.order_passes == _expected_passes(iters, spec), (iters, spec, t.order_passes)
    assert ctx.bandwidth() == O.bandwidth(m.row_offsets, m.col_ids, want)
    ctx.set_param("speculate", 3)


def test_ctx_order_reverse_and_given_start(ctx):
    m = GRAPHS["c2_small"]
    rc, start, iters, _ = O.peripheral(m.row_offsets, m.col_ids)
    ctx.load(m)
    _say("ctx order reverse")
    s, it = ctx.order(flags=P.RCM_FLAG_STABLE | P.RCM_FLAG_REVERSE)
    assert (s, it) == (start, iters)
    assert np.array_equal(ctx.perm(), O.cm(m.row_offsets, m.col_ids, start)[::-1])
    _say("ctx order given start")
    s, it = ctx.order(start=7)
    assert (s, it) == (7, 0)
    assert ctx.timing().order_passes == 1
    assert np.array_equal(ctx.perm(), O.cm(m.row_offsets, m.col_ids, 7))


def test_ctx_order_errors(ctx):
    rp, ci = G.from_edges(10, [(0, 1), (1, 2), (5, 6)])
    ctx.load(P.CSR(rp, ci))
    _say("ctx order disconnected")
    for spec in (0, 1, 3):
        ctx.set_param("speculate", spec)
        with pytest.raises(P.RcmError) as e:
            ctx.order()
        assert e.value.status == P.RCM_ENOTCONNECTED
    ctx.set_param("speculate", 3)
    with pytest.raises(P.RcmError):
        ctx.order(start=10)


def test_one_shot_full_size_c2_and_c3():
    """rcm_gpu_batch (one-shot: the fused search and order) at the full sizes of the 128^3 stencil and the 1M-point
    geometric graph."""
    for which in (2, 3):
        m = P.workloads.config(which, seed=which)
        rc, start, iters, _ = O.peripheral(m.row_offsets, m.col_ids)
        _say("one-shot full size", which)
        perm, bw, used = P.rcm_gpu_batch(m, -1, P.RCM_FLAG_STABLE | P.RCM_FLAG_BANDWIDTH)
        assert used == start
        want = O.cm(m.row_offsets, m.col_ids, start)
        assert np.array_equal(perm, want)
        assert bw == O.bandwidth(m.row_offsets, m.col_ids, want)


@pytest.mark.parametrize("spec", [0, 3])
def test_batch_order_matches_oracle(spec):
    mix = P.workloads.batch_mix(24, n_lo=500, n_hi=20000, seed=8)
    mats = list(GRAPHS.values()) + [mix.matrix(k) for k in range(len(mix))]
    bs = P.BatchSet.from_matrices(mats)
    _say("batch order", len(bs), "speculate", spec)
    with P.RcmBatch(timeout_ms=30000) as b:
        b.set_param("speculate", spec)
        b.load(bs)
        starts, iters, st = b.order()
        assert (st == 0).all()
        perms = b.perms()
        for k, m in enumerate(mats):
            rc, s, it, _ = O.peripheral(m.row_offsets, m.col_ids)
            assert (starts[k], iters[k]) == (s, it), k
            assert np.array_equal(perms[k], O.cm(m.row_offsets, m.col_ids, s)), k
        bw = b.bandwidth()
        for k, m in enumerate(mats):
            assert bw[k] == O.bandwidth(m.row_offsets, m.col_ids, perms[k]), k
        # the same set through the two-call path gives the same starts and orders
        s2, it2, st2 = b.find_start()
        assert np.array_equal(s2, starts) and np.array_equal(it2, iters)
        b.run(s2)
        for k, p in enumerate(b.perms()):
            assert np.array_equal(p, perms[k]), k


def test_batch_order_given_starts_reverse_and_failing_member():
    rp, ci = G.from_edges(10, [(0, 1), (1, 2), (5, 6)])
    mats = [GRAPHS["grid30x7"], P.CSR(rp, ci), GRAPHS["hub5000"], GRAPHS["path1"], GRAPHS["c1"]]
    bs = P.BatchSet.from_matrices(mats)
    given = np.array([5, -1, -1, -1, 17], np.int32)
    _say("batch order given starts")
    with P.RcmBatch(timeout_ms=30000) as b:
        b.load(bs)
        starts, iters, st = b.order(given, flags=P.RCM_FLAG_STABLE | P.RCM_FLAG_REVERSE)
        assert st[1] == P.RCM_ENOTCONNECTED
        assert [st[k] for k in (0, 2, 3, 4)] == [0, 0, 0, 0]
        perms = b.perms()
        for k in (0, 2, 3, 4):
            m = mats[k]
            if given[k] >= 0:
                s = int(given[k])
                assert starts[k] == s and iters[k] == 0
            else:
                _, s, it, _ = O.peripheral(m.row_offsets, m.col_ids)
                assert (starts[k], iters[k]) == (s, it)
            assert np.array_equal(perms[k], O.cm(m.row_offsets, m.col_ids, s)[::-1]), k


def test_gpu_batch_many_uses_given_and_searched_starts():
    mats = [GRAPHS["c1"], GRAPHS["star40"], GRAPHS["random2000"]]
    bs = P.BatchSet.from_matrices(mats)
    _say("gpu_batch_many")
    perm, po, bw, st, rc = P.rcm_gpu_batch_many(bs, [3, -1, -1], P.RCM_FLAG_STABLE | P.RCM_FLAG_BANDWIDTH)
    assert rc == 0 and (st == 0).all()
    for k, m in enumerate(mats):
        s = 3 if k == 0 else O.peripheral(m.row_offsets, m.col_ids)[1]
        want = O.cm(m.row_offsets, m.col_ids, s)
        assert np.array_equal(perm[po[k]:po[k] + m.rows], want), k
        assert bw[k] == O.bandwidth(m.row_offsets, m.col_ids, want)
