"""GPU: the serving form (rcm_server): requests admitted while others run on the device, each result
bit-identical to the oracle and to the same matrix reordered alone."""
from __future__ import annotations

import time

import numpy as np
import pytest

import graphs as G
import oracle_lib as O
import parallelbatchrcm_amd as P

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _mix(count, seed):
    bs = P.workloads.batch_mix(count, n_lo=1000, n_hi=30000, seed=seed)
    mats = [bs.matrix(k) for k in range(len(bs))]
    mats.insert(3, P.connect_components(P.workloads.rmat(14, seed=seed))[0])
    mats.insert(7, P.CSR(*G.hub_graph(5000, 9000, seed=seed)))
    return mats


def test_server_results_match_oracle_in_the_mix():
    mats = _mix(20, 21)
    _say("server mix", len(mats))
    with P.RcmServer(slots=6, lanes=3, job_timeout_ms=20000) as srv:
        ids = [srv.submit(m, -1, P.RCM_FLAG_BANDWIDTH) for m in mats]  # all at once: more than the slots
        out = [srv.wait(i, 60000) for i in ids]
        st = srv.stats()
    assert st.completed == len(mats) and st.failed == 0
    assert st.max_inflight > 1  # requests overlapped on the device
    for m, (rc, perm, js) in zip(mats, out):
        assert rc == 0
        rc0, s, _, _ = O.peripheral(m.row_offsets, m.col_ids)
        want = O.cm(m.row_offsets, m.col_ids, s)
        assert js.start_node == s
        assert np.array_equal(perm, want)
        assert js.bandwidth == O.bandwidth(m.row_offsets, m.col_ids, want)


def test_server_same_result_alone_and_in_the_mix():
    """Each request's order alone (one at a time, nothing else on the GPU) equals its order in a mix that
    arrives while other requests run."""
    mats = _mix(10, 22)
    _say("server alone vs mix")
    with P.RcmServer(slots=4, lanes=2) as srv:
        alone = []
        for m in mats:
            rc, perm, _ = srv.wait(srv.submit(m, -1, P.RCM_FLAG_REVERSE), 60000)
            assert rc == 0
            alone.append(perm)
        ids = []
        for k, m in enumerate(mats):  # staggered arrivals
            ids.append(srv.submit(m, -1, P.RCM_FLAG_REVERSE))
            time.sleep(0.001 * (k % 3))
        mixed = [srv.wait(i, 60000) for i in ids]
    for k in range(len(mats)):
        assert mixed[k][0] == 0 and np.array_equal(mixed[k][1], alone[k]), k
    with P.RcmContext() as c:  # and the plain single-matrix path
        for k in (0, 3, 7):
            c.load(mats[k])
            s = c.find_start()
            c.run(s, P.RCM_FLAG_REVERSE)
            assert np.array_equal(c.perm(), alone[k])


def test_server_errors_are_per_request():
    rp, ci = G.from_edges(10, [(0, 1), (1, 2), (5, 6)])
    good = P.CSR(*G.grid2d(9, 7))
    bad_rp = np.array([0, 2, 1], np.int32)  # decreasing row offsets
    _say("server per-request errors")
    with P.RcmServer(slots=2, lanes=2) as srv:
        a = srv.submit(P.CSR(rp, ci), 0)
        b = srv.submit(P.CSR(bad_rp, np.array([0, 1], np.int32)), 0)
        c = srv.submit(good, 0)
        assert srv.wait(a, 30000)[0] == P.RCM_ENOTCONNECTED
        assert srv.wait(b, 30000)[0] == P.RCM_EINVAL
        rc, perm, _ = srv.wait(c, 30000)
        assert rc == 0 and np.array_equal(perm, O.cm(good.row_offsets, good.col_ids, 0))


def test_server_admission_by_arena_memory():
    """With one lane whose arena holds only a couple of requests at once, the others wait for memory and
    still finish: the slot limit (16) would have taken all 10 in one group."""
    mats = _mix(8, 23)
    need = max(4 * m.nnz + 44 * m.rows + 4096 for m in mats)
    arena = 2 * need + 8192
    _say("server small arena")
    with P.RcmServer(slots=16, lanes=1, arena_bytes=arena) as srv:
        ids = [srv.submit(m, 0) for m in mats]
        res = [srv.wait(i, 60000) for i in ids]
        st = srv.stats()
    assert st.completed == len(mats)
    assert st.loop_iterations >= 2
    assert st.arena_peak_bytes <= arena
    for m, (rc, perm, _) in zip(mats, res):
        assert rc == 0 and np.array_equal(perm, O.cm(m.row_offsets, m.col_ids, 0))


def test_server_borrowed_requests():
    """RCM_FLAG_BORROW: the server reads the caller's arrays in place. Requests whose only reference is the
    wrapper's (the arrays below are dropped at once) match the oracle, mixed with copied ones and a malformed
    borrowed one."""
    mats = _mix(12, 24)
    _say("server borrowed requests")
    with P.RcmServer(slots=5, lanes=2) as srv:
        ids = []
        for k, m in enumerate(mats):
            m2 = P.CSR(m.row_offsets.copy(), m.col_ids.copy())  # the server's reference is the only one left
            ids.append(srv.submit(m2, -1, P.RCM_FLAG_BANDWIDTH, borrow=k % 3 != 1))
            del m2
        bad = srv.submit(P.CSR(np.array([0, 2, 1], np.int32), np.array([0, 1], np.int32)), 0, borrow=True)
        out = [srv.wait(i, 60000) for i in ids]
        assert srv.wait(bad, 30000)[0] == P.RCM_EINVAL
        assert not srv._keep
    for m, (rc, perm, js) in zip(mats, out):
        assert rc == 0
        _, s, _, _ = O.peripheral(m.row_offsets, m.col_ids)
        want = O.cm(m.row_offsets, m.col_ids, s)
        assert js.start_node == s and np.array_equal(perm, want)
        assert js.bandwidth == O.bandwidth(m.row_offsets, m.col_ids, want)


def test_server_wait_limits_unknown_ids_and_shutdown_with_borrowed_requests():
    """rcm_server_wait: RCM_EAGAIN when the request has not finished within the limit (it stays queued and is
    waited for again), RCM_ENOTFOUND for an id the server does not hold (or held and released). Closing a server
    with borrowed requests still queued returns once the admitted groups finish; a new server then works."""
    big = P.connect_components(P.workloads.rmat(17, seed=5))[0]
    _say("server wait limits")
    with P.RcmServer(slots=2, lanes=1) as srv:
        jid = srv.submit(big, -1, P.RCM_FLAG_STABLE, borrow=True)
        try:
            rc, perm, _ = srv.wait(jid, 0)  # finished already only if the device was very quick
        except P.RcmError as e:
            assert e.status == P.RCM_EAGAIN
            rc, perm, _ = srv.wait(jid, 60000)
        assert rc == 0
        _, s, _, _ = O.peripheral(big.row_offsets, big.col_ids)
        assert np.array_equal(perm, O.cm(big.row_offsets, big.col_ids, s))
        lib = P._native.lib()
        assert lib.rcm_server_wait(srv._h, jid, 0, None, None) == P.RCM_ENOTFOUND  # released by the wait above
        assert lib.rcm_server_wait(srv._h, 1 << 40, 0, None, None) == P.RCM_ENOTFOUND
    _say("server shutdown with borrowed requests queued")
    mats = _mix(16, 25)
    srv = P.RcmServer(slots=2, lanes=1)
    for m in mats:
        srv.submit(P.CSR(m.row_offsets.copy(), m.col_ids.copy()), -1, P.RCM_FLAG_STABLE, borrow=True)
    srv.close()  # cancels what is queued, lets the admitted group finish
    assert not srv._keep
    with P.RcmServer(slots=4, lanes=2) as srv2:
        rc, perm, _ = srv2.wait(srv2.submit(mats[0], 0), 60000)
        assert rc == 0 and np.array_equal(perm, O.cm(mats[0].row_offsets, mats[0].col_ids, 0))


def test_server_large_requests_go_through_the_single_matrix_context():
    """Requests of 2^18 rows or more are ordered alone through each lane's single-matrix context: results equal the
    oracle's (reverse and bandwidth flags included, borrowed and copied), and a disconnected or malformed large
    request fails alone."""
    big = P.workloads.config(2, scale=64, seed=7)  # 64^3 = 2^18 rows
    assert big.rows == 1 << 18
    half = P.workloads.config(2, scale=52, seed=8)  # two copies: 281,216 rows
    n2 = half.rows
    two = P.CSR(np.concatenate([half.row_offsets, half.row_offsets[1:] + half.nnz]).astype(np.int32),
                np.concatenate([half.col_ids, half.col_ids + n2]).astype(np.int32))  # two components
    pad = P.workloads.config(2, scale=66, seed=9)  # > 2^18 rows
    bad_rp = pad.row_offsets.copy()
    bad_rp[1000] = bad_rp[1002]  # row offsets decrease at row 1001
    bad = P.CSR(bad_rp, pad.col_ids)
    _say("server large requests")
    with P.RcmServer(slots=4, lanes=2) as srv:
        assert srv.stats().solo_lanes == 2
        a = srv.submit(big, -1, P.RCM_FLAG_REVERSE | P.RCM_FLAG_BANDWIDTH, borrow=True)
        assert two.rows >= 1 << 18
        b = srv.submit(two, 0)
        c = srv.submit(bad, 0)
        d = srv.submit(big, 5)
        ra, rb, rc_, rd = (srv.wait(x, 60000) for x in (a, b, c, d))
    _, s, _, _ = O.peripheral(big.row_offsets, big.col_ids)
    want = O.cm(big.row_offsets, big.col_ids, s)
    assert ra[0] == 0 and ra[2].start_node == s and np.array_equal(ra[1], want[::-1])
    assert ra[2].bandwidth == O.bandwidth(big.row_offsets, big.col_ids, want)
    assert rb[0] == P.RCM_ENOTCONNECTED
    assert rc_[0] == P.RCM_EINVAL
    assert rd[0] == 0 and np.array_equal(rd[1], O.cm(big.row_offsets, big.col_ids, 5))
