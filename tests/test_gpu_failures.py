"""GPU: the error paths of the reorder kernels, forced from the host so that every kernel exits normally.

* RCM_ETIMEOUT: a 1 ms device watchdog on runs that need far longer. Every wait of the persistent kernels polls
  the watchdog; the first workgroup past its deadline records the error and marks the run done, and the others
  leave at their next poll.
* RCM_ECAPACITY: RCM_PARAM_BATCH_CAP leaves fewer batch records than the run needs, so the batch that would
  publish beyond them stops the run, the error path of the reference's oversize-batch __trap
  (include/device/batchWorker.cuh:539-543).

After each failure the same context reorders the same matrix again, bit-exact against the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle_lib as O
import parallelbatchrcm_amd as P

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _want(m):
    rc, s, _, _ = O.peripheral(m.row_offsets, m.col_ids)
    assert rc == 0
    return s, O.cm(m.row_offsets, m.col_ids, s)


def test_ctx_watchdog_timeout_then_reuse():
    m = P.workloads.config(2, seed=41)  # 128^3: 382 BFS levels, tens of ms on 8 workgroups
    s, want = _want(m)
    _say("ctx watchdog 1 ms, 8 workgroups, n", m.rows)
    with P.RcmContext(timeout_ms=1, workers=8) as c:
        c.load(m)
        assert c.run_status(s) == P.RCM_ETIMEOUT
        with pytest.raises(P.RcmError):
            c.perm()  # no result after a failed run
        c.set_param("timeout_ms", 30000)
        c.set_param("workers", 0)
        _say("ctx reuse after the time-out")
        c.run(s)
        assert np.array_equal(c.perm(), want)


@pytest.mark.parametrize("cap", [1, 2, 50])
def test_ctx_batch_capacity_then_reuse(cap):
    m = P.workloads.config(3, scale=20000, seed=42)
    s, want = _want(m)
    _say("ctx batch cap", cap, "n", m.rows)
    with P.RcmContext(timeout_ms=20000) as c:
        c.load(m)
        c.set_param("batch_cap", cap)
        assert c.get_param("batch_cap") == cap
        assert c.run_status(s) == P.RCM_ECAPACITY
        c.set_param("batch_cap", 0)
        _say("ctx reuse after the capacity error")
        c.run(s, P.RCM_FLAG_STABLE | P.RCM_FLAG_REVERSE)
        assert np.array_equal(c.perm(), want[::-1])
        assert c.bandwidth() == O.bandwidth(m.row_offsets, m.col_ids, want)
    with pytest.raises(P.RcmError):
        P.RcmContext(batch_cap=-1)


def test_batch_watchdog_timeout_then_reuse():
    bs = P.workloads.batch_mix(12, n_lo=20000, n_hi=200000, seed=43)
    _say("batch watchdog 1 ms, 4 workgroups", len(bs))
    with P.RcmBatch(timeout_ms=30000, graph=False) as b:
        b.load(bs)
        starts, _, st = b.find_start()
        assert (st == 0).all()
        b.set_param("timeout_ms", 1)
        b.set_param("workers", 4)
        rc, st = b.run_status(starts)
        assert rc == P.RCM_ETIMEOUT
        assert (st == P.RCM_ETIMEOUT).any()
        assert set(np.unique(st)) <= {P.RCM_OK, P.RCM_ETIMEOUT}  # finished ones keep their result
        b.set_param("timeout_ms", 30000)
        b.set_param("workers", 0)
        _say("batch reuse after the time-out")
        assert (b.run(starts) == 0).all()
        perms = b.perms()
    for k in range(len(bs)):
        m = bs.matrix(k)
        assert np.array_equal(perms[k], O.cm(m.row_offsets, m.col_ids, int(starts[k]))), k
