"""GPU: rcm_gpu_batch_noalloc, the one-shot reorder that allocates nothing, through the C ABI.

Its output (order, start, bandwidth) must equal rcm_gpu_batch's on the same matrix -- the existing entry point
it is a variant of -- and the oracle's. Two calls in a row share the device's fixed array; the second is
smaller, so it runs over what the first left behind. Sizes beyond the array's maximum are rejected.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import parallelbatchrcm_amd as P
from hip_buffers import HipBuffers
from parallelbatchrcm_amd import _native

pytestmark = pytest.mark.gpu

MAX_N = 1 << 21


def _say(*a):
    print("[case]", *a, flush=True)


def _noalloc(hip, m, start, flags):
    lib = _native.lib()
    d_rp, d_ci = hip.upload(m.row_offsets), hip.upload(m.col_ids)
    d_perm = hip.alloc(4 * m.rows)
    bw, used = C.c_int64(-1), C.c_int32(-1)
    rc = lib.rcm_gpu_batch_noalloc(m.rows, m.nnz, d_rp, d_ci, int(start), flags, d_perm, C.byref(bw), C.byref(used))
    perm = hip.download(d_perm, m.rows, np.int32) if rc == 0 else None
    return rc, perm, bw.value, used.value


@pytest.mark.parametrize("first,second", [((2, 0, 11), (3, 20000, 12)), ((3, 40000, 13), (1, 0, 14))])
def test_noalloc_equals_rcm_gpu_batch_twice(first, second):
    flags = P.RCM_FLAG_STABLE | P.RCM_FLAG_BANDWIDTH | P.RCM_FLAG_REVERSE
    with HipBuffers() as hip:
        for which, scale, seed in (first, second):
            m = P.workloads.config(which, scale=scale, seed=seed)
            _say("noalloc config", which, scale, "n", m.rows)
            rc, perm, bw, used = _noalloc(hip, m, -1, flags)
            assert rc == 0
            want, wbw, wused = P.rcm_gpu_batch(m, -1, flags)
            assert used == wused
            assert np.array_equal(perm, want)
            assert bw == wbw
            orc_rc, s, _, _ = O.peripheral(m.row_offsets, m.col_ids)
            assert orc_rc == 0 and used == s
            assert np.array_equal(perm, O.cm(m.row_offsets, m.col_ids, s)[::-1])
            # a given start, CM order, no bandwidth
            _say("noalloc given start", which, scale)
            rc, perm2, _, used2 = _noalloc(hip, m, s, P.RCM_FLAG_STABLE)
            assert rc == 0 and used2 == s
            assert np.array_equal(perm2, want[::-1])


def test_noalloc_errors():
    lib = _native.lib()
    with HipBuffers() as hip:
        m = P.workloads.config(1, seed=15)
        d_rp, d_ci = hip.upload(m.row_offsets), hip.upload(m.col_ids)
        d_perm = hip.alloc(4 * m.rows)
        _say("noalloc errors")
        # beyond the fixed array: rejected before any device work
        assert lib.rcm_gpu_batch_noalloc(MAX_N + 1, m.nnz, d_rp, d_ci, 0, 0, d_perm, None, None) == P.RCM_ECAPACITY
        assert lib.rcm_gpu_batch_noalloc(m.rows, m.nnz, d_rp, d_ci, m.rows, 0, d_perm, None, None) == P.RCM_EINVAL
        assert lib.rcm_gpu_batch_noalloc(m.rows, m.nnz, d_rp, d_ci, 0, 0, None, None, None) == P.RCM_EINVAL
        bad = m.col_ids.copy()
        bad[7] = m.rows  # column out of range: caught by the device validation
        d_bad = hip.upload(bad)
        assert lib.rcm_gpu_batch_noalloc(m.rows, m.nnz, d_rp, d_bad, 0, 0, d_perm, None, None) == P.RCM_EINVAL
        # two components: not connected
        two = P.CSR(np.array([0, 1, 2], np.int32), np.array([0, 1], np.int32))
        d_rp2, d_ci2 = hip.upload(two.row_offsets), hip.upload(two.col_ids)
        assert lib.rcm_gpu_batch_noalloc(2, 2, d_rp2, d_ci2, 0, 0, d_perm, None, None) == P.RCM_ENOTCONNECTED
        # and it still works afterwards
        rc, perm, _, used = _noalloc(hip, m, -1, P.RCM_FLAG_STABLE)
        assert rc == 0
        assert np.array_equal(perm, O.cm(m.row_offsets, m.col_ids, used))
