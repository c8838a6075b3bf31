"""GPU check of the wave64 and workgroup primitives the engines are built on (rcm_selftest_primitives): block
exclusive scan and total, wave sum / max / min and inclusive scan, against numpy on several input patterns."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import parallelbatchrcm_amd as P
from parallelbatchrcm_amd import _native

pytestmark = pytest.mark.gpu

PATTERNS = {
    "ones": np.ones(256, np.uint32),
    "iota": np.arange(256, dtype=np.uint32),
    "random": np.random.default_rng(1).integers(0, 1 << 20, 256).astype(np.uint32),
    "sparse": (np.random.default_rng(2).random(256) < 0.1).astype(np.uint32) * 7,
    "large": np.random.default_rng(3).integers(0, 1 << 23, 256).astype(np.uint32),
}


@pytest.mark.parametrize("name", list(PATTERNS))
def test_primitives_match_numpy(name):
    a = PATTERNS[name]
    out = np.zeros(6 * 256, np.uint32)
    print("[case] primitives", name, flush=True)
    rc = _native.lib().rcm_selftest_primitives(0, a.ctypes.data_as(C.POINTER(C.c_uint32)),
                                               out.ctypes.data_as(C.POINTER(C.c_uint32)))
    assert rc == P.RCM_OK
    a64 = a.astype(np.uint64)
    np.testing.assert_array_equal(out[:256], (np.cumsum(a64) - a64).astype(np.uint32))
    assert np.all(out[256:512] == np.uint32(a64.sum() & 0xFFFFFFFF))
    w = a64.reshape(4, 64)
    np.testing.assert_array_equal(out[512:768], np.repeat(w.sum(1), 64).astype(np.uint32))
    np.testing.assert_array_equal(out[768:1024], np.repeat(w.max(1), 64).astype(np.uint32))
    np.testing.assert_array_equal(out[1024:1280], np.repeat(w.min(1), 64).astype(np.uint32))
    np.testing.assert_array_equal(out[1280:1536], np.cumsum(w, axis=1).reshape(-1).astype(np.uint32))


def test_copy_bandwidth_is_plausible():
    """rcm_selftest_copy_bandwidth: a plain device copy reaches a bandwidth between 5% of the MI355X HBM peak and
    the peak itself (with 10% allowance for the clock of the event timer); malformed sizes are rejected on the
    host."""
    print("[case] copy bandwidth", flush=True)
    for runtime in (False, True):
        gbs, ms = P.copy_bandwidth(1 << 28, 3, runtime=runtime)
        assert 400.0 < gbs < 8800.0, (runtime, gbs)
        assert abs(2 * (1 << 28) / (ms * 1e-3) / 1e9 - gbs) < 1e-6 * gbs
    for bad in (0, 15, (1 << 20) + 8):
        with pytest.raises(P.RcmError):
            P.copy_bandwidth(bad, 1)
