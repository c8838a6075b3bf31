"""GPU: matrix files through the device pipeline (the reference's Graph<T>::load chain, include/Graph.h:266-324).

Each fixture (tests/golden/ingest_*.npz, written by tests/golden/make_ingest_golden.py from the reference's own
Graph<double>::load) holds a MatrixMarket file and the matrix the reference loads from it. The file is read on
the host (loadMTX + convert, src/COO.cpp:88-199) and canonicalised on the GPU (prepareRCM); the .hicsr cache
the host Graph::load writes is read and prepared on the GPU as well, and the .rcm.hicsr cache is read as is.
Every route must give the reference's matrix, values included; the GPU reorder of the result must equal the
oracle's order from the oracle's start.
"""
from __future__ import annotations

import os

import numpy as np
import pytest

import matrix_ref as M
import oracle_lib as O
import parallelbatchrcm_amd as P

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ["ingest_sym_grid", "ingest_pattern_sym", "ingest_general_nonsym"]


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def ops():
    o = P.MatrixOps()
    yield o
    o.close()


def _same(m, rp, ci, vals):
    np.testing.assert_array_equal(m.row_offsets, rp)
    np.testing.assert_array_equal(m.col_ids, ci)
    np.testing.assert_array_equal(np.asarray(m.data, np.float64), vals)


def _device_canonical(ops, m, check_sym):
    rows = M.rows_of(m.row_offsets).astype(np.int32)
    data = np.ones(m.nnz) if m.data is None else np.asarray(m.data, np.float64)
    return ops.canonicalize(m.rows, m.cols, rows, m.col_ids, data, check_sym)


@pytest.mark.parametrize("name", NAMES)
def test_file_ingest_through_the_gpu(tmp_path, ops, name):
    with np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    rp, ci, vals = g["row_ptr"], g["col_idx"], g["values"]
    path = tmp_path / "g.mtx"
    path.write_bytes(g["text"].tobytes())
    _say("ingest", name, "host Graph::load")
    loaded = P.Graph.load(path, np.float64)  # also writes g.mtxd_.hicsr and g.mtxd_.rcm.hicsr
    _same(loaded.csr, rp, ci, vals)
    _say("ingest", name, ".mtx -> host parse -> GPU canonicalise")
    m, _ = P.loadMTX(path, np.float64)
    canon, links = _device_canonical(ops, m, True)
    _same(canon, rp, ci, vals)
    assert links == loaded.links
    _say("ingest", name, ".hicsr -> GPU canonicalise")
    h = P.loadCSR(str(path) + "d_.hicsr", np.float64)
    canon2, _ = _device_canonical(ops, h, True)
    _same(canon2, rp, ci, vals)
    _say("ingest", name, ".rcm.hicsr as is, and through the GPU canonicalisation")
    r = P.loadCSR(str(path) + "d_.rcm.hicsr", np.float64)
    _same(r, rp, ci, vals)
    canon3, links3 = _device_canonical(ops, r, True)
    _same(canon3, rp, ci, vals)
    assert links3 == 0
    _say("ingest", name, "GPU reorder")
    rc, start, _, _ = O.peripheral(rp, ci)
    assert rc == 0
    want = O.cm(rp, ci, start)
    perm, bw, used = P.rcm_gpu_batch(canon, -1, P.RCM_FLAG_STABLE | P.RCM_FLAG_BANDWIDTH | P.RCM_FLAG_REVERSE)
    assert used == start
    assert np.array_equal(perm, want[::-1])
    assert bw == O.bandwidth(rp, ci, want)
