"""Golden fixtures written from the reference's own CPU code (tests/golden/make_golden.py): the oracle (CPU)
and the engine (GPU) must reproduce them exactly. These need neither the reference tree nor its build."""
from __future__ import annotations

import glob
import os

import numpy as np
import pytest

import oracle_lib as O
import parallelbatchrcm_amd as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ORDERS = sorted(p for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                if not os.path.basename(p).startswith(("prepare", "ingest")))


def _load(path):
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def test_fixtures_present():
    assert len(ORDERS) >= 6


@pytest.mark.parametrize("path", ORDERS, ids=[os.path.basename(p)[:-4] for p in ORDERS])
def test_oracle_reproduces_reference_fixture(path):
    g = _load(path)
    rc, s, _, _ = O.peripheral(g["row_ptr"], g["col_idx"])
    assert rc == 0 and s == int(g["start"])
    perm = O.cm(g["row_ptr"], g["col_idx"], s)
    assert np.array_equal(perm, g["perm"])
    assert O.bandwidth(g["row_ptr"], g["col_idx"], perm) == int(g["bandwidth"])


def test_prepare_reproduces_reference_fixture():
    g = _load(os.path.join(GOLDEN, "prepare_nonsym_600.npz"))
    rc, rp, ci = O.prepare(g["row_ptr"], g["col_idx"], True)
    assert rc >= 0
    assert np.array_equal(rp, g["out_row_ptr"]) and np.array_equal(ci, g["out_col_idx"])
    m = P.Graph(P.CSR(g["row_ptr"], g["col_idx"]))
    m.prepareRCM(True)
    assert np.array_equal(m.csr.row_offsets, g["out_row_ptr"]) and np.array_equal(m.csr.col_ids, g["out_col_idx"])


@pytest.mark.gpu
def test_engine_reproduces_reference_fixtures(ctx):
    for path in ORDERS:
        g = _load(path)
        print("[case] golden", os.path.basename(path), flush=True)
        ctx.load(P.CSR(g["row_ptr"], g["col_idx"]))
        assert ctx.find_start() == int(g["start"])
        ctx.run(int(g["start"]))
        assert np.array_equal(ctx.perm(), g["perm"])
        assert ctx.bandwidth() == int(g["bandwidth"])


@pytest.mark.gpu
def test_batch_reproduces_reference_fixtures():
    gs = [_load(p) for p in ORDERS]
    bs = P.BatchSet.from_matrices([(g["row_ptr"], g["col_idx"]) for g in gs])
    print("[case] golden batch", len(gs), flush=True)
    with P.RcmBatch(timeout_ms=20000) as b:
        b.load(bs)
        starts, _, st = b.find_start()
        assert (st == 0).all()
        assert (b.run(starts) == 0).all()
        bw = b.bandwidth()
        perms = b.perms()
    for k, g in enumerate(gs):
        assert starts[k] == int(g["start"]) and np.array_equal(perms[k], g["perm"]) and bw[k] == int(g["bandwidth"])


@pytest.mark.parametrize("name", ["ingest_sym_grid", "ingest_pattern_sym", "ingest_general_nonsym"])
def test_ingest_fixtures_match_oracle_and_host_load(tmp_path, name):
    """The file-ingest fixtures (the reference's Graph<double>::load of each file) equal the oracle's
    prepareRCM of the parsed file and the library's host Graph::load."""
    g = _load(os.path.join(GOLDEN, f"{name}.npz"))
    path = tmp_path / "g.mtx"
    path.write_bytes(g["text"].tobytes())
    m, _ = P.loadMTX(path, np.float64)
    data = np.ones(m.nnz) if m.data is None else m.data
    added, rp, ci, vals = O.prepare_values(m.row_offsets, m.col_ids, data, True)
    assert added >= 0
    np.testing.assert_array_equal(rp, g["row_ptr"])
    np.testing.assert_array_equal(ci, g["col_idx"])
    np.testing.assert_array_equal(vals, g["values"])
    loaded = P.Graph.load(path, np.float64, write_cache=False)
    np.testing.assert_array_equal(loaded.csr.col_ids, g["col_idx"])
    np.testing.assert_array_equal(np.asarray(loaded.csr.data, np.float64), g["values"])
