"""CPU tests (no GPU): the C-ABI library loads and exports every declared symbol; the workload generators and
the host canonicalisation are deterministic and agree with the oracle and the reference's prepareRCM."""
from __future__ import annotations

import os
import re

import numpy as np
import pytest

import graphs as G
import oracle_lib as O
import parallelbatchrcm_amd as P
from parallelbatchrcm_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(f for f in os.listdir(os.path.join(ROOT, "include")) if f.endswith(".h")):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(rcm_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    names = declared_functions()
    assert "rcm_gpu_batch" in names and "rcm_ctx_run" in names
    for n in names:
        assert hasattr(lib, n), n
        assert n in _native.exported_symbols(), f"{n} has no ctypes signature"


def test_context_create_without_gpu_fails_cleanly():
    import ctypes as C

    h = C.c_void_p()
    rc = _native.lib().rcm_ctx_create(0, C.byref(h))
    if rc == 0:  # a GPU is visible: nothing to check here
        _native.lib().rcm_ctx_destroy(h)
    else:
        assert rc == P.RCM_EDEVICE


def test_splitmix64_stream():
    import ctypes as C

    st = C.c_uint64(0)
    # published splitmix64 outputs for seed 0
    expect = [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    got = [_native.lib().rcm_splitmix64(C.byref(st)) for _ in range(3)]
    assert got == expect


def test_config_sizes():
    m1 = P.workloads.config(1, seed=1)
    assert (m1.rows, m1.nnz) == (10000, 49600)
    m2 = P.workloads.config(2, scale=32, seed=2)
    assert (m2.rows, m2.nnz) == (32 ** 3, 7 * 32 ** 3 - 6 * 32 ** 2)


def test_generators_deterministic():
    a = P.workloads.config(3, scale=20000, seed=9)
    b = P.workloads.config(3, scale=20000, seed=9)
    np.testing.assert_array_equal(a.row_offsets, b.row_offsets)
    np.testing.assert_array_equal(a.col_ids, b.col_ids)
    c = P.workloads.config(3, scale=20000, seed=10)
    assert not np.array_equal(a.col_ids, c.col_ids)


def _is_symmetric_sorted_with_diag(m):
    rp, ci = m.row_offsets, m.col_ids
    n = m.rows
    rows = [ci[rp[i]:rp[i + 1]] for i in range(n)]
    for i, r in enumerate(rows):
        assert np.all(np.diff(r) > 0), "sorted, no duplicates"
        assert i in r
    s = set()
    for i, r in enumerate(rows):
        for c in r:
            s.add((i, int(c)))
    assert all((c, i) in s for (i, c) in s)


def test_generated_matrices_are_symmetric_with_diagonal():
    _is_symmetric_sorted_with_diag(P.workloads.config(1, seed=3))
    _is_symmetric_sorted_with_diag(P.workloads.config(3, scale=3000, seed=3))
    _is_symmetric_sorted_with_diag(P.workloads.rmat(10, seed=3))


def test_permute_symmetric_is_a_relabelling():
    base = P.workloads.stencil2d(9, 7)
    b, q = P.workloads.permute_symmetric(base, 77)
    assert sorted(q.tolist()) == list(range(63))
    for i in range(63):
        got = sorted(b.col_ids[b.row_offsets[q[i]]:b.row_offsets[q[i] + 1]].tolist())
        want = sorted(q[base.col_ids[base.row_offsets[i]:base.row_offsets[i + 1]]].tolist())
        assert got == want


def test_rgg_mean_degree():
    m = P.workloads.rgg(200000, 12.0, seed=5)
    mean_off = (m.nnz - m.rows) / m.rows
    assert 11.0 < mean_off < 12.5  # slightly below 12 from the boundary of the unit square


def test_connect_matches_oracle_and_reference():
    rp, ci = G.from_edges(40, [(i, i + 1) for i in range(0, 39, 3)])
    m = P.CSR(rp, ci)
    got, added = P.connect_components(m)
    rc, orp, oci = O.prepare(rp, ci, False)
    assert added == rc
    np.testing.assert_array_equal(got.row_offsets, orp)
    np.testing.assert_array_equal(got.col_ids, oci)
    if O.ref_available():
        rrp, rci = O.ref_prepare(rp, ci, False)
        np.testing.assert_array_equal(got.col_ids, rci)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_symmetrize_then_connect_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    n = 150
    rows = [[i] for i in range(n)]
    for _ in range(200):
        a, b = rng.integers(0, n, 2)
        rows[a].append(int(b))  # one direction only: non-symmetric input
    for r in rows:
        r.sort()
    rp = np.zeros(n + 1, np.int32)
    for i in range(n):
        rp[i + 1] = rp[i] + len(rows[i])
    ci = np.array([c for r in rows for c in r], np.int32)
    g = P.Graph(P.CSR(rp, ci))
    added = g.prepareRCM(True)
    rc, orp, oci = O.prepare(rp, ci, True)
    assert added == rc
    np.testing.assert_array_equal(g.csr.row_offsets, orp)
    np.testing.assert_array_equal(g.csr.col_ids, oci)


def test_connect_hazard_is_an_error():
    rp = np.array([0, 1, 2, 2], np.int32)
    ci = np.array([0, 1], np.int32)
    with pytest.raises(P.RcmError):
        P.connect_components(P.CSR(rp, ci))


def test_batch_mix_layout():
    bs = P.workloads.batch_mix(6, 200, 3000, seed=4)
    assert len(bs) == 6
    for k in range(6):
        m = bs.matrix(k)
        assert m.rows == bs.n[k] and m.nnz == bs.nnz[k]
        assert m.row_offsets[0] == 0 and m.row_offsets[-1] == m.nnz
        assert O.cm(m.row_offsets, m.col_ids, 0).shape[0] == m.rows  # connected after repair
    bs2 = P.workloads.batch_mix(6, 200, 3000, seed=4, threads=1)
    np.testing.assert_array_equal(bs.col_idx, bs2.col_idx)  # independent of the thread count


def test_batch_create_without_gpu_fails_cleanly():
    import ctypes as C

    h = C.c_void_p()
    rc = _native.lib().rcm_batch_create(0, C.byref(h))
    if rc == 0:  # a GPU is visible: nothing to check here
        _native.lib().rcm_batch_destroy(h)
    else:
        assert rc == P.RCM_EDEVICE


def test_batchset_from_matrices_round_trip():
    mats = [P.workloads.stencil2d(5, 4), P.CSR(*G.path(7)), P.workloads.stencil3d(3, 3, 2)]
    bs = P.BatchSet.from_matrices(mats)
    assert len(bs) == 3
    for k, m in enumerate(mats):
        got = bs.matrix(k)
        assert np.array_equal(got.row_offsets, m.row_offsets) and np.array_equal(got.col_ids, m.col_ids)
    assert bs.perm_offsets.tolist() == [0, 20, 27]


def test_server_create_without_gpu_fails_cleanly():
    import ctypes as C

    h = C.c_void_p()
    cfg = _native.ServerConfigT(4, 2, 1 << 20, 0, 1000)
    rc = _native.lib().rcm_server_create(0, C.byref(cfg), C.byref(h))
    if rc == 0:
        _native.lib().rcm_server_destroy(h)
    else:
        assert rc == P.RCM_EDEVICE


def test_tune_result_mirror_layout():
    """rcm_tune_result is 40 bytes in the library (a static_assert there); the ctypes mirror must match."""
    import ctypes

    assert ctypes.sizeof(_native.TuneResultT) == 40
    assert _native.TuneResultT.helper_nodes.offset == 32
