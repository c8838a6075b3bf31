"""GPU parity of the steps either side of the reorder (include/rcm_matrix_ops.h), through the C ABI:

* canonicalisation (COO -> CSR + prepareRCM) against the oracle's convert and value-carrying prepareRCM
  (oracle/rcm_oracle.c, pinned to the reference's own Graph<double>::prepareRCM in tests/test_oracle.py) --
  bit-exact, values included;
* B = P A P^T against its numpy definition (tests/matrix_ref.py) -- bit-exact, values included;
* the whole device pipeline (COO in HBM -> canonicalise -> start search -> reorder -> permute) against the host
  pipeline built from the oracle.

Every case prints its name before it touches the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import matrix_ref as M
from hip_buffers import HipBuffers
import oracle_lib as O
import parallelbatchrcm_amd as P
from parallelbatchrcm_amd import _native

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def ops():
    o = P.MatrixOps()
    yield o
    o.close()


def oracle_canonical(n, r, c, v, check_sym):
    """The oracle's convert + prepareRCM with values: (CSR, links)."""
    rp, ci, vals = O.convert(n, r, c, v)
    added, orp, oci, ov = O.prepare_values(rp, ci, vals, check_sym)
    assert added >= 0
    return P.CSR(orp, oci, ov), added


def long_row_matrix(rng, n, k, hubs=2, hub_len=3000):
    r = rng.integers(0, n, k)
    c = rng.integers(0, n, k)
    if hubs:  # hub rows and a hub column
        hr = np.repeat(rng.choice(n, hubs, replace=False), hub_len)
        r = np.concatenate([r, hr, rng.integers(0, n, hub_len)])
        c = np.concatenate([c, rng.integers(0, n, hr.size), np.full(hub_len, min(3, n - 1))])
    return P.convert_coo(n, n, r, c, rng.standard_normal(r.size))


@pytest.mark.parametrize("n,k,vals", [(1, 1, True), (97, 400, True), (5000, 40000, True), (60000, 500000, False),
                                      (200000, 1500000, True)])
def test_permute_matches_definition(ops, n, k, vals):
    rng = np.random.default_rng(n + k)
    m = long_row_matrix(rng, n, k, hubs=2 if n > 100 else 0, hub_len=min(3000, 4 * n))
    if not vals:
        m.data = None
    perm = rng.permutation(n).astype(np.int32)
    _say("permute", n, m.nnz, vals)
    b = ops.permute(m, perm)
    erp, eci, ev = M.permute_ref(m.row_offsets, m.col_ids, m.data, perm)
    np.testing.assert_array_equal(b.row_offsets, erp)
    np.testing.assert_array_equal(b.col_ids, eci)
    if vals:
        np.testing.assert_array_equal(b.data, ev)
    else:
        assert b.data is None


def test_permute_errors(ops):
    rng = np.random.default_rng(3)
    m = long_row_matrix(rng, 300, 900, hubs=1, hub_len=100)
    perm = rng.permutation(300).astype(np.int32)
    bad = perm.copy()
    bad[1] = bad[2]
    _say("permute errors")
    with pytest.raises(P.RcmError) as e:
        ops.permute(m, bad)
    assert e.value.status == P.RCM_EINVAL
    bad = perm.copy()
    bad[0] = -1
    with pytest.raises(P.RcmError):
        ops.permute(m, bad)
    broken = P.CSR(m.row_offsets.copy(), m.col_ids.copy(), m.data)
    broken.col_ids[5] = 300
    with pytest.raises(P.RcmError):
        ops.permute(broken, perm)
    # the handle still works after the errors
    b = ops.permute(m, perm)
    np.testing.assert_array_equal(b.col_ids, M.permute_ref(m.row_offsets, m.col_ids, None, perm)[1])


@pytest.mark.parametrize("n,comps,check_sym", [(50, 1, True), (3000, 4, True), (3000, 4, False),
                                               (60000, 30, True), (250000, 200, True)])
def test_canonicalize_matches_host(ops, n, comps, check_sym):
    rng = np.random.default_rng(n + comps)
    r, c = M.coo_with_components(rng, n, 4 * n, comps=comps)
    v = rng.standard_normal(r.size)
    if not check_sym:
        sym = P.convert_coo(n, n, np.concatenate([r, c]), np.concatenate([c, r]), np.concatenate([v, v]))
        r, c, v = M.rows_of(sym.row_offsets).astype(np.int32), sym.col_ids, sym.data
    _say("canonicalize", n, r.size, comps, check_sym)
    got, links = ops.canonicalize(n, n, r, c, v, check_sym)
    host, hlinks = oracle_canonical(n, r, c, v, check_sym)
    assert links == hlinks
    np.testing.assert_array_equal(got.row_offsets, host.row_offsets)
    np.testing.assert_array_equal(got.col_ids, host.col_ids)
    np.testing.assert_array_equal(got.data, host.data)


def directed_coo(rng, n, k, cycles):
    """Random one-sided entries, a full diagonal and a few directed cycles."""
    r = np.concatenate([rng.integers(0, n, k), np.arange(n)])
    c = np.concatenate([rng.integers(0, n, k), np.arange(n)])
    for _ in range(cycles):
        nodes = rng.choice(n, int(rng.integers(2, 6)), replace=False)
        r = np.concatenate([r, nodes])
        c = np.concatenate([c, np.roll(nodes, -1)])
    return r.astype(np.int32), c.astype(np.int32)


@pytest.mark.parametrize("n,check_sym", [(300, False), (5000, True), (5000, False), (60000, False)])
def test_canonicalize_directed_matches_oracle(ops, n, check_sym):
    """Matrices that stay structurally non-symmetric (one-sided entries without the symmetrisation; pairs
    stored twice in one direction with it): the device repair follows the stored rows like the reference's walk
    (include/Graph.h:160-263), against the oracle's restatement of it."""
    rng = np.random.default_rng(n + check_sym)
    r, c = directed_coo(rng, n, n // 2, cycles=20)
    if check_sym:
        r, c = np.concatenate([r, r[:11]]), np.concatenate([c, c[:11]])
    v = rng.standard_normal(r.size)
    _say("canonicalize directed", n, r.size, check_sym)
    got, links = ops.canonicalize(n, n, r, c, v, check_sym)
    want, wlinks = oracle_canonical(n, r, c, v, check_sym)
    assert links == wlinks and links > 0
    np.testing.assert_array_equal(got.row_offsets, want.row_offsets)
    np.testing.assert_array_equal(got.col_ids, want.col_ids)
    np.testing.assert_array_equal(got.data, want.data)


def test_canonicalize_errors(ops):
    _say("canonicalize errors")
    # a pair stored twice in one direction stays non-symmetric after the parity pass: the directed repair runs
    r = np.array([0, 0, 0, 1, 2, 1], np.int32)
    c = np.array([1, 1, 0, 1, 2, 2], np.int32)
    got, links = ops.canonicalize(3, 3, r, c)
    assert links == 0 and got.col_ids.tolist() == [0, 1, 1, 1, 2, 1, 2]
    # a node that nothing smaller reaches and that lies on no cycle: the reference's walk never ends
    with pytest.raises(P.RcmError) as e:
        ops.canonicalize(3, 3, [0, 2], [0, 0], check_sym=False)
    assert e.value.status == P.RCM_EINVAL
    with pytest.raises(P.RcmError) as e:
        ops.canonicalize(3, 3, [0, 0, 1], [0, 1, 0])
    assert e.value.status == P.RCM_EINVAL
    with pytest.raises(P.RcmError) as e:
        ops.canonicalize(3, 3, [0, 5], [0, 0])
    assert e.value.status == P.RCM_EINVAL
    got, links = ops.canonicalize(3, 3, [2, 1, 0], [2, 1, 0])
    assert links == 2


def test_device_pipeline_matches_host(ops):
    """COO in HBM -> canonicalise -> start search -> reorder (RCM) -> B = P A P^T, with no host round trip between
    the steps; compared with the oracle's convert + prepareRCM + CM and the permutation definition."""
    rng = np.random.default_rng(21)
    side = 40
    grid = P.workloads.stencil2d(side, side)
    n = grid.rows
    rows = M.rows_of(grid.row_offsets)
    lower = rows >= grid.col_ids  # MatrixMarket "symmetric" storage: the lower triangle
    r = rows[lower].astype(np.int32)
    c = grid.col_ids[lower]
    # three nodes cut loose become islands: the repair adds a link for each
    islands = rng.choice(n, 3, replace=False)
    keep = ~(np.isin(r, islands) | np.isin(c, islands)) | (r == c)
    r, c = r[keep], c[keep]
    v = rng.standard_normal(r.size)
    shuffle = rng.permutation(r.size)
    r, c, v = r[shuffle], c[shuffle], v[shuffle]
    _say("device pipeline", n, r.size)
    lib = _native.lib()
    hip = HipBuffers()
    ctx = C.c_void_p()
    try:
        d_r, d_c, d_v = hip.upload(r), hip.upload(c), hip.upload(v)
        links = C.c_int64()
        assert lib.rcm_ops_canonicalize_device(ops._h, n, n, r.size, d_r, d_c, d_v, 8, 1, C.byref(links)) == 0
        nn, nnz = C.c_int32(), C.c_int64()
        d_rp, d_ci, d_vals = C.c_void_p(), C.c_void_p(), C.c_void_p()
        assert lib.rcm_ops_result(ops._h, C.byref(nn), C.byref(nnz), C.byref(d_rp), C.byref(d_ci),
                                  C.byref(d_vals)) == 0
        assert lib.rcm_ctx_create(0, C.byref(ctx)) == 0
        assert lib.rcm_ctx_load_device(ctx, nn.value, nnz.value, C.cast(d_rp, C.POINTER(C.c_int32)),
                                       C.cast(d_ci, C.POINTER(C.c_int32))) == 0
        start, iters = C.c_int32(), C.c_int32()
        assert lib.rcm_ctx_find_start(ctx, C.byref(start), C.byref(iters)) == 0
        assert lib.rcm_ctx_run(ctx, start.value, P.RCM_FLAG_REVERSE | P.RCM_FLAG_STABLE) == 0
        d_perm = C.c_void_p()
        assert lib.rcm_ctx_perm_device(ctx, C.byref(d_perm)) == 0
        out_rp = hip.alloc((n + 1) * 4)
        out_ci = hip.alloc(nnz.value * 4)
        out_v = hip.alloc(nnz.value * 8)
        assert lib.rcm_ops_permute_device(ops._h, nn.value, nnz.value, d_rp, d_ci, d_vals, 8, d_perm, out_rp, out_ci,
                                          out_v) == 0
        perm = np.empty(n, np.int32)
        assert lib.rcm_ctx_read_perm(ctx, perm.ctypes.data_as(C.POINTER(C.c_int32))) == 0
        got_rp = hip.download(out_rp, n + 1, np.int32)
        got_ci = hip.download(out_ci, nnz.value, np.int32)
        got_v = hip.download(out_v, nnz.value, np.float64)
    finally:
        if ctx:
            lib.rcm_ctx_destroy(ctx)
        hip.close()
    # the oracle's pipeline
    host, hlinks = oracle_canonical(n, r, c, v, True)
    assert links.value == hlinks == 3
    rc, hstart, _, _ = O.peripheral(host.row_offsets, host.col_ids)
    assert rc == 0 and hstart == start.value
    cm = O.cm(host.row_offsets, host.col_ids, hstart)
    np.testing.assert_array_equal(perm, cm[::-1])
    erp, eci, ev = M.permute_ref(host.row_offsets, host.col_ids, host.data, perm)
    np.testing.assert_array_equal(got_rp, erp)
    np.testing.assert_array_equal(got_ci, eci)
    np.testing.assert_array_equal(got_v, ev)
    # the permuted matrix's bandwidth is the reorder's bandwidth
    assert O.bandwidth_identity(got_rp, got_ci) == O.bandwidth(host.row_offsets, host.col_ids, perm)


@pytest.mark.slow
def test_full_size_permute_and_canonicalize(ops):
    """128^3 7-point matrix (2,097,152 rows, 14,581,760 entries): canonicalise its shuffled lower triangle, reorder,
    permute; checked against the oracle's convert + prepareRCM and the permutation definition."""
    rng = np.random.default_rng(5)
    m = P.workloads.config(2, seed=7)
    n = m.rows
    rows = M.rows_of(m.row_offsets)
    lower = rows >= m.col_ids
    r = rows[lower].astype(np.int32)
    c = m.col_ids[lower]
    shuffle = rng.permutation(r.size)
    r, c = r[shuffle], c[shuffle]
    v = rng.standard_normal(r.size).astype(np.float32).astype(np.float64)
    _say("full size canonicalize", n, r.size)
    got, links = ops.canonicalize(n, n, r, c, v, True)
    host, hlinks = oracle_canonical(n, r, c, v, True)
    assert links == hlinks
    np.testing.assert_array_equal(got.row_offsets, host.row_offsets)
    np.testing.assert_array_equal(got.col_ids, host.col_ids)
    np.testing.assert_array_equal(got.data, host.data)
    _say("full size permute", n, host.nnz)
    perm, _, _ = P.rcm_gpu_batch(host, -1, P.RCM_FLAG_STABLE | P.RCM_FLAG_REVERSE)
    b = ops.permute(host, perm)
    erp, eci, ev = M.permute_ref(host.row_offsets, host.col_ids, host.data, perm)
    np.testing.assert_array_equal(b.row_offsets, erp)
    np.testing.assert_array_equal(b.col_ids, eci)
    np.testing.assert_array_equal(b.data, ev)
