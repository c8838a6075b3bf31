"""CPU checks of the canonicalisation and permutation kernels (parallelbatchrcm_amd/csrc/matrix_ops.hpp) run in
the wave emulator through the library's own orchestration: every lane is a host thread, ballots and barriers are
collectives. Checked against numpy definitions and the oracle's convert / value-carrying prepareRCM
(oracle/rcm_oracle.c, pinned to the reference's own Graph<double>::prepareRCM in tests/test_oracle.py)."""
from __future__ import annotations

import numpy as np
import pytest

import emu_lib as E
import graphs as G
import matrix_ref as M
import oracle_lib as O
import parallelbatchrcm_amd as P


def test_scan_closed_form():
    for m in (1, 255, 4096, 4097, 9000):
        a = np.random.default_rng(m).integers(0, 1000, m).astype(np.uint32)
        np.testing.assert_array_equal(E.scan(a), np.concatenate([[0], np.cumsum(a)]).astype(np.uint32))


@pytest.mark.parametrize("lo_bits,hi_bits", [(8, 0), (12, 6), (3, 17), (24, 0)])
def test_radix_sort_is_stable(lo_bits, hi_bits):
    rng = np.random.default_rng(lo_bits * 31 + hi_bits)
    n = 7000
    lo = rng.integers(0, 1 << lo_bits, n).astype(np.uint64) if lo_bits else np.zeros(n, np.uint64)
    hi = rng.integers(0, 1 << hi_bits, n).astype(np.uint64) if hi_bits else np.zeros(n, np.uint64)
    lo[::7] = lo[0]  # many ties
    keys = (hi << np.uint64(32)) | lo
    vals = np.arange(n, dtype=np.uint32)
    k, v = E.radix_sort(keys, vals, lo_bits, hi_bits)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])  # equal keys keep their input order


def matrix_with_long_rows(rng, n, k):
    r = rng.integers(0, n, k)
    c = rng.integers(0, n, k)
    # two long rows (> 64 entries, one with repeated columns) and a long column
    r = np.concatenate([r, np.full(300, 5), np.full(90, 11), rng.integers(0, n, 150)])
    c = np.concatenate([c, rng.integers(0, n, 300), rng.integers(0, 7, 90), np.full(150, 9)])
    return P.convert_coo(n, n, r, c, rng.integers(-99, 99, r.size).astype(np.float64))


@pytest.mark.parametrize("seed", range(3))
def test_permute_matches_definition(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(300, 900))
    m = matrix_with_long_rows(rng, n, 5 * n)
    perm = rng.permutation(n).astype(np.int32)
    rc, rp2, ci2, v2 = E.permute(m.row_offsets, m.col_ids, m.data, perm)
    assert rc == 0
    erp, eci, ev = M.permute_ref(m.row_offsets, m.col_ids, m.data, perm)
    np.testing.assert_array_equal(rp2, erp)
    np.testing.assert_array_equal(ci2, eci)
    np.testing.assert_array_equal(v2, ev)
    # dense form: B = A[perm][:, perm]
    np.testing.assert_array_equal(M.dense(rp2, ci2, v2, n), M.dense(m.row_offsets, m.col_ids, m.data, n)[perm][:, perm])


def test_permute_rejects_bad_input():
    rng = np.random.default_rng(7)
    m = matrix_with_long_rows(rng, 200, 600)
    perm = rng.permutation(200).astype(np.int32)
    bad = perm.copy()
    bad[3] = bad[4]
    assert E.permute(m.row_offsets, m.col_ids, m.data, bad)[0] == -1
    bad = perm.copy()
    bad[0] = 200
    assert E.permute(m.row_offsets, m.col_ids, m.data, bad)[0] == -1
    rp = m.row_offsets.copy()
    rp[-1] += 1  # row offsets do not span nnz
    assert E.permute(rp, m.col_ids, m.data, perm)[0] == -1


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("check_sym", [True, False])
def test_canonicalize_matches_host(seed, check_sym):
    rng = np.random.default_rng(50 + seed)
    n = int(rng.integers(60, 500))
    r, c = M.coo_with_components(rng, n, 4 * n, comps=int(rng.integers(1, 6)))
    v = rng.integers(-99, 99, r.size).astype(np.float64)
    if not check_sym:  # a symmetric input, its components still apart
        sym = P.convert_coo(n, n, np.concatenate([r, c]), np.concatenate([c, r]), np.concatenate([v, v]))
        r, c, v = M.rows_of(sym.row_offsets).astype(np.int32), sym.col_ids, sym.data
    rc, res = E.canonicalize(n, n, r, c, v, check_sym)
    assert rc == 0
    nn, rp, ci, vv, links = res
    _expect_oracle(n, r, c, v, check_sym, rp, ci, vv, links)


def _expect_oracle(n, r, c, v, check_sym, rp, ci, vv, links):
    orp, oci, ov = O.convert(n, r, c, v)
    added, erp, eci, ev = O.prepare_values(orp, oci, ov, check_sym)
    assert links == added
    np.testing.assert_array_equal(rp, erp)
    np.testing.assert_array_equal(ci, eci)
    np.testing.assert_array_equal(vv, ev)


def directed_coo(rng, n, k, diag=True, cycles=0):
    """Random directed entries (no mirror), optionally a full diagonal, and `cycles` directed cycles."""
    r = list(rng.integers(0, n, k))
    c = list(rng.integers(0, n, k))
    if diag:
        r += list(range(n))
        c += list(range(n))
    for _ in range(cycles):
        nodes = rng.choice(n, int(rng.integers(2, 6)), replace=False)
        for a, b in zip(nodes, np.roll(nodes, -1)):
            r.append(a)
            c.append(b)
    return np.array(r, np.int32), np.array(c, np.int32)


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("check_sym", [True, False])
def test_canonicalize_directed_matches_oracle(seed, check_sym):
    """Matrices that stay structurally non-symmetric (one-sided entries without the symmetrisation, or pairs
    stored an even number of times in one direction with it): the repair follows the stored rows as the
    reference's walk does (include/Graph.h:160-263)."""
    rng = np.random.default_rng(70 + seed)
    n = int(rng.integers(40, 400))
    r, c = directed_coo(rng, n, n // 2 + int(rng.integers(0, n)), diag=True, cycles=seed)
    if check_sym:  # doubled one-sided entries keep the result non-symmetric after the parity pass
        r = np.concatenate([r, r[:7]])
        c = np.concatenate([c, c[:7]])
    v = rng.integers(-99, 99, r.size).astype(np.float64)
    rc, res = E.canonicalize(n, n, r, c, v, check_sym)
    assert rc == 0, rc
    nn, rp, ci, vv, links = res
    _expect_oracle(n, r, c, v, check_sym, rp, ci, vv, links)


def test_canonicalize_closed_forms():
    # a pair stored twice in one direction is "symmetric" to the reference's parity check but not structurally;
    # 0 -> 1 -> 2 reaches everything: nothing to link
    r = np.array([0, 0, 0, 1, 2, 1], np.int32)
    c = np.array([1, 1, 0, 1, 2, 2], np.int32)
    rc, res = E.canonicalize(3, 3, r, c, None, True)
    assert rc == 0 and res[4] == 0
    # without the symmetrisation: 0 -> 1, 2 alone with its diagonal: one link (1, 2)
    rc, res = E.canonicalize(3, 3, np.array([0, 0, 1, 2], np.int32), np.array([0, 1, 1, 2], np.int32), None, False)
    assert rc == 0 and res[4] == 1
    assert res[2].tolist() == [0, 1, 1, 2, 1, 2] and res[1].tolist() == [0, 2, 4, 6]
    # a directed chain 2 -> 1 -> 0 with diagonals: 0 reaches nothing, so 1 and 2 are each linked
    rc, res = E.canonicalize(3, 3, np.array([0, 1, 1, 2, 2], np.int32), np.array([0, 0, 1, 1, 2], np.int32), None,
                             False)
    assert rc == 0 and res[4] == 2
    # a node no smaller node reaches and on no cycle (no diagonal): the reference's walk never ends
    rc, _ = E.canonicalize(3, 3, np.array([0, 2], np.int32), np.array([0, 0], np.int32), None, False)
    assert rc == -13
    assert O.prepare(np.array([0, 1, 1, 2], np.int32), np.array([0, 0], np.int32), False)[0] == -3
    # an empty row: the reference's repair loop does not terminate
    rc, _ = E.canonicalize(3, 3, np.array([0, 0, 1], np.int32), np.array([0, 1, 0], np.int32), None, True)
    assert rc == -13
    # an index out of range
    rc, _ = E.canonicalize(3, 3, np.array([0, 3], np.int32), np.array([0, 0], np.int32), None, True)
    assert rc == -1
    # non-square with nothing to mirror
    rc, _ = E.canonicalize(2, 3, np.array([0, 1], np.int32), np.array([0, 1], np.int32), None, False)
    assert rc == -1


@pytest.mark.parametrize("seed", range(3))
def test_canonicalize_nonsquare_symmetrised(seed):
    """Fewer rows than columns: the symmetrisation places the transposes of entries in the extra columns into
    rows beyond the input's, and the result is max(rows, cols) square (include/Graph.h:77-157). Checked against the
    host convert + prepareRCM (themselves equal to the reference's code in tests/test_formats.py)."""
    rng = np.random.default_rng(90 + seed)
    cols = int(rng.integers(80, 300))
    rows = cols - int(rng.integers(1, 20))
    k = 4 * cols
    r = rng.integers(0, rows, k).astype(np.int32)
    c = rng.integers(0, cols, k).astype(np.int32)
    r = np.concatenate([r, np.arange(rows, dtype=np.int32)])  # a diagonal, and every extra column used
    c = np.concatenate([c, np.arange(rows, dtype=np.int32)])
    r = np.concatenate([r, rng.integers(0, rows, cols - rows).astype(np.int32)])
    c = np.concatenate([c, np.arange(rows, cols, dtype=np.int32)])
    v = rng.integers(-99, 99, r.size).astype(np.float64)
    rc, res = E.canonicalize(rows, cols, r, c, v, True)
    host, hlinks = P.prepare_rcm(P.convert_coo(rows, cols, r, c, v), True)
    assert rc == 0, rc
    nn, rp, ci, vv, links = res
    assert nn == cols and links == hlinks
    np.testing.assert_array_equal(rp, host.row_offsets)
    np.testing.assert_array_equal(ci, host.col_ids)
    np.testing.assert_array_equal(vv, host.data)


LEVEL_CASES = {
    "path17": G.path(17),
    "star40": G.star(40),
    "grid7x5": G.grid2d(7, 5),
    "rand500": G.random_connected(500, 1500, seed=4),
    "dups300": G.with_duplicates(*G.random_connected(300, 600, seed=5), seed=6),
    "unsorted300": G.shuffle_rows(*G.random_connected(300, 900, seed=7), seed=8),
    "hub900": G.hub_graph(900, 1500, seed=2),  # a level of 900 nodes: the wide path
    "star9000": G.star(9000),  # one parent with 9000 entries: beyond a narrow level's work
}


@pytest.mark.parametrize("name", list(LEVEL_CASES))
def test_level_synchronous_engine_matches_oracle(name):
    """level_cm.hpp: the CM order and the plain BFS order with levels, from several starts."""
    rp, ci = LEVEL_CASES[name]
    n = rp.shape[0] - 1
    for s in sorted({0, n // 2, n - 1}):
        rc, perm, _, _ = E.level_order(rp, ci, s)
        assert rc == 0
        np.testing.assert_array_equal(perm, O.cm(rp, ci, s))
        rc, order, lev, depth = E.level_order(rp, ci, s, bfs=True)
        want, wlev = O.bfs(rp, ci, s)
        np.testing.assert_array_equal(order, want)
        np.testing.assert_array_equal(lev, wlev)
        assert depth == wlev.max() + 1
