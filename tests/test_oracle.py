"""CPU tests that pin the oracle (no GPU).

The oracle restates src/rcm_cpu.cpp:57-104, src/simple_peripheral.cpp:44-141, src/main.cpp:114-226 and
include/Graph.h:74-264. It is pinned by
  * the reference itself where it compiles here (oracle/_ref: RealCPU::rcm serial and CPU_BATCH, prepareRCM),
  * closed forms (path, star, complete graph, grid from a corner),
  * an independent brute-force CM written below in pure Python (tiny graphs, every start node),
  * the order characterisation (orc_verify_order), itself checked to reject planted mistakes.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

import graphs as G
import oracle_lib as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")


# ---- an independent CM, written from the definition (level by level, explicit per-parent sort) ----------
def brute_cm(rp, ci, start):
    n = len(rp) - 1
    placed = {start}
    order = [start]
    head = 0
    while head < len(order):
        v = order[head]
        head += 1
        kids = []
        seen_here = set()
        for k in range(rp[v], rp[v + 1]):
            c = int(ci[k])
            if c in placed or c in seen_here:
                continue
            seen_here.add(c)
            kids.append((int(rp[c + 1] - rp[c]), k, c))  # (degree, first row index, node)
        kids.sort()  # degree, then row position: the stable order
        for _, _, c in kids:
            placed.add(c)
            order.append(c)
    return np.array(order, np.int32)


TINY = {
    "path9": G.path(9),
    "star7": G.star(7),
    "complete6": G.complete(6),
    "grid4x3": G.grid2d(4, 3),
    "rand12a": G.random_connected(12, 10, seed=1),
    "rand12b": G.random_connected(12, 25, seed=2),
    "dups": G.with_duplicates(*G.random_connected(11, 12, seed=3), seed=4),
    "unsorted": G.shuffle_rows(*G.random_connected(12, 18, seed=5), seed=6),
    "nodiag": G.random_connected(10, 8, seed=7, diag=False),
}


@pytest.mark.parametrize("name", list(TINY))
def test_cm_matches_brute_force_every_start(name):
    rp, ci = TINY[name]
    n = len(rp) - 1
    for s in range(n):
        got = O.cm(rp, ci, s)
        np.testing.assert_array_equal(got, brute_cm(rp, ci, s), err_msg=f"{name} start {s}")
        assert O.verify_order(rp, ci, s, got) == 0


def test_cm_closed_forms():
    rp, ci = G.path(20)
    np.testing.assert_array_equal(O.cm(rp, ci, 0), np.arange(20))
    np.testing.assert_array_equal(O.cm(rp, ci, 19), np.arange(19, -1, -1))
    rp, ci = G.star(10)
    np.testing.assert_array_equal(O.cm(rp, ci, 0), np.arange(10))  # centre, then leaves in row order
    # from a leaf: leaf, centre, then the other leaves in row order
    np.testing.assert_array_equal(O.cm(rp, ci, 4), [4, 0, 1, 2, 3, 5, 6, 7, 8, 9])
    rp, ci = G.complete(8)
    np.testing.assert_array_equal(O.cm(rp, ci, 3), [3, 0, 1, 2, 4, 5, 6, 7])
    # 2D grid from a corner under natural numbering: anti-diagonals; boundary nodes (degree 4 incl. diagonal)
    # sort before interior nodes (degree 5) within each parent's children
    nx = ny = 6
    rp, ci = G.grid2d(nx, ny)
    p = O.cm(rp, ci, 0)
    diag_of = [(v % nx) + (v // nx) for v in p]
    assert diag_of == sorted(diag_of)


def test_verifier_rejects_planted_mistakes():
    rp, ci = G.random_connected(200, 500, seed=11)
    s = 17
    p = O.cm(rp, ci, s)
    assert O.verify_order(rp, ci, s, p) == 0
    # swap two adjacent entries with different keys
    for i in range(1, len(p) - 1):
        q = p.copy()
        q[i], q[i + 1] = q[i + 1], q[i]
        assert O.verify_order(rp, ci, s, q) != 0
    # degree tie broken the wrong way / BFS order instead of CM
    b, _ = O.bfs(rp, ci, s)
    if not np.array_equal(b, p):
        assert O.verify_order(rp, ci, s, b) != 0
        assert O.verify_order(rp, ci, s, b, sorted_=False) == 0
    assert O.verify_order(rp, ci, s + 1, p) != 0  # wrong start
    assert O.verify_order(rp, ci, s, p[:-1]) != 0  # not the whole component


def test_bfs_levels_and_order():
    rp, ci = G.grid2d(5, 4)
    order, lev = O.bfs(rp, ci, 0)
    assert lev[0] == 0 and np.all(np.diff(lev) >= 0)
    for v, l in zip(order, lev):
        assert l == (v % 5) + (v // 5)


def test_peripheral_hand_cases():
    rp, ci = G.path(17)  # from rows/2 = 8 to an end, then to the other end, which is returned
    rc, s, it, ecc = O.peripheral(rp, ci)
    assert (rc, s, it, ecc) == (0, 0, 3, 16)
    rp, ci = G.path(1)
    assert O.peripheral(rp, ci)[:2] == (0, 0)
    rp, ci = G.from_edges(6, [(0, 1), (3, 4)])
    assert O.peripheral(rp, ci)[0] == -2  # not connected
    rp, ci = G.star(9)  # rows/2 = 4 is a leaf: BFS reaches the centre then the other leaves at distance 2
    rc, s, it, _ = O.peripheral(rp, ci)
    assert rc == 0 and s != 0


def brute_peripheral(rp, ci):
    """SimplePeripheral restated independently: full BFS distances, last-32 window of discovery order."""
    n = len(rp) - 1
    if n == 1:
        return 0
    cur, lastdist = n // 2, 0
    while True:
        order, lev = O.bfs(rp, ci, cur)
        if len(order) != n:
            return None
        win = list(zip(order[1:], lev[1:]))[-32:]
        if win[-1][1] == lastdist:
            return cur
        mx = max(d for _, d in win)
        first = next(i for i, (_, d) in enumerate(win) if d == mx)
        best = None
        for v, _ in win[first:]:
            dv = int(rp[v + 1] - rp[v])
            if best is None or dv <= best[0]:
                best = (dv, v)
        cur = best[1]
        lastdist = win[-1][1]


@pytest.mark.parametrize("seed", range(12))
def test_peripheral_matches_restatement(seed):
    rp, ci = G.random_connected(40 + 7 * seed, 30 + 5 * seed, seed=seed)
    assert O.peripheral(rp, ci)[1] == brute_peripheral(rp, ci)


def test_bandwidth_closed_forms_and_reversal():
    rp, ci = G.path(30)
    assert O.bandwidth(rp, ci, np.arange(30)) == 1
    rp, ci = G.star(10)
    assert O.bandwidth(rp, ci, np.arange(10)) == 9
    rp, ci = G.grid2d(7, 9)
    assert O.bandwidth(rp, ci, np.arange(63)) == 7 == O.bandwidth_identity(rp, ci)
    rng = np.random.default_rng(5)
    for _ in range(5):
        p = rng.permutation(63).astype(np.int32)
        inv = np.empty(63, np.int64)
        inv[p] = np.arange(63)
        brute = max(abs(int(inv[i]) - int(inv[int(c)])) for i in range(63) for c in ci[rp[i]:rp[i + 1]])
        assert O.bandwidth(rp, ci, p) == brute == O.bandwidth(rp, ci, p[::-1])
    assert O.bandwidth(rp, ci, np.zeros(0, np.int32)) == -1


def test_check_perm():
    assert O.check_perm(np.array([2, 0, 1])) == 0
    assert O.check_perm(np.array([2, 0, 2])) == -1
    assert O.check_perm(np.array([3, 0, 1])) == -1


# ---- the reference's own code (oracle/_ref) ------------------------------------------------------------
@needs_ref
@pytest.mark.parametrize("name", ["rand12b", "dups", "unsorted", "grid4x3"])
def test_oracle_cm_equals_reference_serial_tiny(name):
    rp, ci = TINY[name]
    for s in range(len(rp) - 1):
        ref, _ = O.ref_rcm(rp, ci, s)
        np.testing.assert_array_equal(O.cm(rp, ci, s), ref)


@needs_ref
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_oracle_cm_equals_reference_serial_random(seed):
    rp, ci = G.random_connected(3000, 9000, seed=seed)
    s = int(np.random.default_rng(seed).integers(0, 3000))
    ref, _ = O.ref_rcm(rp, ci, s)
    np.testing.assert_array_equal(O.cm(rp, ci, s), ref)


@needs_ref
def test_reference_cpu_batch_equals_serial():
    """The reference's threaded CPU_BATCH (stable) equals its serial CM (src/main.cpp:465-477)."""
    rp, ci = G.random_connected(20000, 60000, seed=9)
    ref, _ = O.ref_rcm(rp, ci, 5)
    for t in (1, 2, 4):
        got, _ = O.ref_rcm(rp, ci, 5, threads=t, batchsize=512)
        np.testing.assert_array_equal(got, ref)


def _random_csr(n, m, seed, sym):
    rng = np.random.default_rng(seed)
    rows = [[] for _ in range(n)]
    for _ in range(m):
        a, b = rng.integers(0, n, 2)
        rows[a].append(int(b))
        if sym:
            rows[b].append(int(a))
    for i in range(n):
        rows[i].append(i)  # full diagonal (the repair loop needs it)
        rows[i].sort()
    rp = np.zeros(n + 1, np.int32)
    for i in range(n):
        rp[i + 1] = rp[i] + len(rows[i])
    return rp, np.array([c for r in rows for c in r], np.int32)


@needs_ref
@pytest.mark.parametrize("seed,sym", [(1, True), (2, True), (3, False), (4, False), (5, True)])
def test_oracle_prepare_equals_reference(seed, sym):
    rp, ci = _random_csr(300, 120, seed, sym)
    for check_sym in (False, True):
        rc, orp, oci = O.prepare(rp, ci, check_sym)
        assert rc >= 0
        rrp, rci = O.ref_prepare(rp, ci, check_sym)
        np.testing.assert_array_equal(orp, rrp)
        np.testing.assert_array_equal(oci, rci)


def test_prepare_hazard_detected():
    # node 3 has no entries at all: the reference's repair loop would re-add the same edge forever
    rp = np.array([0, 1, 2, 3, 3], np.int32)
    ci = np.array([0, 1, 2], np.int32)
    assert O.prepare(rp, ci, False)[0] == -3


def test_prepare_connects_components_in_order():
    rp, ci = G.from_edges(8, [(0, 1), (2, 3), (5, 6), (6, 7)])
    rc, r2, c2 = O.prepare(rp, ci, False)
    assert rc == 3  # components {2,3}, {4}, {5,6,7} each joined to the node below their lowest node
    rows = {i: set(c2[r2[i]:r2[i + 1]].tolist()) for i in range(8)}
    assert 1 in rows[2] and 3 in rows[4] and 4 in rows[5]
    assert O.cm(r2, c2, 0).shape[0] == 8


def _coo_without_duplicates(rng, n, k, sym):
    r = rng.integers(0, n, k)
    c = rng.integers(0, n, k)
    if sym:
        r, c = np.concatenate([r, c]), np.concatenate([c, r])
    r = np.concatenate([r, np.arange(n)])  # full diagonal (the repair loop needs it)
    c = np.concatenate([c, np.arange(n)])
    _, first = np.unique(r.astype(np.int64) * n + c, return_index=True)
    keep = np.zeros(r.size, bool)
    keep[first] = True  # no duplicates: the reference's unstable sort then has a unique result
    order = rng.permutation(int(keep.sum()))
    return r[keep][order].astype(np.int32), c[keep][order].astype(np.int32)


@needs_ref
@pytest.mark.parametrize("seed,sym", [(11, True), (12, False), (13, False), (14, True)])
def test_oracle_prepare_values_equals_reference(seed, sym):
    """orc_prepare_values against the reference's own Graph<double>::prepareRCM with values."""
    rng = np.random.default_rng(seed)
    n = int(rng.integers(50, 400))
    r, c = _coo_without_duplicates(rng, n, n // 3, sym)
    v = rng.standard_normal(r.size)
    rp, ci, vals = O.convert(n, r, c, v)
    for check_sym in (False, True):
        if not check_sym and not sym:
            continue  # the reference's repair on a directed pattern is covered by test_oracle_prepare_equals_reference
        rc, orp, oci, ov = O.prepare_values(rp, ci, vals, check_sym)
        assert rc >= 0
        _, _, rrp, rci, rv = O.ref_prepare_values(n, n, rp, ci, vals, check_sym, True)
        np.testing.assert_array_equal(orp, rrp)
        np.testing.assert_array_equal(oci, rci)
        np.testing.assert_array_equal(ov, rv)
        rc2, prp, pci = O.prepare(rp, ci, check_sym)  # the pattern-only form agrees
        assert rc2 == rc
        np.testing.assert_array_equal(prp, orp)
        np.testing.assert_array_equal(pci, oci)


def test_oracle_prepare_values_closed_form():
    """Two one-sided entries get their transposes with the same value; three components get two links of value 1."""
    # (0,1) stored once, (3,2) stored once; diagonal everywhere; components {0,1}, {2,3}, {4}
    r = np.array([0, 0, 1, 2, 3, 3, 4], np.int32)
    c = np.array([0, 1, 1, 2, 2, 3, 4], np.int32)
    v = np.array([5.0, -7.0, 6.0, 8.0, -9.0, 4.0, 3.0])
    rp, ci, vals = O.convert(5, r, c, v)
    rc, orp, oci, ov = O.prepare_values(rp, ci, vals, True)
    assert rc == 2
    want = {(0, 0): 5.0, (0, 1): -7.0, (1, 0): -7.0, (1, 1): 6.0, (1, 2): 1.0, (2, 1): 1.0, (2, 2): 8.0,
            (2, 3): -9.0, (3, 2): -9.0, (3, 3): 4.0, (3, 4): 1.0, (4, 3): 1.0, (4, 4): 3.0}
    got = {(i, int(oci[k])): float(ov[k]) for i in range(5) for k in range(orp[i], orp[i + 1])}
    assert got == want


def test_oracle_convert_is_stable_lexicographic():
    """orc_convert equals a stable numpy lexsort on (row, col), duplicates in input order with their values."""
    rng = np.random.default_rng(15)
    n, k = 300, 6000
    r = rng.integers(0, n, k).astype(np.int32)
    c = rng.integers(0, 40, k).astype(np.int32)
    v = np.arange(k, dtype=np.float64)  # distinct values expose any reordering of duplicates
    rp, ci, vals = O.convert(n, r, c, v)
    order = np.lexsort((np.arange(k), c, r))
    np.testing.assert_array_equal(ci, c[order])
    np.testing.assert_array_equal(vals, v[order])
    np.testing.assert_array_equal(rp, np.concatenate([[0], np.cumsum(np.bincount(r, minlength=n))]))
    rp2, ci2, none = O.convert(n, r, c)
    assert none is None
    np.testing.assert_array_equal(rp2, rp)
    np.testing.assert_array_equal(ci2, ci)
