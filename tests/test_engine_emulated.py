"""CPU: the engines' device code, compiled for the host and run under the wave emulator (tests/emu), against
the oracle. Lanes are threads, so every cross-lane operation and barrier is a real rendezvous: the speculative
scheme (claims racing across workers, look-back, in-order commit) runs with true concurrency. This checks the
device code's logic without a GPU; it says nothing about gfx950 code generation."""
import numpy as np
import pytest

import graphs as G
import oracle_lib as O
import parallelbatchrcm_amd as P

emu_lib = pytest.importorskip("emu_lib")


def _cases():
    rp, ci = G.random_connected(300, 500, seed=7)
    drp, dci = G.with_duplicates(rp, ci, seed=3)
    m, _ = P.workloads.permute_symmetric(P.workloads.stencil2d(24, 20), seed=11)
    return {
        "path17": G.path(17),
        "star40": G.star(40),
        "complete12": G.complete(12),
        "grid9x7": G.grid2d(9, 7),
        "random300": (rp, ci),
        "duplicates300": (drp, dci),
        "stencil24x20": (m.row_offsets, m.col_ids),
    }


CASES = _cases()


def _starts(rp, ci):
    n = rp.shape[0] - 1
    rc, s, _, _ = O.peripheral(rp, ci)
    assert rc == 0
    return sorted({0, n - 1, n // 2, s})


@pytest.mark.parametrize("engine,workers", [("block", 1), ("block", 3), ("wave", 1), ("wave", 4)])
@pytest.mark.parametrize("name", list(CASES))
def test_cm_matches_oracle(engine, workers, name):
    rp, ci = CASES[name]
    for start in _starts(rp, ci):
        perm, _, info = emu_lib.run(rp, ci, start, workers=workers, engine=engine)
        assert info[0] == 0, f"device error {info.tolist()}"
        assert info[1] == rp.shape[0] - 1
        want = O.cm(rp, ci, start)
        assert np.array_equal(perm, want), (name, start)


@pytest.mark.parametrize("engine", ["block", "wave"])
@pytest.mark.parametrize("sort_mode", [1, 2])
def test_sort_modes_agree(engine, sort_mode):
    rp, ci = CASES["random300"]
    perm, _, info = emu_lib.run(rp, ci, 5, workers=2, sort_mode=sort_mode, engine=engine)
    assert info[0] == 0
    assert np.array_equal(perm, O.cm(rp, ci, 5))


@pytest.mark.parametrize("engine", ["block", "wave"])
def test_bfs_levels(engine):
    rp, ci = CASES["stencil24x20"]
    perm, lev, info = emu_lib.run(rp, ci, 3, workers=2, bfs=True, engine=engine)
    assert info[0] == 0
    want_order, want_level = O.bfs(rp, ci, 3)
    assert np.array_equal(perm, want_order)
    assert np.array_equal(lev, want_level)


def test_disconnected_places_component_only():
    rp, ci = G.from_edges(10, [(0, 1), (1, 2), (5, 6)])
    perm, _, info = emu_lib.run(rp, ci, 0, workers=2, engine="block")
    assert info[0] == 0 and info[1] == 3
    assert perm[:3].tolist() == [0, 1, 2]


def _many_set():
    mats = [CASES[k] for k in ("path17", "star40", "complete12", "grid9x7", "random300", "duplicates300",
                               "stencil24x20")]
    mats += [G.random_connected(80 + 40 * i, 120 + 30 * i, seed=20 + i) for i in range(5)]
    mats += [G.path(1), G.path(2)]
    return mats


@pytest.mark.parametrize("workers", [1, 3, 7])
def test_many_matrices_one_launch(workers):
    mats = _many_set()
    starts = [(len(rp) - 1) // 3 for rp, _ in mats]
    perms, _, info, done = emu_lib.run_many(mats, starts, workers=workers)
    assert done == len(mats)
    for k, ((rp, ci), s) in enumerate(zip(mats, starts)):
        assert info[k, 0] == 0 and info[k, 3] == 1, (k, info[k].tolist())
        assert info[k, 1] == rp.shape[0] - 1
        assert np.array_equal(perms[k], O.cm(rp, ci, s)), k


def test_many_matrices_bfs_levels():
    mats = _many_set()
    starts = [0 for _ in mats]
    perms, levs, info, done = emu_lib.run_many(mats, starts, workers=4, bfs=True)
    assert done == len(mats)
    for k, (rp, ci) in enumerate(mats):
        order, level = O.bfs(rp, ci, 0)
        assert np.array_equal(perms[k], order), k
        assert np.array_equal(levs[k], level), k


def _dist_by_node(rp, ci, start):
    """BFS distance of every node from start, from the oracle's plain BFS (order, level per position)."""
    order, level = O.bfs(rp, ci, start)
    dist = np.full(rp.shape[0] - 1, -1, np.int64)
    dist[order] = level
    return dist


@pytest.mark.parametrize("name", ["stencil24x20", "duplicates300", "grid9x7"])
def test_cm_run_records_levels(name):
    """A CM run with a level buffer (the start search's speculative sweep) places the CM order unchanged and
    records each position's BFS distance from the start; its last entry is the start's eccentricity."""
    rp, ci = CASES[name]
    for start in _starts(rp, ci):
        perm, lev, info = emu_lib.run(rp, ci, start, workers=3, levels=True)
        assert info[0] == 0
        assert np.array_equal(perm, O.cm(rp, ci, start)), (name, start)
        dist = _dist_by_node(rp, ci, start)
        assert np.array_equal(lev, dist[perm]), (name, start)
        assert lev[-1] == dist.max()


def test_cm_run_records_levels_long_rows():
    rp, ci = LONG["twohubs2600"]()
    for start in (5, int(np.argmax(np.diff(rp)))):
        perm, lev, info = emu_lib.run(rp, ci, start, workers=2, levels=True, timeout_s=60)
        assert info[0] == 0
        assert np.array_equal(perm, O.cm(rp, ci, start))
        assert np.array_equal(lev, _dist_by_node(rp, ci, start)[perm])


def test_many_matrices_cm_levels():
    mats = _many_set() + [LONG["star3000"]()]
    starts = [0 for _ in mats]
    perms, levs, info, done = emu_lib.run_many(mats, starts, workers=4, levels=True, timeout_s=60)
    assert done == len(mats)
    for k, (rp, ci) in enumerate(mats):
        assert np.array_equal(perms[k], O.cm(rp, ci, 0)), k
        assert np.array_equal(levs[k], _dist_by_node(rp, ci, 0)[perms[k]]), k


def test_many_matrices_disconnected_member():
    rp, ci = G.from_edges(10, [(0, 1), (1, 2), (5, 6)])
    mats = [G.path(9), (rp, ci), G.star(12)]
    perms, _, info, done = emu_lib.run_many(mats, [0, 0, 0], workers=2)
    assert done == 3
    assert info[1, 0] == 0 and info[1, 1] == 3
    assert np.array_equal(perms[0], O.cm(*mats[0], 0)) and np.array_equal(perms[2], O.cm(*mats[2], 0))


LONG = {
    "star3000": lambda: G.star(3000),
    "hub5000": lambda: G.hub_graph(5000, 9000, seed=1),
    "twohubs2600": lambda: G.hub_graph(2600, 3000, seed=2, hubs=2),
    "hubdups": lambda: G.with_duplicates(*G.hub_graph(2500, 2000, seed=3), seed=4),
}


@pytest.mark.parametrize("name", list(LONG))
@pytest.mark.parametrize("workers", [1, 3])
def test_long_rows(name, workers):
    rp, ci = LONG[name]()
    n = rp.shape[0] - 1
    deg = np.diff(rp)
    hub = int(np.argmax(deg))
    assert deg[hub] > 2048
    for start in sorted({hub, int(np.argmin(deg)), n // 2}):
        perm, _, info = emu_lib.run(rp, ci, start, workers=workers, engine="block", timeout_s=60)
        assert info[0] == 0, f"device error {info.tolist()}"
        assert np.array_equal(perm, O.cm(rp, ci, start)), (name, start)


def test_long_rows_bfs_levels():
    rp, ci = LONG["twohubs2600"]()
    perm, lev, info = emu_lib.run(rp, ci, 5, workers=2, bfs=True, engine="block", timeout_s=60)
    assert info[0] == 0
    order, level = O.bfs(rp, ci, 5)
    assert np.array_equal(perm, order) and np.array_equal(lev, level)


def test_long_rows_many():
    mats = [LONG["hub5000"](), G.path(30), LONG["star3000"]()]
    starts = [0, 0, 0]
    perms, _, info, done = emu_lib.run_many(mats, starts, workers=3, timeout_s=60)
    assert done == 3
    for k, (rp, ci) in enumerate(mats):
        assert info[k, 0] == 0 and np.array_equal(perms[k], O.cm(rp, ci, 0)), k


@pytest.mark.parametrize("flt", [1, 2])
@pytest.mark.parametrize("name", ["random300", "duplicates300", "stencil24x20"])
def test_claim_filter(flt, name):
    rp, ci = CASES[name]
    for start in _starts(rp, ci):
        perm, _, info = emu_lib.run(rp, ci, start, workers=3, engine="block", claim_filter=flt)
        assert info[0] == 0 and np.array_equal(perm, O.cm(rp, ci, start)), (name, start)


def test_claim_filter_long_rows():
    rp, ci = LONG["twohubs2600"]()
    perm, _, info = emu_lib.run(rp, ci, 0, workers=3, engine="block", claim_filter=1, timeout_s=60)
    assert info[0] == 0 and np.array_equal(perm, O.cm(rp, ci, 0))


@pytest.mark.parametrize("slots", [4096, 1024, 2049, 4097], ids=["v1_256x4096", "v2_128x1024", "v3_digit6",
                                                                  "v4_512x4096"])
@pytest.mark.parametrize("name", ["random300", "duplicates300", "stencil24x20"])
def test_engine_variants(slots, name):
    rp, ci = CASES[name]
    for start in _starts(rp, ci)[:2]:
        for sort_mode in (0, 1):
            perm, _, info = emu_lib.run(rp, ci, start, workers=2, slots=slots, sort_mode=sort_mode, engine="block")
            assert info[0] == 0 and np.array_equal(perm, O.cm(rp, ci, start)), (name, start, sort_mode)


@pytest.mark.parametrize("slots", [4096, 1024, 2049, 4097])
def test_engine_variants_long_rows(slots):
    rp, ci = LONG["hub5000"]()
    perm, _, info = emu_lib.run(rp, ci, 3, workers=2, slots=slots, engine="block", timeout_s=60)
    assert info[0] == 0 and np.array_equal(perm, O.cm(rp, ci, 3))


def test_one_workgroup_engine_matches_oracle():
    """engine_wg.hpp (one workgroup per matrix, level by level) in the emulator: CM and plain-BFS orders of a
    mixed set from several starts, including duplicates, unsorted rows, a 1-node and a 2-node graph."""
    mats = [G.path(17), G.star(40), G.complete(9), G.grid2d(7, 5), G.random_connected(500, 1500, seed=4),
            G.with_duplicates(*G.random_connected(300, 600, seed=5), seed=6),
            G.shuffle_rows(*G.random_connected(300, 900, seed=7), seed=8), G.path(1), G.path(2)]
    m = P.workloads.config(3, scale=3000, seed=5)
    mats.append((m.row_offsets, m.col_ids))
    for pick in (lambda n: n // 2, lambda n: 0, lambda n: n - 1):
        starts = [pick(rp.shape[0] - 1) for rp, _ in mats]
        for bfs in (False, True):
            rc, perms, levs, info = emu_lib.run_wg(mats, starts, workers=3, bfs=bfs)
            assert rc == 0
            for k, (rp, ci) in enumerate(mats):
                if bfs:
                    order, lev = O.bfs(rp, ci, starts[k])
                    np.testing.assert_array_equal(perms[k], order)
                    np.testing.assert_array_equal(levs[k], lev)
                else:
                    np.testing.assert_array_equal(perms[k], O.cm(rp, ci, starts[k]))
                assert info[2 * k] == rp.shape[0] - 1


DIRECTED = {"directed3000": (3000, 3000, 40, 1), "directed2000": (2000, 1500, 24, 2)}


@pytest.mark.parametrize("name", list(DIRECTED))
def test_directed_graph_with_empty_rows(name):
    """Rows are out-neighbour lists (not symmetric) and half the nodes have empty rows, so batches hold parents
    with no adjacency entries among parents with rows of up to 40 (both sort paths): the parent marks, the
    chunked compaction and the cut against the serial algorithm, CM and plain BFS, one and many matrices."""
    rp, ci = G.directed_with_empty_rows(*DIRECTED[name])
    assert (np.diff(rp) == 0).sum() > rp.shape[0] // 3
    ref = O.cm(rp, ci, 0)
    for sort_mode in (0, 1, 2):
        perm, _, info = emu_lib.run(rp, ci, 0, workers=3, sort_mode=sort_mode, engine="block")
        assert info[0] == 0 and np.array_equal(perm, ref), (name, sort_mode)
    want_order, want_level = O.bfs(rp, ci, 0)
    perm, lev, info = emu_lib.run(rp, ci, 0, workers=3, bfs=True, engine="block")
    assert info[0] == 0 and np.array_equal(perm, want_order) and np.array_equal(lev, want_level)
    perms, _, info, done = emu_lib.run_many([(rp, ci), CASES["grid9x7"]], [0, 0], workers=3)
    assert done == 2 and np.array_equal(perms[0], ref)
    assert np.array_equal(perms[1], O.cm(*CASES["grid9x7"], 0))
