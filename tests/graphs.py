"""Small hand-built and seeded graphs shared by the CPU and GPU tests (numpy CSR, int32)."""
from __future__ import annotations

import numpy as np


def from_edges(n, edges, diag=True, sym=True):
    rows = [[] for _ in range(n)]
    for a, b in edges:
        rows[a].append(b)
        if sym and a != b:
            rows[b].append(a)
    if diag:
        for i in range(n):
            rows[i].append(i)
    rp = np.zeros(n + 1, np.int32)
    cols = []
    for i in range(n):
        r = sorted(set(rows[i])) if sym else sorted(rows[i])
        cols.extend(r)
        rp[i + 1] = rp[i] + len(r)
    return rp, np.array(cols, np.int32)


def path(n, diag=True):
    return from_edges(n, [(i, i + 1) for i in range(n - 1)], diag)


def star(n, diag=True):
    return from_edges(n, [(0, i) for i in range(1, n)], diag)


def complete(n, diag=True):
    return from_edges(n, [(i, j) for i in range(n) for j in range(i + 1, n)], diag)


def grid2d(nx, ny, diag=True):
    e = []
    for y in range(ny):
        for x in range(nx):
            v = y * nx + x
            if x + 1 < nx:
                e.append((v, v + 1))
            if y + 1 < ny:
                e.append((v, v + nx))
    return from_edges(nx * ny, e, diag)


def random_graph(n, p, seed, diag=True):
    rng = np.random.default_rng(seed)
    e = [(i, j) for i in range(n) for j in range(i + 1, n) if rng.random() < p]
    return from_edges(n, e, diag)


def random_connected(n, extra, seed, diag=True):
    """Random spanning tree plus `extra` random edges: connected, irregular degrees."""
    rng = np.random.default_rng(seed)
    order = rng.permutation(n)
    e = [(int(order[i]), int(order[rng.integers(0, i)])) for i in range(1, n)]
    for _ in range(extra):
        a, b = rng.integers(0, n, 2)
        if a != b:
            e.append((int(a), int(b)))
    return from_edges(n, e, diag)


def with_duplicates(rp, ci, seed):
    """Insert repeated column entries (the reference keeps duplicates; degree counts them)."""
    rng = np.random.default_rng(seed)
    n = rp.shape[0] - 1
    rows = []
    for i in range(n):
        r = list(ci[rp[i]:rp[i + 1]])
        if r and rng.random() < 0.3:
            r.append(r[rng.integers(0, len(r))])
            r.sort()
        rows.append(r)
    nrp = np.zeros(n + 1, np.int32)
    for i in range(n):
        nrp[i + 1] = nrp[i] + len(rows[i])
    return nrp, np.array([c for r in rows for c in r], np.int32)


def shuffle_rows(rp, ci, seed):
    """Same graph, columns of every row in a random (unsorted) order."""
    rng = np.random.default_rng(seed)
    ci = ci.copy()
    for i in range(rp.shape[0] - 1):
        seg = ci[rp[i]:rp[i + 1]]
        rng.shuffle(seg)
        ci[rp[i]:rp[i + 1]] = seg
    return rp, ci


def hub_graph(leaves, extra, seed, hubs=1, diag=True):
    """`hubs` nodes adjacent to every one of `leaves` other nodes (rows longer than a batch's slots), plus
    `extra` random edges among the leaves so that the leaves' degrees vary; node ids shuffled."""
    rng = np.random.default_rng(seed)
    n = hubs + leaves
    e = [(h, hubs + i) for h in range(hubs) for i in range(leaves)]
    e += [(h, h + 1) for h in range(hubs - 1)]
    for _ in range(extra):
        a, b = rng.integers(hubs, n, 2)
        if a != b:
            e.append((int(a), int(b)))
    perm = rng.permutation(n)
    e = [(int(perm[a]), int(perm[b])) for a, b in e]
    return from_edges(n, e, diag)


def directed_with_empty_rows(n, extra, max_row, seed):
    """A directed graph (rows are out-neighbour lists, not symmetric) in which every node is reachable from node 0
    through a random tree, and the tree's leaves keep empty rows unless an extra edge lands on them: batches then
    hold parents with no adjacency entries between parents with long rows (up to max_row entries)."""
    rng = np.random.default_rng(seed)
    rows = [[] for _ in range(n)]
    order = rng.permutation(np.arange(1, n))
    placed = [0]
    for v in order:  # attach each node below a random placed node that still has room
        while True:
            p = placed[int(rng.integers(len(placed)))]
            if len(rows[p]) < max_row:
                break
        rows[p].append(int(v))
        placed.append(int(v))
    inner = [i for i in range(n) if rows[i]]
    heavy = inner[: max(1, len(inner) // 20)]  # extra edges go to a few nodes, whose rows grow to max_row
    for _ in range(extra):
        a = heavy[int(rng.integers(len(heavy)))]
        if len(rows[a]) < max_row:
            rows[a].append(int(rng.integers(n)))
    rp = np.zeros(n + 1, np.int32)
    cols = []
    for i in range(n):
        cols.extend(sorted(rows[i]))
        rp[i + 1] = rp[i] + len(rows[i])
    return rp, np.array(cols, np.int32)
