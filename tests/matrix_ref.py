"""numpy restatements used as checkers for the canonicalisation and permutation kernels. Test infrastructure.

permute_ref: B = P A P^T as the definition states it -- entry k = (r, c, v) of A lands in row inv[r] of B with
column inv[c]; rows sorted by column, equal columns in their order in A (a lexsort on (inv[r], inv[c], k)).
"""
from __future__ import annotations

import numpy as np


def rows_of(rp):
    rp = np.asarray(rp, np.int64)
    return np.repeat(np.arange(rp.size - 1), np.diff(rp))


def permute_ref(rp, ci, vals, perm):
    rp = np.asarray(rp, np.int64)
    ci = np.asarray(ci, np.int64)
    perm = np.asarray(perm, np.int64)
    n = rp.size - 1
    inv = np.empty(n, np.int64)
    inv[perm] = np.arange(n)
    r = inv[rows_of(rp)]
    c = inv[ci]
    order = np.lexsort((np.arange(ci.size), c, r))
    rp2 = np.concatenate([[0], np.cumsum(np.bincount(r, minlength=n))]).astype(np.int32)
    v2 = None if vals is None else np.asarray(vals)[order]
    return rp2, c[order].astype(np.int32), v2


def dense(rp, ci, vals, n):
    d = np.zeros((n, n))
    np.add.at(d, (rows_of(rp), np.asarray(ci)), np.ones(len(ci)) if vals is None else np.asarray(vals, np.float64))
    return d


def coo_with_components(rng, n, k, comps=3, mirrored_dups=5):
    """A COO whose symmetrisation is structurally symmetric: unique unordered pairs stored one way, the other
    way, or both; a few pairs duplicated in both directions; the full diagonal; `comps` groups of nodes with no
    edges between groups. Entries shuffled."""
    lab = rng.integers(0, comps, n)
    r = rng.integers(0, n, k)
    c = rng.integers(0, n, k)
    keep = (lab[r] == lab[c]) & (r != c)
    r, c = r[keep], c[keep]
    key = np.unique(np.minimum(r, c).astype(np.int64) * n + np.maximum(r, c))
    lo, hi = key // n, key % n
    mode = rng.integers(0, 3, lo.size)  # 0: lo -> hi only, 1: hi -> lo only, 2: both
    R = np.concatenate([lo[mode != 1], hi[mode != 0]])
    C = np.concatenate([hi[mode != 1], lo[mode != 0]])
    both = np.flatnonzero(mode == 2)
    if both.size:
        d = rng.choice(both, size=min(mirrored_dups, both.size), replace=False)
        R = np.concatenate([R, lo[d], hi[d]])
        C = np.concatenate([C, hi[d], lo[d]])
    R = np.concatenate([R, np.arange(n)])
    C = np.concatenate([C, np.arange(n)])
    p = rng.permutation(R.size)
    return R[p].astype(np.int32), C[p].astype(np.int32)
