"""Writes the file-ingest fixtures of this directory: small MatrixMarket files and what the reference's own
Graph<double>::load (include/Graph.h:266-324: loadMTX, convert, prepareRCM, with its .hicsr / .rcm.hicsr caches)
makes of them, through oracle/_ref/libref_rcm.so built from the reference sources by oracle/Makefile.

Run in the build container (where the reference tree is present):  python tests/golden/make_ingest_golden.py
Each fixture (numpy .npz, no pickles) holds the file text and the reference's loaded CSR with values;
tests/test_gpu_ingest.py feeds the files through the host parser and the GPU canonicalisation and reorder.
"""
from __future__ import annotations

import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_lib as O  # noqa: E402


def mtx_text(banner: str, n: int, entries) -> str:
    lines = [f"%%MatrixMarket matrix coordinate {banner}", "% file-ingest fixture", f"{n} {n} {len(entries)}"]
    for e in entries:
        lines.append(" ".join(str(x) for x in e))
    return "\n".join(lines) + "\n"


def grid_lower(nx, ny, rng, pattern=False, drop=()):
    """Lower triangle of a 2D 5-point grid (1-based), nodes in `drop` cut loose (diagonal kept)."""
    idx = lambda x, y: y * nx + x  # noqa: E731
    out = []
    for y in range(ny):
        for x in range(nx):
            i = idx(x, y)
            out.append((i, i))
            for j in ([idx(x - 1, y)] if x else []) + ([idx(x, y - 1)] if y else []):
                if i in drop or j in drop:
                    continue
                out.append((i, j))
    rng.shuffle(out)
    if pattern:
        return [(i + 1, j + 1) for i, j in out]
    return [(i + 1, j + 1, f"{v:.6f}") for (i, j), v in zip(out, rng.standard_normal(len(out)))]


def cases():
    rng = np.random.default_rng(31)
    out = {}
    out["ingest_sym_grid"] = mtx_text("real symmetric", 40 * 50, grid_lower(40, 50, rng, drop={77, 901, 1502}))
    out["ingest_pattern_sym"] = mtx_text("pattern symmetric", 30 * 30, grid_lower(30, 30, rng, pattern=True,
                                                                                 drop={0, 450}))
    n = 1200  # general storage: one-sided entries the symmetrisation mirrors, a full diagonal, three clusters
    seen = set()
    ent = []
    for i in range(n):
        seen.add((i, i))
        ent.append((i, i))
    for _ in range(2 * n):
        a = int(rng.integers(0, n))
        b = int(a + rng.integers(-15, 16))
        if 0 <= b < n and b // 400 == a // 400 and (a, b) not in seen and (b, a) not in seen:
            seen.add((a, b))
            ent.append((a, b))
    rng.shuffle(ent)
    vals = rng.standard_normal(len(ent))
    out["ingest_general_nonsym"] = mtx_text("real general", n, [(a + 1, b + 1, f"{v:.6f}") for (a, b), v in
                                                                   zip(ent, vals)])
    return out


def main() -> int:
    if not O.ref_available():
        print("the reference build oracle/_ref/libref_rcm.so is missing (run make -C oracle here)")
        return 1
    for name, text in cases().items():
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "g.mtx")
            with open(path, "w") as f:
                f.write(text)
            ret, ref = O.ref_graph_load(path, True)
            assert ret == 0 and ref is not None, name
            rows, cols, rp, ci, vals = ref
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), text=np.frombuffer(text.encode(), np.uint8),
                            row_ptr=rp.astype(np.int32), col_idx=ci.astype(np.int32), values=vals)
        print(f"{name}: n={rows} nnz={ci.shape[0]} ({len(text)} bytes of text)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
