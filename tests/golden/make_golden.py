"""Writes the golden fixtures of this directory from the reference's own CPU code (oracle/_ref/libref_rcm.so,
built by oracle/Makefile from the reference sources): inputs, the reference's RealCPU::rcm<STABLE=true> order
from the oracle's pseudo-peripheral start, and the reference's Graph<float>::prepareRCM repair.

Run in the build container (where the reference tree is present):  python tests/golden/make_golden.py
The fixtures are data (numpy .npz, no pickles); tests/test_golden.py checks the oracle and the GPU against them.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import graphs as G  # noqa: E402
import oracle_lib as O  # noqa: E402
import parallelbatchrcm_amd as P  # noqa: E402


def cases():
    out = {
        "c1_stencil2d_100": P.workloads.config(1, seed=1),
        "c2_stencil3d_side20": P.workloads.config(2, scale=20, seed=2),
        "c3_rgg_8000": P.workloads.config(3, scale=8000, seed=3),
        "hub_3000": P.CSR(*G.hub_graph(3000, 4000, seed=11)),
        "dups_unsorted_400": P.CSR(*G.shuffle_rows(*G.with_duplicates(*G.random_connected(400, 900, seed=12),
                                                                      seed=13), seed=14)),
    }
    m = P.connect_components(P.workloads.rmat(12, seed=15))[0]
    out["rmat12"] = m
    return out


def main() -> int:
    if not O.ref_available():
        print("the reference build oracle/_ref/libref_rcm.so is missing (run make -C oracle here)")
        return 1
    for name, m in cases().items():
        rp, ci = m.row_offsets, m.col_ids
        rc, start, _, _ = O.peripheral(rp, ci)
        assert rc == 0, name
        perm, _ = O.ref_rcm(rp, ci, start, threads=1, batchsize=1, stable=True)
        assert perm.shape[0] == rp.shape[0] - 1, name
        bw = O.bandwidth(rp, ci, perm)
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), row_ptr=rp, col_idx=ci, start=np.int32(start),
                            perm=perm.astype(np.int32), bandwidth=np.int64(bw))
        print(f"{name}: n={rp.shape[0] - 1} nnz={ci.shape[0]} start={start} bandwidth={bw}")
    # prepareRCM: a disconnected, non-symmetric input and the reference's repaired CSR
    rng = np.random.default_rng(16)
    n = 600
    e = [(int(a), int(b)) for a, b in rng.integers(0, n, (700, 2)) if a != b]
    rp, ci = G.from_edges(n, e, diag=True, sym=False)
    rrp, rci = O.ref_prepare(rp, ci, True)
    np.savez_compressed(os.path.join(HERE, "prepare_nonsym_600.npz"), row_ptr=rp, col_idx=ci, out_row_ptr=rrp,
                        out_col_idx=rci)
    print(f"prepare_nonsym_600: n={n} nnz={ci.shape[0]} -> nnz {rci.shape[0]}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
