"""ctypes wrapper of the CPU wave emulator builds of the engines (tests/emu). Test infrastructure only.

engine="block": engine.hpp, the library's 256-thread workgroup engine (2048 adjacency slots per batch).
engine="wave":  engine_wave.hpp, the one-wave worker variant (512, 1024 or 2048 slots), not in the library.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "emu")
I32P = C.POINTER(C.c_int32)
_libs: dict = {}


def lib(engine: str = "block"):
    if engine not in _libs:
        subprocess.run(["make", "-s", "-C", EMU_DIR], check=True)
        L = C.CDLL(os.path.join(EMU_DIR, f"libemu_{engine}.so"))
        if engine in ("many", "ops"):
            _libs[engine] = L
            return L
        L.emu_run.restype = C.c_int
        L.emu_run.argtypes = [C.c_int32, C.c_int64, I32P, I32P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                              C.c_int32, I32P, I32P, I32P, C.c_double, C.c_int32]
        _libs[engine] = L
    return _libs[engine]


def run(rp, ci, start, workers=1, slots=None, sort_mode=0, bfs=False, timeout_s=60.0, engine="block",
        claim_filter=0, levels=False):
    """Returns (perm, level, info) with info = [error, placed, batches, diag...]. levels: a CM run that also
    records BFS levels (block engine)."""
    if slots is None:
        slots = 2048 if engine == "block" else 1024
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    n = rp.shape[0] - 1
    perm = np.full(n, -1, np.int32)
    lev = np.full(n, -1, np.int32)
    info = np.zeros(12, np.int32)
    rc = lib(engine).emu_run(n, ci.shape[0], rp.ctypes.data_as(I32P), ci.ctypes.data_as(I32P), int(start), workers,
                             slots, sort_mode, 1 if bfs else (2 if levels else 0), perm.ctypes.data_as(I32P),
                             lev.ctypes.data_as(I32P), info.ctypes.data_as(I32P), timeout_s, claim_filter)
    assert rc == 0, "emulator rejected the arguments"
    return perm, lev, info


def _many_lib():
    L = lib("many")
    if not hasattr(L, "_many_ready"):
        I64P = C.POINTER(C.c_int64)
        L.emu_run_many.restype = C.c_int
        L.emu_run_many.argtypes = [C.c_int32, I32P, I64P, I64P, I64P, I32P, I32P, I32P, C.c_int32, C.c_int32, I32P,
                                   I32P, I32P, C.c_double]
        L._many_ready = True
    return L


def run_many(mats, starts, workers=2, bfs=False, timeout_s=60.0, levels=False):
    """mats: list of (rp, ci). Returns (list of perms, list of levels, info[count, 4], jobs_done). levels: CM
    runs that also record BFS levels."""
    I64P = C.POINTER(C.c_int64)
    n = np.array([m[0].shape[0] - 1 for m in mats], np.int32)
    nnz = np.array([m[1].shape[0] for m in mats], np.int64)
    rp_off = np.concatenate([[0], np.cumsum(n.astype(np.int64) + 1)[:-1]]).astype(np.int64)
    ci_off = np.concatenate([[0], np.cumsum(nnz)[:-1]]).astype(np.int64)
    rp_cat = np.concatenate([np.asarray(m[0], np.int32) for m in mats])
    ci_cat = np.concatenate([np.asarray(m[1], np.int32) for m in mats] + [np.zeros(1, np.int32)])
    st = np.ascontiguousarray(starts, np.int32)
    tot = int(n.sum())
    perm = np.full(tot, -1, np.int32)
    lev = np.full(tot, -1, np.int32)
    info = np.zeros((len(mats), 4), np.int32)
    done = _many_lib().emu_run_many(len(mats), n.ctypes.data_as(I32P), nnz.ctypes.data_as(I64P),
                                    rp_off.ctypes.data_as(I64P), ci_off.ctypes.data_as(I64P),
                                    rp_cat.ctypes.data_as(I32P), ci_cat.ctypes.data_as(I32P), st.ctypes.data_as(I32P),
                                    workers, 1 if bfs else (2 if levels else 0), perm.ctypes.data_as(I32P),
                                    lev.ctypes.data_as(I32P),
                                    info.ctypes.data_as(I32P), timeout_s)
    offs = np.concatenate([[0], np.cumsum(n)])
    perms = [perm[offs[k]:offs[k + 1]] for k in range(len(mats))]
    levs = [lev[offs[k]:offs[k + 1]] for k in range(len(mats))]
    return perms, levs, info, done


# ---- matrix_ops.hpp kernels (canonicalisation, permutation, radix sort, scan) ----------------------------
def _ops_lib():
    L = lib("ops")
    if not hasattr(L, "_ops_ready"):
        I64P = C.POINTER(C.c_int64)
        DP = C.POINTER(C.c_double)
        U64P = C.POINTER(C.c_uint64)
        U32P = C.POINTER(C.c_uint32)
        L.emu_permute.restype = C.c_int
        L.emu_permute.argtypes = [C.c_int32, C.c_int64, I32P, I32P, DP, I32P, I32P, I32P, DP]
        L.emu_canonicalize.restype = C.c_int
        L.emu_canonicalize.argtypes = [C.c_int32, C.c_int32, C.c_int64, I32P, I32P, DP, C.c_int32, C.c_int64, I32P,
                                       I64P, I64P, I32P, I32P, DP]
        L.emu_radix_sort.restype = C.c_int
        L.emu_radix_sort.argtypes = [C.c_int64, U64P, U32P, C.c_int32, C.c_int32]
        L.emu_level_order.restype = C.c_int
        L.emu_level_order.argtypes = [C.c_int32, C.c_int64, I32P, I32P, C.c_int32, C.c_int32, I32P, I32P, I32P, I32P]
        L.emu_scan.restype = C.c_int
        L.emu_scan.argtypes = [C.c_int64, U32P, U32P]
        L._ops_ready = True
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def scan(a):
    a = np.ascontiguousarray(a, np.uint32)
    out = np.empty(a.size + 1, np.uint32)
    _ops_lib().emu_scan(a.size, a.ctypes.data_as(C.POINTER(C.c_uint32)), out.ctypes.data_as(C.POINTER(C.c_uint32)))
    return out


def radix_sort(keys, vals, lo_bits, hi_bits):
    k = np.ascontiguousarray(keys, np.uint64).copy()
    v = np.ascontiguousarray(vals, np.uint32).copy()
    _ops_lib().emu_radix_sort(k.size, k.ctypes.data_as(C.POINTER(C.c_uint64)), v.ctypes.data_as(C.POINTER(C.c_uint32)),
                              lo_bits, hi_bits)
    return k, v


def permute(rp, ci, vals, perm):
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    perm = np.ascontiguousarray(perm, np.int32)
    vals = None if vals is None else np.ascontiguousarray(vals, np.float64)
    n, nnz = rp.size - 1, ci.size
    rp2 = np.empty(n + 1, np.int32)
    ci2 = np.empty(max(nnz, 1), np.int32)
    v2 = None if vals is None else np.empty(max(nnz, 1), np.float64)
    rc = _ops_lib().emu_permute(n, nnz, rp.ctypes.data_as(I32P), ci.ctypes.data_as(I32P), _dp(vals),
                                perm.ctypes.data_as(I32P), rp2.ctypes.data_as(I32P), ci2.ctypes.data_as(I32P), _dp(v2))
    return rc, rp2, ci2[:nnz], (None if v2 is None else v2[:nnz])


def canonicalize(rows, cols, r, c, vals, check_sym):
    r = np.ascontiguousarray(r, np.int32)
    c = np.ascontiguousarray(c, np.int32)
    vals = None if vals is None else np.ascontiguousarray(vals, np.float64)
    cap = 2 * r.size + 2 * max(rows, cols) + 16
    rp = np.empty(max(rows, cols) + 1, np.int32)
    ci = np.empty(cap, np.int32)
    v = None if vals is None else np.empty(cap, np.float64)
    n, nnz, links = C.c_int32(), C.c_int64(), C.c_int64()
    rc = _ops_lib().emu_canonicalize(rows, cols, r.size, r.ctypes.data_as(I32P), c.ctypes.data_as(I32P), _dp(vals),
                                     1 if check_sym else 0, cap, C.byref(n), C.byref(nnz), C.byref(links),
                                     rp.ctypes.data_as(I32P), ci.ctypes.data_as(I32P), _dp(v))
    if rc:
        return rc, None
    m = nnz.value
    return 0, (n.value, rp[: n.value + 1].copy(), ci[:m].copy(), None if v is None else v[:m].copy(), links.value)


def level_order(rp, ci, start, bfs=False):
    """Level-synchronous engine (level_cm.hpp): (status, perm[:placed], level[:placed] or None, levels)."""
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    n = rp.size - 1
    perm = np.full(n, -1, np.int32)
    lev = np.full(n, -1, np.int32)
    placed, levels = C.c_int32(), C.c_int32()
    rc = _ops_lib().emu_level_order(n, ci.size, rp.ctypes.data_as(I32P), ci.ctypes.data_as(I32P), int(start),
                                    1 if bfs else 0, perm.ctypes.data_as(I32P), lev.ctypes.data_as(I32P),
                                    C.byref(placed), C.byref(levels))
    k = placed.value
    return rc, perm[:k], (lev[:k] if bfs else None), levels.value


def run_wg(mats, starts, workers=3, bfs=False):
    """One workgroup per matrix (engine_wg.hpp). mats: list of (rp, ci). Returns (rc, [perm], [level], info)."""
    L = _many_lib()
    if not hasattr(L, "_wg_ready"):
        I64P = C.POINTER(C.c_int64)
        L.emu_run_wg.restype = C.c_int
        L.emu_run_wg.argtypes = [C.c_int32, I32P, I64P, I64P, I32P, I32P, I32P, C.c_int32, C.c_int32, I32P, I32P,
                                 I32P]
        L._wg_ready = True
    n = np.array([rp.shape[0] - 1 for rp, _ in mats], np.int32)
    rp_off = np.concatenate([[0], np.cumsum([rp.shape[0] for rp, _ in mats])[:-1]]).astype(np.int64)
    ci_off = np.concatenate([[0], np.cumsum([ci.shape[0] for _, ci in mats])[:-1]]).astype(np.int64)
    rpc = np.ascontiguousarray(np.concatenate([rp for rp, _ in mats]), np.int32)
    cic = np.ascontiguousarray(np.concatenate([ci for _, ci in mats]), np.int32)
    st = np.ascontiguousarray(starts, np.int32)
    tot = int(n.sum())
    perm = np.full(tot, -1, np.int32)
    lev = np.full(tot, -1, np.int32)
    info = np.zeros(2 * len(mats), np.int32)
    I64 = C.POINTER(C.c_int64)
    rc = L.emu_run_wg(len(mats), n.ctypes.data_as(I32P), rp_off.ctypes.data_as(I64), ci_off.ctypes.data_as(I64),
                      rpc.ctypes.data_as(I32P), cic.ctypes.data_as(I32P), st.ctypes.data_as(I32P), workers,
                      1 if bfs else 0, perm.ctypes.data_as(I32P), lev.ctypes.data_as(I32P), info.ctypes.data_as(I32P))
    off = np.concatenate([[0], np.cumsum(n)])
    return rc, [perm[off[k]:off[k + 1]] for k in range(len(mats))], [lev[off[k]:off[k + 1]] for k in range(len(mats))], info
