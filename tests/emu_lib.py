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
        if engine == "many":
            _libs[engine] = L
            return L
        L.emu_run.restype = C.c_int
        L.emu_run.argtypes = [C.c_int32, C.c_int64, I32P, I32P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                              C.c_int32, I32P, I32P, I32P, C.c_double, C.c_int32]
        _libs[engine] = L
    return _libs[engine]


def run(rp, ci, start, workers=1, slots=None, sort_mode=0, bfs=False, timeout_s=20.0, engine="block",
        claim_filter=0):
    """Returns (perm, level, info) with info = [error, placed, batches, diag...]."""
    if slots is None:
        slots = 2048 if engine == "block" else 1024
    rp = np.ascontiguousarray(rp, np.int32)
    ci = np.ascontiguousarray(ci, np.int32)
    n = rp.shape[0] - 1
    perm = np.full(n, -1, np.int32)
    lev = np.full(n, -1, np.int32)
    info = np.zeros(12, np.int32)
    rc = lib(engine).emu_run(n, ci.shape[0], rp.ctypes.data_as(I32P), ci.ctypes.data_as(I32P), int(start), workers,
                             slots, sort_mode, 1 if bfs else 0, perm.ctypes.data_as(I32P),
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


def run_many(mats, starts, workers=2, bfs=False, timeout_s=30.0):
    """mats: list of (rp, ci). Returns (list of perms, list of levels, info[count, 4], jobs_done)."""
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
                                    workers, 1 if bfs else 0, perm.ctypes.data_as(I32P), lev.ctypes.data_as(I32P),
                                    info.ctypes.data_as(I32P), timeout_s)
    offs = np.concatenate([[0], np.cumsum(n)])
    perms = [perm[offs[k]:offs[k + 1]] for k in range(len(mats))]
    levs = [lev[offs[k]:offs[k + 1]] for k in range(len(mats))]
    return perms, levs, info, done
