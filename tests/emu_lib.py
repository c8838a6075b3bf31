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
        L.emu_run.restype = C.c_int
        L.emu_run.argtypes = [C.c_int32, C.c_int64, I32P, I32P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                              C.c_int32, I32P, I32P, I32P, C.c_double]
        _libs[engine] = L
    return _libs[engine]


def run(rp, ci, start, workers=1, slots=None, sort_mode=0, bfs=False, timeout_s=20.0, engine="block"):
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
                             lev.ctypes.data_as(I32P), info.ctypes.data_as(I32P), timeout_s)
    assert rc == 0, "emulator rejected the arguments"
    return perm, lev, info
