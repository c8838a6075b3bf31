"""ctypes bindings of the oracle (oracle/liboracle.so) and of the reference build (oracle/_ref/libref_rcm.so).

Test infrastructure: the checker for the HIP engine.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")
REF_LIB = os.path.join(ORACLE_DIR, "_ref", "libref_rcm.so")

I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)


def _p32(a):
    return a.ctypes.data_as(I32P)


_orc = None
_ref = None


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True, capture_output=True)


def orc() -> C.CDLL:
    global _orc
    if _orc is None:
        src = os.path.join(ORACLE_DIR, "rcm_oracle.c")
        if not os.path.exists(ORACLE_LIB) or os.path.getmtime(ORACLE_LIB) < os.path.getmtime(src):
            build_oracle()
        L = C.CDLL(ORACLE_LIB)
        L.orc_cm.restype = C.c_int32
        L.orc_cm.argtypes = [C.c_int32, I32P, I32P, C.c_int32, I32P]
        L.orc_bfs.restype = C.c_int32
        L.orc_bfs.argtypes = [C.c_int32, I32P, I32P, C.c_int32, I32P, I32P]
        L.orc_peripheral.restype = C.c_int32
        L.orc_peripheral.argtypes = [C.c_int32, I32P, I32P, I32P, I32P, I32P]
        L.orc_bandwidth.restype = C.c_int64
        L.orc_bandwidth.argtypes = [C.c_int32, I32P, I32P, I32P, C.c_int32]
        L.orc_bandwidth_identity.restype = C.c_int64
        L.orc_bandwidth_identity.argtypes = [C.c_int32, I32P, I32P]
        L.orc_check_perm.restype = C.c_int32
        L.orc_check_perm.argtypes = [I32P, C.c_int32]
        L.orc_verify_order.restype = C.c_int64
        L.orc_verify_order.argtypes = [C.c_int32, I32P, I32P, C.c_int32, I32P, C.c_int32, C.c_int32]
        L.orc_prepare.restype = C.c_int64
        L.orc_prepare.argtypes = [C.c_int32, C.c_int64, I32P, I32P, C.c_int32, C.POINTER(I32P), C.POINTER(I32P), I64P]
        DP = C.POINTER(C.c_double)
        L.orc_prepare_values.restype = C.c_int64
        L.orc_prepare_values.argtypes = [C.c_int32, C.c_int64, I32P, I32P, DP, C.c_int32, C.POINTER(I32P),
                                         C.POINTER(I32P), C.POINTER(DP), I64P]
        L.orc_convert.restype = C.c_int64
        L.orc_convert.argtypes = [C.c_int32, C.c_int64, I32P, I32P, DP, C.POINTER(I32P), C.POINTER(I32P),
                                  C.POINTER(DP)]
        L.orc_free.restype = None
        L.orc_free.argtypes = [C.c_void_p]
        _orc = L
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref() -> C.CDLL:
    global _ref
    if _ref is None:
        L = C.CDLL(REF_LIB)
        L.ref_rcm.restype = C.c_int64
        L.ref_rcm.argtypes = [C.c_int32, C.c_int64, I32P, I32P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, I32P,
                              C.POINTER(C.c_double), C.c_int32]
        L.ref_prepare.restype = C.c_int64
        L.ref_prepare.argtypes = [C.c_int32, C.c_int64, I32P, I32P, C.c_int32, C.POINTER(I32P), C.POINTER(I32P),
                                  I64P, C.c_int32]
        L.ref_free.restype = None
        L.ref_free.argtypes = [I32P]
        DP = C.POINTER(C.c_double)
        out3 = [I32P, I32P, C.POINTER(I32P), C.POINTER(I32P), C.POINTER(DP)]
        L.ref_mtx_load.restype = C.c_int64
        L.ref_mtx_load.argtypes = [C.c_char_p, C.c_int32, *out3, C.POINTER(C.c_uint32), C.c_int32]
        L.ref_mtx_store.restype = C.c_int32
        L.ref_mtx_store.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int64, I32P, I32P, DP]
        L.ref_hicsr_load.restype = C.c_int64
        L.ref_hicsr_load.argtypes = [C.c_char_p, C.c_int32, *out3]
        L.ref_hicsr_store.restype = C.c_int32
        L.ref_hicsr_store.argtypes = [C.c_char_p, C.c_int32, C.c_int32, C.c_int32, C.c_int64, I32P, I32P, DP]
        L.ref_graph_load.restype = C.c_int64
        L.ref_graph_load.argtypes = [C.c_char_p, C.c_int32, *out3, I32P, C.c_int32]
        L.ref_prepare_values.restype = C.c_int64
        L.ref_prepare_values.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int64, I32P, I32P, DP, C.c_int32,
                                         *out3, C.c_int32]
        L.ref_free_values.restype = None
        L.ref_free_values.argtypes = [DP]
        _ref = L
    return _ref


# ---- thin numpy wrappers ---------------------------------------------------------------------------------
def as_csr(rp, ci):
    return np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32)


def cm(rp, ci, start):
    rp, ci = as_csr(rp, ci)
    n = rp.shape[0] - 1
    out = np.empty(n, np.int32)
    k = orc().orc_cm(n, _p32(rp), _p32(ci), int(start), _p32(out))
    assert k >= 0
    return out[:k]


def bfs(rp, ci, start):
    rp, ci = as_csr(rp, ci)
    n = rp.shape[0] - 1
    out = np.empty(n, np.int32)
    lev = np.empty(n, np.int32)
    k = orc().orc_bfs(n, _p32(rp), _p32(ci), int(start), _p32(out), _p32(lev))
    assert k >= 0
    return out[:k], lev[:k]


def peripheral(rp, ci):
    """(status, start, iterations, eccentricity)"""
    rp, ci = as_csr(rp, ci)
    n = rp.shape[0] - 1
    s, it, ecc = C.c_int32(-1), C.c_int32(-1), C.c_int32(-1)
    rc = orc().orc_peripheral(n, _p32(rp), _p32(ci), C.byref(s), C.byref(it), C.byref(ecc))
    return rc, s.value, it.value, ecc.value


def bandwidth(rp, ci, perm):
    rp, ci = as_csr(rp, ci)
    perm = np.ascontiguousarray(perm, np.int32)
    return int(orc().orc_bandwidth(rp.shape[0] - 1, _p32(rp), _p32(ci), _p32(perm), perm.shape[0]))


def bandwidth_identity(rp, ci):
    rp, ci = as_csr(rp, ci)
    return int(orc().orc_bandwidth_identity(rp.shape[0] - 1, _p32(rp), _p32(ci)))


def check_perm(perm):
    perm = np.ascontiguousarray(perm, np.int32)
    return int(orc().orc_check_perm(_p32(perm), perm.shape[0]))


def verify_order(rp, ci, start, perm, sorted_=True):
    rp, ci = as_csr(rp, ci)
    perm = np.ascontiguousarray(perm, np.int32)
    return int(orc().orc_verify_order(rp.shape[0] - 1, _p32(rp), _p32(ci), int(start), _p32(perm), perm.shape[0],
                                      1 if sorted_ else 0))


def prepare(rp, ci, check_sym):
    """Oracle prepareRCM: (status_or_added, rp, ci)."""
    rp, ci = as_csr(rp, ci)
    n = rp.shape[0] - 1
    orp, oci, onnz = I32P(), I32P(), C.c_int64()
    rc = orc().orc_prepare(n, ci.shape[0], _p32(rp), _p32(ci), 1 if check_sym else 0, C.byref(orp), C.byref(oci),
                           C.byref(onnz))
    if rc < 0:
        return int(rc), None, None
    r = np.ctypeslib.as_array(orp, shape=(n + 1,)).copy()
    c = np.ctypeslib.as_array(oci, shape=(max(onnz.value, 1),))[: onnz.value].copy()
    orc().orc_free(C.cast(orp, C.c_void_p))
    orc().orc_free(C.cast(oci, C.c_void_p))
    return int(rc), r, c


def prepare_values(rp, ci, vals, check_sym):
    """Oracle prepareRCM with float64 values: (status_or_added, rp, ci, vals)."""
    rp, ci = as_csr(rp, ci)
    v = np.ascontiguousarray(vals, np.float64)
    n = rp.shape[0] - 1
    orp, oci, ov, onnz = I32P(), I32P(), C.POINTER(C.c_double)(), C.c_int64()
    rc = orc().orc_prepare_values(n, ci.shape[0], _p32(rp), _p32(ci), v.ctypes.data_as(C.POINTER(C.c_double)),
                                  1 if check_sym else 0, C.byref(orp), C.byref(oci), C.byref(ov), C.byref(onnz))
    if rc < 0:
        return int(rc), None, None, None
    m = onnz.value
    r = np.ctypeslib.as_array(orp, shape=(n + 1,)).copy()
    c = np.ctypeslib.as_array(oci, shape=(max(m, 1),))[:m].copy()
    w = np.ctypeslib.as_array(ov, shape=(max(m, 1),))[:m].copy()
    for p in (orp, oci, ov):
        orc().orc_free(C.cast(p, C.c_void_p))
    return int(rc), r, c, w


def convert(rows, row_ids, col_ids, vals=None):
    """Oracle convert (COO -> CSR sorted by (row, col), duplicates in input order): (rp, ci, vals or None)."""
    r = np.ascontiguousarray(row_ids, np.int32)
    c = np.ascontiguousarray(col_ids, np.int32)
    v = None if vals is None else np.ascontiguousarray(vals, np.float64)
    orp, oci, ov = I32P(), I32P(), C.POINTER(C.c_double)()
    m = orc().orc_convert(int(rows), r.shape[0], _p32(r), _p32(c),
                          None if v is None else v.ctypes.data_as(C.POINTER(C.c_double)), C.byref(orp), C.byref(oci),
                          C.byref(ov))
    assert m >= 0
    rp = np.ctypeslib.as_array(orp, shape=(rows + 1,)).copy()
    ci = np.ctypeslib.as_array(oci, shape=(max(m, 1),))[:m].copy()
    orc().orc_free(C.cast(orp, C.c_void_p))
    orc().orc_free(C.cast(oci, C.c_void_p))
    w = None
    if v is not None:
        w = np.ctypeslib.as_array(ov, shape=(max(m, 1),))[:m].copy()
        orc().orc_free(C.cast(ov, C.c_void_p))
    return rp, ci, w


def ref_rcm(rp, ci, start, threads=1, batchsize=1, stable=True):
    """The reference's own RealCPU::rcm (25 timed iterations inside). Returns (perm, mean_ms)."""
    rp, ci = as_csr(rp, ci)
    n = rp.shape[0] - 1
    out = np.empty(n, np.int32)
    ms = C.c_double(-1)
    k = ref().ref_rcm(n, ci.shape[0], _p32(rp), _p32(ci), int(start), threads, batchsize, 1 if stable else 0,
                      _p32(out), C.byref(ms), 1)
    assert k >= 0
    return out[:k], ms.value


def ref_prepare(rp, ci, check_sym):
    rp, ci = as_csr(rp, ci)
    n = rp.shape[0] - 1
    orp, oci, onnz = I32P(), I32P(), C.c_int64()
    rows = ref().ref_prepare(n, ci.shape[0], _p32(rp), _p32(ci), 1 if check_sym else 0, C.byref(orp), C.byref(oci),
                             C.byref(onnz), 1)
    r = np.ctypeslib.as_array(orp, shape=(rows + 1,)).copy()
    c = np.ctypeslib.as_array(oci, shape=(max(onnz.value, 1),))[: onnz.value].copy()
    ref().ref_free(orp)
    ref().ref_free(oci)
    return r, c


# ---- the reference's file formats (src/COO.cpp, src/CSR.cpp, include/Graph.h) ----------------------------
def _take_ref_csr(nnz, rows, cols, rp, ci, vals):
    r = np.ctypeslib.as_array(rp, shape=(rows.value + 1,)).copy()
    c = np.ctypeslib.as_array(ci, shape=(max(nnz, 1),))[:nnz].copy()
    v = np.ctypeslib.as_array(vals, shape=(max(nnz, 1),))[:nnz].copy()
    ref().ref_free(rp)
    ref().ref_free(ci)
    ref().ref_free_values(vals)
    return rows.value, cols.value, r, c, v


def _out3():
    return C.c_int32(), C.c_int32(), I32P(), I32P(), C.POINTER(C.c_double)()


def ref_mtx_load(path, is_double=False):
    """loadMTX + convert. Returns None where the reference throws, else (rows, cols, rp, ci, vals, info)."""
    rows, cols, rp, ci, vals = _out3()
    info = C.c_uint32()
    nnz = ref().ref_mtx_load(os.fsencode(path), int(is_double), C.byref(rows), C.byref(cols), C.byref(rp),
                             C.byref(ci), C.byref(vals), C.byref(info), 1)
    if nnz < 0:
        return None
    return (*_take_ref_csr(nnz, rows, cols, rp, ci, vals), info.value)


def _dp(v):
    return v.ctypes.data_as(C.POINTER(C.c_double))


def ref_mtx_store(path, rows, cols, rp, ci, vals, is_double=False):
    rp, ci = as_csr(rp, ci)
    vals = np.ascontiguousarray(vals, np.float64)
    return int(ref().ref_mtx_store(os.fsencode(path), int(is_double), rows, cols, ci.shape[0], _p32(rp), _p32(ci),
                                   _dp(vals)))


def ref_hicsr_load(path, is_double=False):
    rows, cols, rp, ci, vals = _out3()
    nnz = ref().ref_hicsr_load(os.fsencode(path), int(is_double), C.byref(rows), C.byref(cols), C.byref(rp),
                               C.byref(ci), C.byref(vals))
    if nnz < 0:
        return None
    return _take_ref_csr(nnz, rows, cols, rp, ci, vals)


def ref_hicsr_store(path, rows, cols, rp, ci, vals, is_double=False):
    rp, ci = as_csr(rp, ci)
    vals = np.ascontiguousarray(vals, np.float64)
    return int(ref().ref_hicsr_store(os.fsencode(path), int(is_double), rows, cols, ci.shape[0], _p32(rp), _p32(ci),
                                     _dp(vals)))


def ref_graph_load(file, is_double=False):
    """Graph<T>::load. Returns (ret, rows, cols, rp, ci, vals) or (ret, None...) when it failed."""
    rows, cols, rp, ci, vals = _out3()
    ret = C.c_int32(-1)
    nnz = ref().ref_graph_load(os.fsencode(file), int(is_double), C.byref(rows), C.byref(cols), C.byref(rp),
                               C.byref(ci), C.byref(vals), C.byref(ret), 1)
    if nnz < 0:
        return ret.value, None
    return ret.value, _take_ref_csr(nnz, rows, cols, rp, ci, vals)


def ref_prepare_values(rows, cols, rp, ci, vals, check_sym, is_double=False):
    rp, ci = as_csr(rp, ci)
    vals = np.ascontiguousarray(vals, np.float64)
    orows, ocols, orp, oci, ovals = _out3()
    nnz = ref().ref_prepare_values(int(is_double), rows, cols, ci.shape[0], _p32(rp), _p32(ci), _dp(vals),
                                   1 if check_sym else 0, C.byref(orows), C.byref(ocols), C.byref(orp), C.byref(oci),
                                   C.byref(ovals), 1)
    return _take_ref_csr(nnz, orows, ocols, orp, oci, ovals)
