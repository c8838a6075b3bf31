"""parallelbatchrcm_amd -- MI355X-native Cuthill-McKee reordering (drop-in for the GPU_BATCH path).

Python host layer over the C ABI in include/rcm_gpu_batch_abi.h. It mirrors the reference's host interface for
this path (include/rcm_gpu_batch.h:40-41, include/simple_peripheral.h:45-46, src/main.cpp:143-226):

    GPU_BATCH.rcm(graph, csv_file, permutation, start_node)   -> CM order on the GPU
    SimplePeripheral.findPeripheral(graph, csv_file, writeOutput, outputnodes=32)
    computeBandwidth(graph, permutation), checkPermutation(permutation)

plus a device-resident `RcmContext`, the one-shot `rcm_gpu_batch`, the batched `rcm_gpu_batch_many`, and the
seeded workload generators (`workloads`). Every compute call goes to the HIP engine; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native
from ._native import (
    I32P,
    I64P,
    BatchTimingT,
    CsrT,
    JobStatsT,
    MatrixT,
    ServerConfigT,
    ServerStatsT,
    TimingT,
    TuneResultT,
)

RCM_FLAG_REVERSE = 1
RCM_FLAG_STABLE = 2
RCM_FLAG_BANDWIDTH = 4
RCM_FLAG_BORROW = 8

RCM_OK = 0
RCM_EINVAL = -1
RCM_ENOTCONNECTED = -2
RCM_ENOMEM = -3
RCM_EDEVICE = -4
RCM_ETIMEOUT = -5
RCM_ECAPACITY = -6
RCM_ENOTLOADED = -7
RCM_ECANCELED = -8
RCM_ENOTFOUND = -9
RCM_EAGAIN = -10
RCM_EIO = -11
RCM_ENOTSYMMETRIC = -12

MTX_PATTERN = 1
MTX_COMPLEX = 2
MTX_SYMMETRIC = 4
MTX_HERMITIAN = 8
SOURCE_RCM_HICSR = 1
SOURCE_HICSR = 2
SOURCE_MTX = 3

PARAM_WORKERS = 1
PARAM_TIMEOUT_MS = 2
PARAM_SORT = 3
PARAM_DIGIT_BITS = 4
PARAM_VARIANT = 5
PARAM_STRICT = 6
PARAM_PROFILE = 7
PARAM_GRAPH = 8
PARAM_CLAIM_FILTER = 9
PARAM_WG_MAX_N = 10
PARAM_HELPER_NODES = 11
PARAM_BATCH_CAP = 12
PARAM_SPECULATE = 13
PROFILE_PHASES = ("acquire", "parents", "claim", "finalise", "sort", "cut", "prefix", "write")


class RcmError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        msg = _native.lib().rcm_status_string(status).decode()
        super().__init__(f"{what}: {msg} ({status})" if what else f"{msg} ({status})")
        self.status = status


def _check(status: int, what: str) -> None:
    if status != RCM_OK:
        raise RcmError(status, what)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _p32(a: np.ndarray):
    return a.ctypes.data_as(I32P)


def _p64(a: np.ndarray):
    return a.ctypes.data_as(I64P)


# ---------------------------------------------------------------------------------------------------------
# Containers mirroring include/CSR.h:42-75 and include/Graph.h:42-70
# ---------------------------------------------------------------------------------------------------------
@dataclass
class CSR:
    """CSR<T> (include/CSR.h:42-75): rows, cols, nnz, row_offsets[rows+1], col_ids[nnz], data[nnz]."""

    row_offsets: np.ndarray
    col_ids: np.ndarray
    data: Optional[np.ndarray] = None
    cols: int = -1

    def __post_init__(self):
        self.row_offsets = _i32(self.row_offsets)
        self.col_ids = _i32(self.col_ids)
        if self.cols < 0:
            self.cols = self.rows

    @property
    def rows(self) -> int:
        return int(self.row_offsets.shape[0]) - 1

    @property
    def nnz(self) -> int:
        return int(self.col_ids.shape[0])


@dataclass
class Graph:
    """Graph<T> (include/Graph.h:42-333): the CSR plus the host helpers the RCM path uses."""

    csr: CSR

    @property
    def size(self) -> int:
        return self.csr.cols

    def neighbour_count(self, node: int) -> int:  # Graph.h:58-61
        return int(self.csr.row_offsets[node + 1] - self.csr.row_offsets[node])

    def prepareRCM(self, check_sym: bool = True) -> int:
        """Graph<T>::prepareRCM (Graph.h:74-264): symmetrise (optional) and connect components.

        Returns the number of connecting edges added. Raises RcmError(RCM_EINVAL) where the reference's
        repair loop would not terminate (a component whose lowest node has no entries). Values, when the CSR
        carries them, follow their entries; added entries get the value 1, as in the reference.
        """
        m = self.csr
        if m.data is not None:
            self.csr, added = prepare_rcm(m, check_sym)
            return added
        if check_sym:
            m = symmetrize(m)
        m, added = connect_components(m)
        self.csr = m
        return added

    @classmethod
    def load(cls, file: str, dtype=np.float32, write_cache: bool = True) -> "Graph":
        """Graph<T>::load (Graph.h:266-324): `file`+ext+".rcm.hicsr", else +".hicsr", else the MatrixMarket
        file, prepared with prepareRCM(!symmetric) and cached as the reference caches it. ext is "" for
        float32 and "d_" for float64. Raises RcmError on failure; `graph.source` tells where it was found
        (SOURCE_RCM_HICSR, SOURCE_HICSR, SOURCE_MTX) and `graph.links` how many links prepareRCM added."""
        vb = _value_bytes(dtype)
        out = MatrixT()
        src = C.c_int32()
        links = C.c_int64()
        msg = C.create_string_buffer(512)
        rc = _native.lib().rcm_graph_load(
            _path(file), vb, 1 if write_cache else 0, C.byref(out), C.byref(src), C.byref(links), msg, len(msg)
        )
        if rc != RCM_OK:
            raise RcmError(rc, msg.value.decode(errors="replace") or "Graph.load")
        g = cls(_take_matrix(out, dtype))
        g.source = int(src.value)
        g.links = int(links.value)
        return g


# ---------------------------------------------------------------------------------------------------------
# Workload generators and host canonicalisation (include/rcm_workloads.h)
# ---------------------------------------------------------------------------------------------------------
def _take_csr(t: CsrT) -> CSR:
    n, nnz = t.n, t.nnz
    rp = np.ctypeslib.as_array(t.row_ptr, shape=(n + 1,)).copy()
    ci = np.ctypeslib.as_array(t.col_idx, shape=(max(nnz, 1),))[:nnz].copy() if nnz else np.zeros(0, np.int32)
    _native.lib().rcm_csr_free(C.byref(t))
    return CSR(rp, ci)


def _as_csr_t(m: CSR) -> CsrT:
    return CsrT(m.rows, m.nnz, _p32(m.row_offsets), _p32(m.col_ids))


class workloads:
    """Seeded synthetic inputs (splitmix64 + Fisher-Yates; deterministic across hosts)."""

    @staticmethod
    def stencil2d(nx: int, ny: int) -> CSR:
        t = CsrT()
        _check(_native.lib().rcm_gen_stencil2d(nx, ny, C.byref(t)), "stencil2d")
        return _take_csr(t)

    @staticmethod
    def stencil3d(nx: int, ny: int, nz: int) -> CSR:
        t = CsrT()
        _check(_native.lib().rcm_gen_stencil3d(nx, ny, nz, C.byref(t)), "stencil3d")
        return _take_csr(t)

    @staticmethod
    def rgg(npts: int, mean_degree: float, seed: int) -> CSR:
        t = CsrT()
        _check(_native.lib().rcm_gen_rgg(npts, float(mean_degree), seed & (2**64 - 1), C.byref(t)), "rgg")
        return _take_csr(t)

    @staticmethod
    def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
             seed: int = 1) -> CSR:
        t = CsrT()
        _check(_native.lib().rcm_gen_rmat(scale, edge_factor, a, b, c, seed & (2**64 - 1), C.byref(t)), "rmat")
        return _take_csr(t)

    @staticmethod
    def permute_symmetric(m: CSR, seed: int):
        """B = P A P^T with a seeded Fisher-Yates P; returns (B, q) where node i of A is node q[i] of B."""
        t = CsrT()
        q = np.empty(m.rows, np.int32)
        src = _as_csr_t(m)
        _check(_native.lib().rcm_permute_symmetric(C.byref(src), seed & (2**64 - 1), C.byref(t), _p32(q)), "permute")
        return _take_csr(t), q

    @staticmethod
    def config(which: int, scale: int = 0, seed: int = 0) -> CSR:
        """Benchmark configurations: 1 = 2D 5-point 100^2, 2 = 3D 7-point 128^3, 3 = RGG 10^6 points deg 12.
        `scale` shrinks them for tests (side length, or number of points)."""
        t = CsrT()
        _check(_native.lib().rcm_gen_config(which, scale, seed & (2**64 - 1), C.byref(t)), f"config {which}")
        return _take_csr(t)

    @staticmethod
    def batch_mix(count: int, n_lo: int = 1000, n_hi: int = 200000, seed: int = 4, threads: int = 0):
        """Batched mix (stencils and RGGs of log-uniform size). Returns a BatchSet."""
        lib = _native.lib()
        ns, nnzs, rpo, cio, rpc, cic = I32P(), I64P(), I64P(), I64P(), I32P(), I32P()
        _check(
            lib.rcm_gen_batch_mix(count, n_lo, n_hi, seed & (2**64 - 1), threads, C.byref(ns), C.byref(nnzs),
                                  C.byref(rpo), C.byref(cio), C.byref(rpc), C.byref(cic)),
            "batch_mix",
        )
        n = np.ctypeslib.as_array(ns, shape=(count,)).copy()
        nnz = np.ctypeslib.as_array(nnzs, shape=(count,)).copy()
        rp_off = np.ctypeslib.as_array(rpo, shape=(count,)).copy()
        ci_off = np.ctypeslib.as_array(cio, shape=(count,)).copy()
        trp = int(rp_off[-1] + n[-1] + 1)
        tci = int(ci_off[-1] + nnz[-1])
        row_ptr = np.ctypeslib.as_array(rpc, shape=(trp,)).copy()
        col_idx = np.ctypeslib.as_array(cic, shape=(max(tci, 1),))[:tci].copy()
        for p in (ns, nnzs, rpo, cio, rpc, cic):
            lib.rcm_free(C.cast(p, C.c_void_p))
        return BatchSet(n.astype(np.int32), nnz.astype(np.int64), rp_off.astype(np.int64), ci_off.astype(np.int64),
                        row_ptr, col_idx)


@dataclass
class BatchSet:
    """Matrices stored back to back: matrix k has n[k] rows, local row_ptr at row_ptr[rp_off[k]:...]."""

    n: np.ndarray
    nnz: np.ndarray
    rp_off: np.ndarray
    ci_off: np.ndarray
    row_ptr: np.ndarray
    col_idx: np.ndarray

    def __len__(self) -> int:
        return int(self.n.shape[0])

    def matrix(self, k: int) -> CSR:
        n, z = int(self.n[k]), int(self.nnz[k])
        rp = self.row_ptr[self.rp_off[k]: self.rp_off[k] + n + 1]
        ci = self.col_idx[self.ci_off[k]: self.ci_off[k] + z]
        return CSR(rp, ci)

    @property
    def perm_offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.n.astype(np.int64))[:-1]]).astype(np.int64)

    @staticmethod
    def from_matrices(mats: Sequence) -> "BatchSet":
        """Pack CSR matrices (or (row_ptr, col_idx) pairs) back to back."""
        pairs = [(m.row_offsets, m.col_ids) if isinstance(m, CSR) else (m[0], m[1]) for m in mats]
        n = np.array([len(rp) - 1 for rp, _ in pairs], np.int32)
        nnz = np.array([len(ci) for _, ci in pairs], np.int64)
        rp_off = np.concatenate([[0], np.cumsum(n.astype(np.int64) + 1)[:-1]]).astype(np.int64)
        ci_off = np.concatenate([[0], np.cumsum(nnz)[:-1]]).astype(np.int64)
        row_ptr = np.concatenate([_i32(rp) for rp, _ in pairs])
        col_idx = np.concatenate([_i32(ci) for _, ci in pairs] + [np.zeros(0, np.int32)])
        return BatchSet(n, nnz, rp_off, ci_off, row_ptr, col_idx)


def symmetrize(m: CSR) -> CSR:
    t = CsrT()
    src = _as_csr_t(m)
    r = _native.lib().rcm_symmetrize(C.byref(src), C.byref(t))
    if r < 0:
        raise RcmError(RCM_EINVAL, "symmetrize")
    return _take_csr(t)


def connect_components(m: CSR):
    """Connectivity repair of prepareRCM (Graph.h:160-263). Returns (csr, edges_added)."""
    t = CsrT()
    src = _as_csr_t(m)
    r = _native.lib().rcm_connect_components(C.byref(src), C.byref(t))
    if r < 0:
        raise RcmError(RCM_EINVAL, "connect_components (the reference repair loop would not terminate)")
    return _take_csr(t), int(r)


# ---------------------------------------------------------------------------------------------------------
# File formats (include/rcm_formats.h): .hicsr and MatrixMarket, COO -> CSR, prepareRCM with values
# ---------------------------------------------------------------------------------------------------------
def _value_bytes(dtype) -> int:
    dt = np.dtype(dtype)
    if dt == np.float32:
        return 4
    if dt == np.float64:
        return 8
    raise ValueError("dtype must be float32 (CSR<float>) or float64 (CSR<double>)")


def _path(p) -> bytes:
    import os

    return os.fsencode(p)


def _take_matrix(t: MatrixT, dtype=np.float64) -> CSR:
    rows, cols, nnz = t.rows, t.cols, t.nnz
    try:
        rp = np.ctypeslib.as_array(t.row_ptr, shape=(rows + 1,)).copy()
        if nnz:
            ci = np.ctypeslib.as_array(t.col_idx, shape=(nnz,)).copy()
            data = np.ctypeslib.as_array(t.values, shape=(nnz,)).astype(dtype) if t.values else None
        else:
            ci = np.zeros(0, np.int32)
            data = np.zeros(0, dtype) if t.values else None
    finally:
        _native.lib().rcm_matrix_free(C.byref(t))
    return CSR(rp, ci, data, cols=int(cols))


def _as_matrix_t(m: CSR):
    """MatrixT view of a CSR; returns (struct, keep-alive tuple)."""
    vals = None if m.data is None else np.ascontiguousarray(m.data, dtype=np.float64)
    t = MatrixT(
        m.rows,
        m.cols,
        m.nnz,
        _p32(m.row_offsets),
        _p32(m.col_ids),
        vals.ctypes.data_as(C.POINTER(C.c_double)) if vals is not None else None,
    )
    return t, (m.row_offsets, m.col_ids, vals)


def loadMTX(path, dtype=np.float32):
    """loadMTX<T> + convert (src/COO.cpp:88-199, src/CSR.cpp:170-208). Returns (csr, info) where info has
    the bits MTX_PATTERN, MTX_COMPLEX, MTX_SYMMETRIC, MTX_HERMITIAN."""
    out = MatrixT()
    info = C.c_uint32()
    msg = C.create_string_buffer(512)
    rc = _native.lib().rcm_mtx_load(_path(path), _value_bytes(dtype), C.byref(out), C.byref(info), msg, len(msg))
    if rc != RCM_OK:
        raise RcmError(rc, msg.value.decode(errors="replace") or "loadMTX")
    return _take_matrix(out, dtype), int(info.value)


def storeMTX(m: CSR, path, dtype=np.float32) -> None:
    """convert(COO, CSR) + storeMTX<T> (src/CSR.cpp:210-231, src/COO.cpp:213-251)."""
    t, keep = _as_matrix_t(m)
    msg = C.create_string_buffer(512)
    rc = _native.lib().rcm_mtx_store(_path(path), C.byref(t), _value_bytes(dtype), msg, len(msg))
    del keep
    if rc != RCM_OK:
        raise RcmError(rc, msg.value.decode(errors="replace") or "storeMTX")


def loadCSR(path, dtype=np.float32) -> CSR:
    """loadCSR<T> (src/CSR.cpp:119-151): the .hicsr binary CSR."""
    out = MatrixT()
    msg = C.create_string_buffer(512)
    rc = _native.lib().rcm_hicsr_load(_path(path), _value_bytes(dtype), C.byref(out), msg, len(msg))
    if rc != RCM_OK:
        raise RcmError(rc, msg.value.decode(errors="replace") or "loadCSR")
    return _take_matrix(out, dtype)


def storeCSR(m: CSR, path, dtype=np.float32) -> None:
    """storeCSR<T> (src/CSR.cpp:153-168)."""
    t, keep = _as_matrix_t(m)
    msg = C.create_string_buffer(512)
    rc = _native.lib().rcm_hicsr_store(_path(path), C.byref(t), _value_bytes(dtype), msg, len(msg))
    del keep
    if rc != RCM_OK:
        raise RcmError(rc, msg.value.decode(errors="replace") or "storeCSR")


def convert_coo(rows: int, cols: int, row_ids, col_ids, values=None) -> CSR:
    """convert(CSR, COO) (src/CSR.cpp:170-208): stable sort by (row, column), duplicates kept."""
    ri = _i32(row_ids)
    ci = _i32(col_ids)
    if ri.shape != ci.shape:
        raise ValueError("row_ids and col_ids differ in length")
    vals = None if values is None else np.ascontiguousarray(values, dtype=np.float64)
    out = MatrixT()
    rc = _native.lib().rcm_coo_to_csr(
        rows, cols, ri.size, _p32(ri), _p32(ci),
        vals.ctypes.data_as(C.POINTER(C.c_double)) if vals is not None else None, C.byref(out),
    )
    _check(rc, "convert")
    return _take_matrix(out)


def prepare_rcm(m: CSR, check_sym: bool = True):
    """prepareRCM on a CSR with values (Graph.h:74-264). Returns (csr, links_added)."""
    t, keep = _as_matrix_t(m)
    out = MatrixT()
    r = _native.lib().rcm_prepare(C.byref(t), 1 if check_sym else 0, C.byref(out))
    del keep
    if r < 0:
        raise RcmError(int(r), "prepareRCM")
    dtype = m.data.dtype if m.data is not None else np.float64
    return _take_matrix(out, dtype), int(r)


# ---------------------------------------------------------------------------------------------------------
# GPU canonicalisation and permutation (include/rcm_matrix_ops.h)
# ---------------------------------------------------------------------------------------------------------
class MatrixOps:
    """An rcm_ops handle: COO -> CSR + prepareRCM, and B = P A P^T with values, on the GPU."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(_native.lib().rcm_ops_create(device, C.byref(self._h)), "rcm_ops_create")

    def close(self) -> None:
        if self._h:
            _native.lib().rcm_ops_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def last_ms(self) -> float:
        ms = C.c_double()
        _check(_native.lib().rcm_ops_last_ms(self._h, C.byref(ms)), "rcm_ops_last_ms")
        return ms.value

    def permute(self, m: CSR, perm) -> CSR:
        """B = P A P^T: row i of B is row perm[i] of A, columns renamed to their positions in perm (sorted,
        duplicates in their order in A). The device side of CuSolverRCM::reorder (rcm_cusolver.cuh:154-185)."""
        perm = _i32(perm)
        if perm.shape[0] != m.rows:
            raise ValueError("perm length differs from the row count")
        vals = None if m.data is None else np.ascontiguousarray(m.data, np.float64)
        rp2 = np.empty(m.rows + 1, np.int32)
        ci2 = np.empty(max(m.nnz, 1), np.int32)
        v2 = None if vals is None else np.empty(max(m.nnz, 1), np.float64)
        dp = C.POINTER(C.c_double)
        rc = _native.lib().rcm_ops_permute(
            self._h, m.rows, m.nnz, _p32(m.row_offsets), _p32(m.col_ids),
            vals.ctypes.data_as(dp) if vals is not None else None, _p32(perm), _p32(rp2), _p32(ci2),
            v2.ctypes.data_as(dp) if v2 is not None else None,
        )
        _check(rc, "rcm_ops_permute")
        data = None if v2 is None else v2[: m.nnz].astype(m.data.dtype)
        return CSR(rp2, ci2[: m.nnz], data, cols=m.rows)

    def canonicalize(self, rows: int, cols: int, row_ids, col_ids, values=None, check_sym: bool = True):
        """convert + prepareRCM(check_sym) on the GPU. Returns (csr, links_added)."""
        ri, ci = _i32(row_ids), _i32(col_ids)
        if ri.shape != ci.shape:
            raise ValueError("row_ids and col_ids differ in length")
        vals = None if values is None else np.ascontiguousarray(values, np.float64)
        out = MatrixT()
        links = C.c_int64()
        rc = _native.lib().rcm_ops_canonicalize(
            self._h, rows, cols, ri.size, _p32(ri), _p32(ci),
            vals.ctypes.data_as(C.POINTER(C.c_double)) if vals is not None else None, 1 if check_sym else 0,
            C.byref(out), C.byref(links),
        )
        _check(rc, "rcm_ops_canonicalize")
        return _take_matrix(out), int(links.value)


# ---------------------------------------------------------------------------------------------------------
# Device-resident context
# ---------------------------------------------------------------------------------------------------------
@dataclass
class Timing:
    run_kernel_ms: float
    init_kernel_ms: float
    start_search_ms: float
    bandwidth_ms: float
    load_ms: float
    workers: int
    batches: int
    start_iterations: int
    placed: int
    error_site: int = 0
    error_info: tuple = ()
    order_ms: float = 0.0
    order_passes: int = 0


class RcmContext:
    """Owns device buffers for one matrix at a time (rcm_ctx in include/rcm_gpu_batch_abi.h)."""

    def __init__(self, device: int = 0, **params):
        self._lib = _native.lib()
        h = C.c_void_p()
        _check(self._lib.rcm_ctx_create(device, C.byref(h)), "rcm_ctx_create")
        self._h = h
        self.n = 0
        for k, v in params.items():
            self.set_param(k, v)

    _PARAMS = {"workers": PARAM_WORKERS, "timeout_ms": PARAM_TIMEOUT_MS, "sort": PARAM_SORT,
               "digit_bits": PARAM_DIGIT_BITS, "variant": PARAM_VARIANT, "strict": PARAM_STRICT,
               "profile": PARAM_PROFILE, "claim_filter": PARAM_CLAIM_FILTER, "batch_cap": PARAM_BATCH_CAP,
               "speculate": PARAM_SPECULATE}

    def set_param(self, name: str, value: int) -> None:
        _check(self._lib.rcm_ctx_set_param(self._h, self._PARAMS[name], int(value)), f"set_param {name}")

    def get_param(self, name: str) -> int:
        v = C.c_int64()
        _check(self._lib.rcm_ctx_get_param(self._h, self._PARAMS[name], C.byref(v)), f"get_param {name}")
        return int(v.value)

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.rcm_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, m: CSR) -> None:
        self._rp, self._ci = m.row_offsets, m.col_ids  # keep host arrays alive during the copy
        _check(self._lib.rcm_ctx_load(self._h, m.rows, m.nnz, _p32(m.row_offsets), _p32(m.col_ids)), "load")
        self.n = m.rows

    def autotune(self, start: int, repeats: int = 3) -> TuneResultT:
        """Time every engine variant x workgroup count on the loaded matrix and keep the fastest."""
        r = TuneResultT()
        _check(self._lib.rcm_ctx_autotune(self._h, int(start), repeats, C.byref(r)), "rcm_ctx_autotune")
        return r

    def find_start(self) -> int:
        s, it = C.c_int32(), C.c_int32()
        _check(self._lib.rcm_ctx_find_start(self._h, C.byref(s), C.byref(it)), "find_start")
        return int(s.value)

    def run(self, start: int, flags: int = RCM_FLAG_STABLE) -> None:
        rc = self._lib.rcm_ctx_run(self._h, int(start), flags)
        if rc != RCM_OK:
            t = self.timing()
            raise RcmError(rc, f"run (device diagnostics: site {t.error_site}, {t.error_info})")

    def run_status(self, start: int, flags: int = RCM_FLAG_STABLE) -> int:
        return int(self._lib.rcm_ctx_run(self._h, int(start), flags))

    def order(self, start: int = -1, flags: int = RCM_FLAG_STABLE):
        """Start search (when start < 0) and reorder in one call; returns (start used, sweeps)."""
        s, it = C.c_int32(-1), C.c_int32(0)
        rc = self._lib.rcm_ctx_order(self._h, int(start), flags, C.byref(s), C.byref(it))
        if rc != RCM_OK:
            t = self.timing()
            raise RcmError(rc, f"order (device diagnostics: site {t.error_site}, {t.error_info})")
        return int(s.value), int(it.value)

    def bandwidth(self) -> int:
        bw = C.c_int64()
        _check(self._lib.rcm_ctx_bandwidth(self._h, C.byref(bw)), "bandwidth")
        return int(bw.value)

    def bandwidth_of(self, perm) -> int:
        """Bandwidth of the loaded matrix under an external ordering perm[new] = old (validated)."""
        p = _i32(perm)
        bw = C.c_int64()
        _check(self._lib.rcm_ctx_bandwidth_of(self._h, _p32(p), C.byref(bw)), "bandwidth_of")
        return int(bw.value)

    def verify(self, start: int, perm=None, flags: int = 0) -> int:
        """0 when perm (host; None: the last result) is exactly the stable CM order of the loaded matrix from
        start (its reverse with RCM_FLAG_REVERSE), else 1 + a position where it departs. Checked on the GPU."""
        res = C.c_int64()
        p = None if perm is None else _i32(perm)
        if p is not None and p.shape[0] != self.n:
            return 1
        _check(self._lib.rcm_ctx_verify(self._h, int(start), None if p is None else _p32(p), flags, C.byref(res)),
               "verify")
        return int(res.value)

    def perm(self) -> np.ndarray:
        out = np.empty(self.n, np.int32)
        _check(self._lib.rcm_ctx_read_perm(self._h, _p32(out)), "read_perm")
        return out

    def profile(self) -> dict:
        """Per-phase profile of the last instrumented run: microseconds summed over batches per phase,
        and counts (batches, early, children, slots)."""
        buf = (C.c_uint64 * 16)()
        _check(self._lib.rcm_ctx_profile(self._h, buf, 16), "profile")
        out = {ph: buf[i] * 0.01 for i, ph in enumerate(PROFILE_PHASES)}  # 100 MHz ticks -> us
        out.update(batches=int(buf[8]), early=int(buf[9]), children=int(buf[10]), slots=int(buf[11]))
        out["compact"] = buf[12] * 0.01  # compaction, before "sort" (the ordering itself)
        out["kept"] = int(buf[13])  # batches processed by the workgroup that produced them
        return out

    def timing(self) -> Timing:
        t = TimingT()
        _check(self._lib.rcm_ctx_timing(self._h, C.byref(t)), "timing")
        return Timing(t.run_kernel_ms, t.init_kernel_ms, t.start_search_ms, t.bandwidth_ms, t.load_ms, t.workers,
                      t.batches, t.start_iterations, t.placed, t.error_site, tuple(t.error_info), t.order_ms,
                      t.order_passes)


def rcm_gpu_batch(m: CSR, start_node: int = -1, flags: int = RCM_FLAG_STABLE):
    """One-shot call of the C ABI. Returns (perm, bandwidth or None, start_used)."""
    perm = np.empty(m.rows, np.int32)
    bw = C.c_int64(-1)
    st = C.c_int32(-1)
    _check(
        _native.lib().rcm_gpu_batch(m.rows, m.nnz, _p32(m.row_offsets), _p32(m.col_ids), int(start_node), flags,
                                    _p32(perm), C.byref(bw), C.byref(st)),
        "rcm_gpu_batch",
    )
    return perm, (int(bw.value) if flags & RCM_FLAG_BANDWIDTH else None), int(st.value)


def rcm_gpu_batch_many(bs: BatchSet, starts: Optional[Sequence[int]] = None, flags: int = RCM_FLAG_STABLE):
    """Reorder every matrix of a BatchSet. Returns (perm_cat, per-matrix offsets, bandwidths, statuses)."""
    k = len(bs)
    po = bs.perm_offsets
    perm = np.empty(int(bs.n.astype(np.int64).sum()), np.int32)
    bw = np.full(k, -1, np.int64)
    st = np.zeros(k, np.int32)
    s = _i32(starts if starts is not None else np.full(k, -1))
    rc = _native.lib().rcm_gpu_batch_many(
        k, _p32(bs.n), _p64(bs.nnz), _p64(bs.rp_off), _p64(bs.ci_off), _p32(bs.row_ptr), _p32(bs.col_idx),
        _p32(s), flags, _p64(po), _p32(perm), _p64(bw), _p32(st))
    return perm, po, bw, st, int(rc)


class RcmBatch:
    """Device-resident set of matrices reordered by one persistent launch (rcm_batch in
    include/rcm_gpu_batch_abi.h). Results are bit-identical to reordering each matrix alone."""

    def __init__(self, device: int = 0, timeout_ms: int = 60000, workers: int = 0, graph: bool = True):
        h = C.c_void_p()
        _check(_native.lib().rcm_batch_create(device, C.byref(h)), "rcm_batch_create")
        self._h = h
        self.set_param("timeout_ms", timeout_ms)
        self.set_param("workers", workers)
        self.set_param("graph", 1 if graph else 0)
        self._bs: Optional[BatchSet] = None

    _PARAMS = {"workers": PARAM_WORKERS, "timeout_ms": PARAM_TIMEOUT_MS, "sort": PARAM_SORT, "graph": PARAM_GRAPH,
               "claim_filter": PARAM_CLAIM_FILTER, "variant": PARAM_VARIANT, "wg_max_n": PARAM_WG_MAX_N,
               "helper_nodes": PARAM_HELPER_NODES, "speculate": PARAM_SPECULATE}

    def autotune(self, starts, repeats: int = 1) -> TuneResultT:
        """Time every engine variant x workgroup count on the loaded set and keep the fastest."""
        r = TuneResultT()
        _check(_native.lib().rcm_batch_autotune(self._h, _p32(_i32(starts)), repeats, C.byref(r)), "rcm_batch_autotune")
        return r

    def set_param(self, name: str, value: int) -> None:
        _check(_native.lib().rcm_batch_set_param(self._h, self._PARAMS[name], int(value)), f"set_param {name}")

    def load(self, bs: BatchSet) -> np.ndarray:
        """Upload and validate; returns per-matrix status (raises if any matrix is malformed)."""
        st = np.zeros(len(bs), np.int32)
        rc = _native.lib().rcm_batch_load(self._h, len(bs), _p32(bs.n), _p64(bs.nnz), _p64(bs.rp_off),
                                          _p64(bs.ci_off), _p32(bs.row_ptr), _p32(bs.col_idx), _p32(st))
        _check(rc, "rcm_batch_load")
        self._bs = bs
        return st

    def find_start(self):
        """Returns (starts, iterations, statuses)."""
        k = len(self._bs)
        s, it, st = np.zeros(k, np.int32), np.zeros(k, np.int32), np.zeros(k, np.int32)
        rc = _native.lib().rcm_batch_find_start(self._h, _p32(s), _p32(it), _p32(st))
        if rc in (RCM_EDEVICE, RCM_ETIMEOUT, RCM_ENOTLOADED, RCM_EINVAL):
            raise RcmError(rc, "rcm_batch_find_start")
        return s, it, st

    def run(self, starts, flags: int = RCM_FLAG_STABLE) -> np.ndarray:
        """Reorder every matrix; returns per-matrix status (device-level failures raise)."""
        st = np.zeros(len(self._bs), np.int32)
        rc = _native.lib().rcm_batch_run(self._h, _p32(_i32(starts)), flags, _p32(st))
        if rc in (RCM_EDEVICE, RCM_ETIMEOUT, RCM_ENOTLOADED, RCM_EINVAL):
            raise RcmError(rc, "rcm_batch_run")
        return st

    def order(self, starts=None, flags: int = RCM_FLAG_STABLE):
        """Start search and reorder of every matrix in one call (starts[k] >= 0 skips the search for matrix k).
        Returns (starts used, sweeps, per-matrix status); device-level failures raise."""
        k = len(self._bs)
        s, it, st = np.zeros(k, np.int32), np.zeros(k, np.int32), np.zeros(k, np.int32)
        sarr = None if starts is None else _i32(starts)
        rc = _native.lib().rcm_batch_order(self._h, None if sarr is None else _p32(sarr), flags, _p32(s), _p32(it),
                                           _p32(st))
        if rc in (RCM_EDEVICE, RCM_ETIMEOUT, RCM_ENOTLOADED, RCM_EINVAL):
            raise RcmError(rc, "rcm_batch_order")
        return s, it, st

    def run_status(self, starts, flags: int = RCM_FLAG_STABLE):
        """Reorder every matrix; returns (return code, per-matrix status) without raising."""
        st = np.zeros(len(self._bs), np.int32)
        rc = _native.lib().rcm_batch_run(self._h, _p32(_i32(starts)), flags, _p32(st))
        return int(rc), st

    def bandwidth(self) -> np.ndarray:
        bw = np.zeros(len(self._bs), np.int64)
        _check(_native.lib().rcm_batch_bandwidth(self._h, _p64(bw)), "rcm_batch_bandwidth")
        return bw

    def perms(self) -> list:
        bs = self._bs
        po = bs.perm_offsets
        cat = np.empty(int(bs.n.astype(np.int64).sum()), np.int32)
        _check(_native.lib().rcm_batch_read_perm(self._h, _p64(po), _p32(cat)), "rcm_batch_read_perm")
        return [cat[po[k]: po[k] + int(bs.n[k])] for k in range(len(bs))]

    def timing(self) -> BatchTimingT:
        t = BatchTimingT()
        _check(_native.lib().rcm_batch_timing(self._h, C.byref(t)), "rcm_batch_timing")
        return t

    def close(self) -> None:
        if self._h:
            _native.lib().rcm_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 10, device: int = 0,
                   runtime: bool = False) -> tuple[float, float]:
    """(GB/s, ms per copy) of a plain device-to-device copy of nbytes (rcm_selftest_copy_bandwidth): the HBM
    bandwidth this GPU reaches on streaming reads and writes; the library's copy kernel, or with runtime the HIP
    runtime's own device-to-device copy."""
    gbs, ms = C.c_double(), C.c_double()
    _check(_native.lib().rcm_selftest_copy_bandwidth(device, int(nbytes), int(reps), 1 if runtime else 0,
                                                     C.byref(gbs), C.byref(ms)), "rcm_selftest_copy_bandwidth")
    return gbs.value, ms.value


class RcmServer:
    """Continuous batching of reorder requests on one GPU (rcm_server in include/rcm_gpu_batch_abi.h): lanes of
    many-matrix contexts run side by side; a free lane takes every waiting request that fits its HBM arena and
    reorders them in one persistent launch, so requests that arrive while others run overlap with them on the
    device. Results are bit-identical to rcm_gpu_batch."""

    def __init__(self, device: int = 0, slots: int = 256, lanes: int = 4, arena_bytes: int = 0, workers: int = 0,
                 job_timeout_ms: int = 60000):
        cfg = ServerConfigT(slots, lanes, arena_bytes, workers, job_timeout_ms)
        h = C.c_void_p()
        _check(_native.lib().rcm_server_create(device, C.byref(cfg), C.byref(h)), "rcm_server_create")
        self._h = h
        self._n = {}
        self._keep = {}  # borrowed arrays of requests not yet waited for

    def submit(self, m: CSR, start: int = -1, flags: int = RCM_FLAG_STABLE, borrow: bool = False) -> int:
        """Queue m. With borrow the server reads m's arrays in place (RCM_FLAG_BORROW): no copy at submit; this
        object keeps them alive until the request's wait returns, and they must not be changed before then."""
        jid = C.c_int64()
        rp, ci = m.row_offsets, m.col_ids
        if borrow:
            flags |= RCM_FLAG_BORROW
        _check(_native.lib().rcm_server_submit(self._h, m.rows, m.nnz, _p32(rp), _p32(ci), int(start), flags,
                                               C.byref(jid)), "rcm_server_submit")
        self._n[jid.value] = m.rows
        if flags & RCM_FLAG_BORROW:
            self._keep[jid.value] = (rp, ci)
        return jid.value

    def wait(self, jid: int, timeout_ms: int = -1):
        """Returns (status, perm or None, JobStatsT); raises RcmError(RCM_EAGAIN) on time-out."""
        n = self._n[jid]
        perm = np.empty(n, np.int32)
        st = JobStatsT()
        rc = _native.lib().rcm_server_wait(self._h, jid, int(timeout_ms), _p32(perm), C.byref(st))
        if rc == RCM_EAGAIN:
            raise RcmError(rc, "rcm_server_wait")
        del self._n[jid]
        self._keep.pop(jid, None)
        return rc, (perm if rc == RCM_OK else None), st

    def stats(self) -> ServerStatsT:
        st = ServerStatsT()
        _check(_native.lib().rcm_server_stats_get(self._h, C.byref(st)), "rcm_server_stats_get")
        return st

    def close(self) -> None:
        if self._h:
            _native.lib().rcm_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------------------------------------
# Reference-shaped entry points
# ---------------------------------------------------------------------------------------------------------
class GPU_BATCH:
    """Mirror of namespace GPU_BATCH (include/rcm_gpu_batch.h:37-45)."""

    warmups = 20  # src/device/rcm_gpu_batch.cu:47-48
    runs = 20

    @classmethod
    def rcm(cls, graph: Graph, csv_file, permutation: list, start_node: int) -> None:
        """GPU_BATCH::rcm (src/device/rcm_gpu_batch.cu:42-112): CM order (not reversed) into `permutation`
        (resized to n), mean kernel milliseconds of `runs` timed launches after `warmups` written to csv_file."""
        with RcmContext() as ctx:
            ctx.load(graph.csr)
            total = 0.0
            for r in range(cls.warmups + cls.runs):
                ctx.run(start_node, RCM_FLAG_STABLE)
                if r >= cls.warmups:
                    total += ctx.timing().run_kernel_ms
            mean = total / max(cls.runs, 1)
            print(f"{'GPU-DYN-RCM':>27} duration: {mean} ms")
            if csv_file is not None:
                csv_file.write(f"{mean}")
                csv_file.flush()
            permutation[:] = ctx.perm().tolist()

    @staticmethod
    def findPeripheral(graph: Graph, csv_file=None, writeOutput: bool = False) -> int:
        return SimplePeripheral.findPeripheral(graph, csv_file, writeOutput)


class SimplePeripheral:
    """Mirror of namespace SimplePeripheral (include/simple_peripheral.h:39-47), computed on the GPU."""

    @staticmethod
    def findPeripheral(graph: Graph, csv_file=None, writeOutput: bool = False, outputnodes: int = 32) -> int:
        if outputnodes != 32:
            raise ValueError("the GPU start search implements outputnodes = 32 (the reference default)")
        with RcmContext() as ctx:
            ctx.load(graph.csr)
            s = ctx.find_start()
            if writeOutput and csv_file is not None:
                csv_file.write(f",{ctx.timing().start_search_ms}")
                csv_file.flush()
            return s


def computeBandwidth(graph: Graph, permutation) -> int:
    """computeBandwidth (src/main.cpp:143-184) of an arbitrary full ordering, as a GPU reduction.
    Returns -1 for an empty permutation (as the reference)."""
    p = _i32(permutation)
    if p.size == 0:
        return -1
    with RcmContext() as ctx:
        ctx.load(graph.csr)
        return ctx.bandwidth_of(p)


def checkPermutation(permutation) -> int:
    """checkPermutation (src/main.cpp:186-226) as a strict bijection check: 0 if valid, -1 otherwise."""
    p = np.asarray(permutation, dtype=np.int64)
    n = p.shape[0]
    if n == 0:
        return 0
    if p.min() < 0 or p.max() >= n:
        return -1
    seen = np.zeros(n, np.int8)
    seen[p] = 1
    return 0 if int(seen.sum()) == n else -1


__all__ = [
    "CSR", "Graph", "BatchSet", "workloads", "symmetrize", "connect_components", "RcmContext", "Timing",
    "rcm_gpu_batch", "rcm_gpu_batch_many", "GPU_BATCH", "SimplePeripheral", "computeBandwidth",
    "checkPermutation", "RcmError",
]
