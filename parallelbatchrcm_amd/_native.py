"""ctypes bindings of the engine library (include/rcm_gpu_batch_abi.h, include/rcm_workloads.h,
include/rcm_formats.h, include/rcm_matrix_ops.h).

The library is built in-tree by parallelbatchrcm_amd/build.py. Loading fails loudly when it is missing:
there is no CPU fallback behind this package.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpbrcm.so")

I32P = C.POINTER(C.c_int32)
I64P = C.POINTER(C.c_int64)
U64P = C.POINTER(C.c_uint64)


class CsrT(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int64), ("row_ptr", I32P), ("col_idx", I32P)]


class MatrixT(C.Structure):
    _fields_ = [
        ("rows", C.c_int32),
        ("cols", C.c_int32),
        ("nnz", C.c_int64),
        ("row_ptr", I32P),
        ("col_idx", I32P),
        ("values", C.POINTER(C.c_double)),
    ]


class TimingT(C.Structure):
    _fields_ = [
        ("run_kernel_ms", C.c_double),
        ("init_kernel_ms", C.c_double),
        ("start_search_ms", C.c_double),
        ("bandwidth_ms", C.c_double),
        ("load_ms", C.c_double),
        ("workers", C.c_int32),
        ("batches", C.c_int32),
        ("start_iterations", C.c_int32),
        ("placed", C.c_int32),
        ("error_site", C.c_int32),
        ("error_info", C.c_int32 * 8),
        ("order_ms", C.c_double),
        ("order_passes", C.c_int32),
    ]


class BatchTimingT(C.Structure):
    _fields_ = [
        ("load_ms", C.c_double),
        ("start_search_ms", C.c_double),
        ("run_ms", C.c_double),
        ("bandwidth_ms", C.c_double),
        ("count", C.c_int32),
        ("workers", C.c_int32),
        ("start_sweeps", C.c_int32),
        ("graph_replays", C.c_int32),
        ("batches", C.c_int64),
        ("total_n", C.c_int64),
        ("total_nnz", C.c_int64),
        ("order_ms", C.c_double),
        ("order_passes", C.c_int32),
    ]


class ServerConfigT(C.Structure):
    _fields_ = [
        ("slots", C.c_int32),
        ("lanes", C.c_int32),
        ("arena_bytes", C.c_int64),
        ("workers", C.c_int32),
        ("job_timeout_ms", C.c_int64),
    ]


class JobStatsT(C.Structure):
    _fields_ = [
        ("status", C.c_int32),
        ("start_node", C.c_int32),
        ("sweeps", C.c_int32),
        ("lane", C.c_int32),
        ("queue_ms", C.c_double),
        ("device_ms", C.c_double),
        ("latency_ms", C.c_double),
        ("bandwidth", C.c_int64),
    ]


class ServerStatsT(C.Structure):
    _fields_ = [
        ("submitted", C.c_int64),
        ("completed", C.c_int64),
        ("failed", C.c_int64),
        ("inflight", C.c_int32),
        ("pending", C.c_int32),
        ("max_inflight", C.c_int32),
        ("slots", C.c_int32),
        ("arena_bytes", C.c_int64),
        ("arena_peak_bytes", C.c_int64),
        ("workers", C.c_int32),
        ("kernel_running", C.c_int32),
        ("admitted", C.c_int64),
        ("loop_iterations", C.c_int64),
        ("probe_rounds", C.c_int64),
        ("claims", C.c_int64),
        ("group_host_copy_ms", C.c_double),
        ("group_load_ms", C.c_double),
        ("group_order_ms", C.c_double),
        ("group_readback_ms", C.c_double),
        ("group_deliver_ms", C.c_double),
        ("group_launches", C.c_int64),
        ("pinned_lanes", C.c_int32),
        ("solo_lanes", C.c_int32),
    ]


class TuneResultT(C.Structure):
    _fields_ = [
        ("variant", C.c_int32),
        ("workers", C.c_int32),
        ("candidates", C.c_int32),
        ("threads", C.c_int32),
        ("best_ms", C.c_double),
        ("default_ms", C.c_double),
        ("helper_nodes", C.c_int32),
        ("reserved", C.c_int32),
    ]


# name -> (restype, argtypes)
_SIGNATURES = {
    # rcm_gpu_batch_abi.h
    "rcm_gpu_batch": (C.c_int, [C.c_int32, C.c_int64, I32P, I32P, C.c_int32, C.c_uint32, I32P, I64P, I32P]),
    "rcm_ctx_verify": (C.c_int, [C.c_void_p, C.c_int32, I32P, C.c_uint32, I64P]),
    "rcm_ctx_trace": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_int64]),
    "rcm_gpu_batch_noalloc": (C.c_int, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32,
                                        C.c_void_p, I64P, I32P]),
    "rcm_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "rcm_ctx_destroy": (C.c_int, [C.c_void_p]),
    "rcm_ctx_set_param": (C.c_int, [C.c_void_p, C.c_int, C.c_int64]),
    "rcm_ctx_get_param": (C.c_int, [C.c_void_p, C.c_int, I64P]),
    "rcm_ctx_load": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, I32P, I32P]),
    "rcm_ctx_load_device": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p]),
    "rcm_ctx_find_start": (C.c_int, [C.c_void_p, I32P, I32P]),
    "rcm_ctx_run": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32]),
    "rcm_ctx_order": (C.c_int, [C.c_void_p, C.c_int32, C.c_uint32, I32P, I32P]),
    "rcm_ctx_bandwidth": (C.c_int, [C.c_void_p, I64P]),
    "rcm_ctx_bandwidth_of": (C.c_int, [C.c_void_p, I32P, I64P]),
    "rcm_ctx_read_perm": (C.c_int, [C.c_void_p, I32P]),
    "rcm_ctx_perm_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "rcm_ctx_timing": (C.c_int, [C.c_void_p, C.POINTER(TimingT)]),
    "rcm_ctx_profile": (C.c_int, [C.c_void_p, U64P, C.c_int32]),
    "rcm_batch_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "rcm_batch_destroy": (C.c_int, [C.c_void_p]),
    "rcm_batch_set_param": (C.c_int, [C.c_void_p, C.c_int, C.c_int64]),
    "rcm_batch_load": (C.c_int, [C.c_void_p, C.c_int32, I32P, I64P, I64P, I64P, I32P, I32P, I32P]),
    "rcm_batch_find_start": (C.c_int, [C.c_void_p, I32P, I32P, I32P]),
    "rcm_batch_run": (C.c_int, [C.c_void_p, I32P, C.c_uint32, I32P]),
    "rcm_batch_order": (C.c_int, [C.c_void_p, I32P, C.c_uint32, I32P, I32P, I32P]),
    "rcm_batch_bandwidth": (C.c_int, [C.c_void_p, I64P]),
    "rcm_batch_read_perm": (C.c_int, [C.c_void_p, I64P, I32P]),
    "rcm_batch_perm_device": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "rcm_batch_timing": (C.c_int, [C.c_void_p, C.POINTER(BatchTimingT)]),
    "rcm_server_create": (C.c_int, [C.c_int, C.POINTER(ServerConfigT), C.POINTER(C.c_void_p)]),
    "rcm_server_submit": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, I32P, I32P, C.c_int32, C.c_uint32, I64P]),
    "rcm_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, I32P, C.POINTER(JobStatsT)]),
    "rcm_server_stats_get": (C.c_int, [C.c_void_p, C.POINTER(ServerStatsT)]),
    "rcm_server_destroy": (C.c_int, [C.c_void_p]),
    "rcm_ctx_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(TuneResultT)]),
    "rcm_batch_autotune": (C.c_int, [C.c_void_p, I32P, C.c_int32, C.POINTER(TuneResultT)]),
    "rcm_gpu_batch_many": (
        C.c_int,
        [C.c_int32, I32P, I64P, I64P, I64P, I32P, I32P, I32P, C.c_uint32, I64P, I32P, I64P, I32P],
    ),
    "rcm_status_string": (C.c_char_p, [C.c_int]),
    "rcm_selftest_primitives": (C.c_int, [C.c_int, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "rcm_selftest_copy_bandwidth": (C.c_int, [C.c_int, C.c_int64, C.c_int32, C.c_int32, C.POINTER(C.c_double),
                                              C.POINTER(C.c_double)]),
    # rcm_workloads.h
    "rcm_csr_free": (None, [C.POINTER(CsrT)]),
    "rcm_gen_stencil2d": (C.c_int32, [C.c_int32, C.c_int32, C.POINTER(CsrT)]),
    "rcm_gen_stencil3d": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.POINTER(CsrT)]),
    "rcm_gen_rgg": (C.c_int32, [C.c_int32, C.c_double, C.c_uint64, C.POINTER(CsrT)]),
    "rcm_gen_rmat": (
        C.c_int32,
        [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_double, C.c_uint64, C.POINTER(CsrT)],
    ),
    "rcm_permute_symmetric": (C.c_int32, [C.POINTER(CsrT), C.c_uint64, C.POINTER(CsrT), I32P]),
    "rcm_connect_components": (C.c_int64, [C.POINTER(CsrT), C.POINTER(CsrT)]),
    "rcm_symmetrize": (C.c_int64, [C.POINTER(CsrT), C.POINTER(CsrT)]),
    "rcm_gen_config": (C.c_int32, [C.c_int32, C.c_int64, C.c_uint64, C.POINTER(CsrT)]),
    "rcm_gen_batch_mix": (
        C.c_int32,
        [
            C.c_int32, C.c_int32, C.c_int32, C.c_uint64, C.c_int32,
            C.POINTER(I32P), C.POINTER(I64P), C.POINTER(I64P), C.POINTER(I64P), C.POINTER(I32P), C.POINTER(I32P),
        ],
    ),
    "rcm_free": (None, [C.c_void_p]),
    # rcm_formats.h
    "rcm_matrix_free": (None, [C.POINTER(MatrixT)]),
    "rcm_mtx_load": (C.c_int32, [C.c_char_p, C.c_int32, C.POINTER(MatrixT), C.POINTER(C.c_uint32), C.c_char_p, C.c_size_t]),
    "rcm_mtx_store": (C.c_int32, [C.c_char_p, C.POINTER(MatrixT), C.c_int32, C.c_char_p, C.c_size_t]),
    "rcm_hicsr_load": (C.c_int32, [C.c_char_p, C.c_int32, C.POINTER(MatrixT), C.c_char_p, C.c_size_t]),
    "rcm_hicsr_store": (C.c_int32, [C.c_char_p, C.POINTER(MatrixT), C.c_int32, C.c_char_p, C.c_size_t]),
    "rcm_coo_to_csr": (
        C.c_int32,
        [C.c_int32, C.c_int32, C.c_int64, I32P, I32P, C.POINTER(C.c_double), C.POINTER(MatrixT)],
    ),
    "rcm_prepare": (C.c_int64, [C.POINTER(MatrixT), C.c_int32, C.POINTER(MatrixT)]),
    # rcm_matrix_ops.h
    "rcm_ops_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "rcm_ops_destroy": (C.c_int, [C.c_void_p]),
    "rcm_ops_last_ms": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "rcm_ops_permute_device": (
        C.c_int,
        [C.c_void_p, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
         C.c_void_p, C.c_void_p],
    ),
    "rcm_ops_permute": (
        C.c_int,
        [C.c_void_p, C.c_int32, C.c_int64, I32P, I32P, C.POINTER(C.c_double), I32P, I32P, I32P,
         C.POINTER(C.c_double)],
    ),
    "rcm_ops_canonicalize_device": (
        C.c_int,
        [C.c_void_p, C.c_int32, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
         I64P],
    ),
    "rcm_ops_result": (
        C.c_int,
        [C.c_void_p, I32P, I64P, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)],
    ),
    "rcm_ops_canonicalize": (
        C.c_int,
        [C.c_void_p, C.c_int32, C.c_int32, C.c_int64, I32P, I32P, C.POINTER(C.c_double), C.c_int32,
         C.POINTER(MatrixT), I64P],
    ),
    "rcm_graph_load": (
        C.c_int32,
        [C.c_char_p, C.c_int32, C.c_int32, C.POINTER(MatrixT), I32P, I64P, C.c_char_p, C.c_size_t],
    ),
    "rcm_splitmix64": (C.c_uint64, [U64P]),
}

_lib = None


def lib() -> C.CDLL:
    """The loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"engine library {LIB_PATH} is missing; build it with "
                "`python -m parallelbatchrcm_amd.build` (there is no CPU fallback)"
            )
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def exported_symbols() -> list[str]:
    return list(_SIGNATURES)
