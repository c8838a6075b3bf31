"""Device buffers for tests that call device-pointer entry points of the C ABI: allocated, filled and read
through the HIP runtime the engine library itself loaded (one runtime per process). Test infrastructure."""
from __future__ import annotations

import ctypes as C

import numpy as np

from parallelbatchrcm_amd import _native


class HipBuffers:
    def __init__(self):
        _native.lib()
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.restype = C.c_int
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.restype = C.c_int
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.restype = C.c_int
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L = L
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(int(nbytes), 1)) == 0
        self.bufs.append(p)
        return p

    def upload(self, a):
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # host to device
        return p

    def download(self, p, count, dtype):
        a = np.empty(count, dtype)
        assert self.L.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, a.nbytes, 2) == 0  # device to host
        return a

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
