"""Minimal device buffers for the GPU tests, through the same HIP runtime the library links.

(torch ships its own HIP runtime; two runtimes cannot share the device in one process, so the tests
do not use torch for device memory.)
"""
import ctypes as C

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        from vebisph_amd import _lib
        _lib.lib()                                  # loads libamdhip64 as a dependency
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return _hip


class DeviceArray:
    """A float64 device buffer with host copies in and out."""

    def __init__(self, host):
        self.host = np.ascontiguousarray(host, dtype=np.float64)
        self.ptr = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.ptr), max(self.host.nbytes, 8)) == 0
        assert hip().hipMemcpy(self.ptr, self.host.ctypes.data, self.host.nbytes, 1) == 0   # H2D

    def numpy(self):
        out = np.empty_like(self.host)
        assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0             # D2H
        return out

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = None
