"""Device buffers for GPU tests of the device-pointer entry points, allocated through the HIP
runtime that libmeshopt_hip.so itself links (so the test process holds one runtime only)."""
import ctypes

import numpy as np

import meshoptimizer_amd as M

_H2D, _D2H = 1, 2
_hip = None


def hip():
    global _hip
    if _hip is None:
        M.library()  # loads the library and, with it, the runtime it was linked against
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
    return _hip


class DeviceArray:
    """A device copy of a numpy array (or an uninitialised buffer of `nbytes`); freed on close."""

    def __init__(self, host=None, nbytes=None):
        self.nbytes = host.nbytes if host is not None else int(nbytes)
        p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(p), max(self.nbytes, 4)) == 0
        self.ptr = p.value
        if host is not None and self.nbytes:
            host = np.ascontiguousarray(host)
            assert hip().hipMemcpy(self.ptr, host.ctypes.data, self.nbytes, _H2D) == 0

    def to_host(self, dtype, count):
        out = np.empty(count, dtype=dtype)
        if count:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, _D2H) == 0
        return out

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
