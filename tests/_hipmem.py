"""Device buffers for the tests of the entry points on caller-provided device memory:
hipMalloc / hipMemcpy / hipFree of the HIP runtime the engine library links (libamdhip64.so.7,
so the buffers belong to the same runtime as the engine's calls), through ctypes."""
import ctypes as C

import numpy as np

import dqmc_amd as D

_hip = None
H2D, D2H = 1, 2


def hip():
    global _hip
    if _hip is None:
        D.lib()   # loads the engine and, with it, its HIP runtime
        h = C.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DevBuf:
    """A device copy of a numpy array in its memory order (column-major matrices stay so)."""

    def __init__(self, a=None, nbytes=None, dtype=np.float64):
        self.p = C.c_void_p()
        if a is not None:
            a = np.asarray(a)
            self.dtype = a.dtype
            flat = np.ascontiguousarray(a.T if a.ndim == 2 else a).reshape(-1)
            nbytes = flat.nbytes
        else:
            self.dtype = np.dtype(dtype)
        self.nbytes = int(nbytes)
        assert hip().hipMalloc(C.byref(self.p), max(self.nbytes, 8)) == 0
        if a is not None:
            assert hip().hipMemcpy(self.p, flat.ctypes.data_as(C.c_void_p), self.nbytes, H2D) == 0

    @property
    def ptr(self):
        return self.p.value

    def get(self, shape=None):
        hip().hipDeviceSynchronize()
        out = np.zeros(self.nbytes // self.dtype.itemsize, self.dtype)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, self.nbytes, D2H) == 0
        if shape is None:
            return out
        return out.reshape(shape[1], shape[0]).T.copy()   # column-major matrix

    def __del__(self):
        if getattr(self, "p", None) is not None and self.p.value:
            hip().hipFree(self.p)
            self.p = None


def synchronize():
    assert hip().hipDeviceSynchronize() == 0
