"""Device arrays for ctypes callers of the library, over the HIP runtime the
library itself links (libamdhip64.so.7, already loaded with it).

    dev = DeviceArray.from_numpy(np_array)      # hipMalloc + copy
    lib.some_entry(..., dev.ptr, ...)
    host = dev.numpy()                          # synchronous copy back
    dev.free()

Only what the trainer and the tests need: allocation, host<->device
copies and a device-wide synchronize. Copies use the null stream, which
waits for the library's non-blocking streams only through an explicit
synchronize (DeviceArray.numpy() calls hipDeviceSynchronize first).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

_H2D, _D2H, _D2D = 1, 2, 3
_rt = None


def runtime():
    global _rt
    if _rt is None:
        from oceanlib_amd import capi
        capi.load()  # the library's own dependency resolves the runtime
        rt = C.CDLL("libamdhip64.so.7")
        rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        rt.hipFree.argtypes = [C.c_void_p]
        rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        rt.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        for f in (rt.hipMalloc, rt.hipFree, rt.hipMemcpy, rt.hipMemset, rt.hipDeviceSynchronize):
            f.restype = C.c_int
        _rt = rt
    return _rt


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


def synchronize():
    _check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    def __init__(self, shape, dtype):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _check(runtime().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self._p = p

    @property
    def ptr(self) -> C.c_void_p:
        return self._p

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        d.upload(a)
        return d

    @classmethod
    def zeros(cls, shape, dtype) -> "DeviceArray":
        d = cls(shape, dtype)
        _check(runtime().hipMemset(d._p, 0, d.nbytes), "hipMemset")
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        _check(runtime().hipMemcpy(self._p, a.ctypes.data_as(C.c_void_p), self.nbytes, _H2D),
               "hipMemcpy H2D")

    def numpy(self) -> np.ndarray:
        synchronize()
        out = np.empty(self.shape, self.dtype)
        _check(runtime().hipMemcpy(out.ctypes.data_as(C.c_void_p), self._p, self.nbytes, _D2H),
               "hipMemcpy D2H")
        return out

    def offset(self, nbytes: int) -> C.c_void_p:
        return C.c_void_p(self._p.value + int(nbytes))

    def free(self):
        if self._p is not None and self._p.value:
            runtime().hipFree(self._p)
        self._p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
