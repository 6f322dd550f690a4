"""Minimal device-memory helper over the HIP runtime that libvoronoid_hip.so itself links, so
inputs can be placed in HBM before a timed region without loading a second HIP runtime.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import lib

_hip = None


def hip():
    global _hip
    if _hip is None:
        lib()  # loads libamdhip64.so.7 as a dependency; the soname below resolves to that copy
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMalloc.restype = ctypes.c_int
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipFree.restype = ctypes.c_int
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpy.restype = ctypes.c_int
        h.hipDeviceSynchronize.restype = ctypes.c_int
        h.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
        h.hipGetDeviceCount.restype = ctypes.c_int
        h.hipMemGetInfo.argtypes = [ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_size_t)]
        h.hipMemGetInfo.restype = ctypes.c_int
        _hip = h
    return _hip


def device_count() -> int:
    n = ctypes.c_int(0)
    if hip().hipGetDeviceCount(ctypes.byref(n)) != 0:
        return 0
    return n.value


def mem_free() -> int:
    """Free device memory of the current device in bytes (hipMemGetInfo)."""
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    if hip().hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) != 0:
        raise RuntimeError("hipMemGetInfo failed")
    return int(free.value)


def synchronize() -> None:
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


class DeviceArray:
    """A float64 array copied to device memory (hipMalloc), freed on close()/deletion."""

    def __init__(self, host: np.ndarray):
        a = np.ascontiguousarray(host, dtype=np.float64)
        self.nbytes = a.nbytes
        self.shape = a.shape
        self.ptr = ctypes.c_void_p()
        if hip().hipMalloc(ctypes.byref(self.ptr), max(self.nbytes, 16)) != 0:
            raise MemoryError("hipMalloc failed")
        if self.nbytes and hip().hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, 1) != 0:
            raise RuntimeError("hipMemcpy H2D failed")

    @classmethod
    def empty(cls, nbytes: int) -> "DeviceArray":
        """Uninitialised device memory of nbytes bytes."""
        return cls(np.empty((max(int(nbytes), 8) + 7) // 8, dtype=np.float64))

    def download(self, dtype, count: int) -> np.ndarray:
        """The first count elements of dtype, copied to the host."""
        out = np.empty(int(count), dtype=dtype)
        if out.nbytes and hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, out.nbytes, 2) != 0:
            raise RuntimeError("hipMemcpy D2H failed")
        return out

    def close(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
