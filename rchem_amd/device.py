"""Device buffers for the device-pointer entry points (rchem_jk_device and its
no-allocation variant), allocated through the HIP runtime librchem_gpu.so
itself links, so a process that never imports torch still has one runtime.
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import lib

_hip = None
H2D, D2H = 1, 2


def hip():
    """The HIP runtime the library is linked against (loaded with it)."""
    global _hip
    if _hip is None:
        lib()  # loads libamdhip64 as the library's dependency
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        _hip = h
    return _hip


class DeviceMatrix:
    """An n x n float64 matrix in device memory."""

    def __init__(self, n: int):
        self.n = n
        self.nbytes = 8 * n * n
        self.ptr = ctypes.c_void_p()
        rc = hip().hipMalloc(ctypes.byref(self.ptr), self.nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc of {self.nbytes} bytes failed ({rc})")

    def upload(self, a: np.ndarray) -> None:
        c = np.ascontiguousarray(a, dtype=np.float64)
        if c.shape != (self.n, self.n):
            raise ValueError(f"expected ({self.n}, {self.n})")
        rc = hip().hipMemcpy(self.ptr, c.ctypes.data, self.nbytes, H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy to device failed ({rc})")

    def download(self) -> np.ndarray:
        out = np.empty((self.n, self.n))
        synchronize()
        rc = hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy from device failed ({rc})")
        return out

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize() -> None:
    rc = hip().hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")
