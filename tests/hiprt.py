"""Device buffers for tests that call entry points on device arrays: ctypes over the HIP
runtime that libsmac_amd.so itself links (so its pointers are valid in the library)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from minismac_amd import solver

_hip = None


def hip():
    global _hip
    if _hip is None:
        solver.lib()                       # loads libamdhip64.so.7 as the library's dependency
        H = C.CDLL("libamdhip64.so.7")
        H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        H.hipFree.argtypes = [C.c_void_p]
        H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        H.hipDeviceSynchronize.argtypes = []
        for f in (H.hipMalloc, H.hipFree, H.hipMemcpy, H.hipDeviceSynchronize):
            f.restype = C.c_int
        _hip = H
    return _hip


H2D, D2H = 1, 2


class DeviceArray:
    """A device copy of a numpy array (bytes in the array's memory order)."""

    def __init__(self, host: np.ndarray):
        self.shape, self.dtype = host.shape, host.dtype
        self.order = "F" if host.flags.f_contiguous and not host.flags.c_contiguous else "C"
        self.nbytes = host.nbytes
        p = C.c_void_p()
        rc = hip().hipMalloc(C.byref(p), max(self.nbytes, 16))
        if rc:
            raise RuntimeError(f"hipMalloc failed ({rc})")
        self.ptr = p.value
        self.upload(host)

    def upload(self, host: np.ndarray) -> None:
        h = np.require(host, dtype=self.dtype, requirements=[self.order])
        rc = hip().hipMemcpy(self.ptr, h.ctypes.data, self.nbytes, H2D)
        if rc:
            raise RuntimeError(f"hipMemcpy H2D failed ({rc})")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype, order=self.order)
        rc = hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H)
        if rc:
            raise RuntimeError(f"hipMemcpy D2H failed ({rc})")
        return out

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = 0

    __del__ = free


def synchronize() -> None:
    rc = hip().hipDeviceSynchronize()
    if rc:
        raise RuntimeError(f"hipDeviceSynchronize failed ({rc})")
