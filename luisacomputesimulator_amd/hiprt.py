"""Device buffers and synchronisation through the HIP runtime that liblcsx.so links.

The standalone LBVH entry points take device pointers; this module gives Python callers
plain device allocations and copies without another GPU framework in the process.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

_HIP: Optional[C.CDLL] = None


def _hip() -> C.CDLL:
    global _HIP
    if _HIP is None:
        from . import load_library

        load_library()  # resolves the same libamdhip64 the solver uses
        lib = C.CDLL("libamdhip64.so.7")
        lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        lib.hipFree.argtypes = [C.c_void_p]
        lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        lib.hipDeviceSynchronize.argtypes = []
        lib.hipGetErrorString.restype = C.c_char_p
        lib.hipGetErrorString.argtypes = [C.c_int]
        _HIP = lib
    return _HIP


def _ok(code: int, what: str) -> None:
    if code != 0:
        raise RuntimeError(f"{what} failed: {_hip().hipGetErrorString(code).decode()}")


def copy_device(dst: "DeviceArray", src: "DeviceArray") -> None:
    """Device-to-device copy of src into dst (same size)."""
    if dst.nbytes != src.nbytes:
        raise ValueError("copy_device: sizes differ")
    _ok(_hip().hipMemcpy(dst.ptr, src.ptr, src.nbytes, 3), "hipMemcpy")


def synchronize() -> None:
    _ok(_hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device copy of a numpy array (freed with the object)."""

    H2D, D2H = 1, 2

    def __init__(self, host: np.ndarray):
        self.host = np.ascontiguousarray(host)
        self.nbytes = self.host.nbytes
        self.ptr = C.c_void_p()
        _ok(_hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 4)), "hipMalloc")
        if self.nbytes:
            _ok(_hip().hipMemcpy(self.ptr, self.host.ctypes.data_as(C.c_void_p), self.nbytes, self.H2D), "hipMemcpy")

    def numpy(self) -> np.ndarray:
        out = np.empty_like(self.host)
        if self.nbytes:
            _ok(_hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, self.D2H), "hipMemcpy")
        return out

    def __del__(self):
        try:
            if self.ptr:
                _hip().hipFree(self.ptr)
        except Exception:
            pass
