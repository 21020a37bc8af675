"""Minimal device buffers on the HIP runtime libbaorec_hip.so itself uses.

Plays the role AMDGPU.jl's ROCArray would play for a Julia caller: the device
pointers handed to the C ABI with BAOREC_MEM_DEVICE.  It binds the already
loaded libamdhip64.so.7 (the library's own dependency) so that no second HIP
runtime enters the process.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_hip = None


def hip():
    global _hip
    if _hip is None:
        _lib.load()  # makes libamdhip64.so.7 resident
        h = ctypes.CDLL("libamdhip64.so.7")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMalloc.restype = ctypes.c_int
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipFree.restype = ctypes.c_int
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpy.restype = ctypes.c_int
        h.hipDeviceSynchronize.restype = ctypes.c_int
        h.hipSetDevice.argtypes = [ctypes.c_int]
        h.hipSetDevice.restype = ctypes.c_int
        _hip = h
    return _hip


H2D, D2H, D2D = 1, 2, 3


class DeviceArray:
    """A float32 (or other dtype) device allocation with a numpy-like shape."""

    is_cuda = True

    def __init__(self, shape, dtype=np.float32, device: int = 0):
        self.shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.device = device
        h = hip()
        h.hipSetDevice(device)
        p = ctypes.c_void_p()
        rc = h.hipMalloc(ctypes.byref(p), max(self.nbytes, 4))
        if rc != 0:
            raise MemoryError(f"hipMalloc({self.nbytes}) failed: {rc}")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray, device: int = 0) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype, device)
        rc = hip().hipMemcpy(ctypes.c_void_p(d.ptr), a.ctypes.data_as(ctypes.c_void_p), a.nbytes, H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed: {rc}")
        return d

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        rc = hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr), self.nbytes, D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed: {rc}")
        return out

    def data_ptr(self) -> int:
        return self.ptr

    def is_contiguous(self) -> bool:
        return True

    def __len__(self):
        return self.shape[0]

    def free(self):
        if self.ptr:
            hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
