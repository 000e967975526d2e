"""Device buffers for the device-pointer entry points, through
libbddevmem.so (same HIP runtime as libbdcudatag.so)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbddevmem.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.bdm_alloc.argtypes = [ctypes.c_size_t]
        L.bdm_alloc.restype = ctypes.c_void_p
        L.bdm_free.argtypes = [ctypes.c_void_p]
        L.bdm_upload.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.bdm_download.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        L.bdm_copy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        for f in (L.bdm_free, L.bdm_upload, L.bdm_download, L.bdm_copy, L.bdm_sync, L.bdm_device_count):
            f.restype = ctypes.c_int32
        _lib = L
    return _lib


def copy(dst: int, src: int, nbytes: int) -> None:
    """Device-to-device copy of nbytes."""
    if lib().bdm_copy(dst, src, nbytes) != 0:
        raise RuntimeError("bdm_copy failed")


def sync():
    if lib().bdm_sync() != 0:
        raise RuntimeError("hipDeviceSynchronize failed")


class DeviceBuffer:
    """A device copy of a numpy array (contiguous bytes)."""

    def __init__(self, host: np.ndarray):
        host = np.ascontiguousarray(host)
        self.nbytes = host.nbytes
        self.shape = host.shape
        self.dtype = host.dtype
        self.ptr = lib().bdm_alloc(self.nbytes)
        if not self.ptr:
            raise MemoryError("hipMalloc failed")
        if lib().bdm_upload(self.ptr, host.ctypes.data, self.nbytes) != 0:
            raise RuntimeError("hipMemcpy to device failed")

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        if lib().bdm_download(out.ctypes.data, self.ptr, self.nbytes) != 0:
            raise RuntimeError("hipMemcpy to host failed")
        return out

    def close(self):
        if getattr(self, "ptr", None):
            lib().bdm_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()
