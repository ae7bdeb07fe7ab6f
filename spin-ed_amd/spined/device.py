"""Device buffers owned through libspined's own HIP runtime.

The host process uses exactly one HIP runtime -- libspined's.  (PyTorch-ROCm
bundles a second HIP/HSA runtime; mixing the two in one process fails, so the
GPU paths here do not touch torch.)"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .internal import check_status


class DeviceArray:
    def __init__(self, shape, dtype):
        self.dtype = np.dtype(dtype)
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check_status(_lib.load().spined_device_malloc(C.byref(p), self.nbytes))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check_status(_lib.load().spined_memcpy(d.ptr, a.ctypes.data, a.nbytes, 0))
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        d = cls(shape, dtype)
        check_status(_lib.load().spined_memset(d.ptr, 0, d.nbytes))
        return d

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        check_status(_lib.load().spined_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1))
        return out

    def data_ptr(self):
        return self.ptr

    def free(self):
        if self.ptr:
            _lib.load().spined_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_synchronize():
    check_status(_lib.load().spined_device_synchronize())
