"""Device-resident blocks allocated through libhelmholtz_amd (no torch).

A :class:`DeviceArray` is a column-major N x ncol block in HBM; the solver and
operator accept it in place of a numpy array and then nothing crosses PCIe.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


class DeviceArray:
    def __init__(self, shape, dtype=np.complex128):
        self.shape = tuple(int(v) for v in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        L.check(L.lib().hh_device_alloc(self.nbytes, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.asfortranarray(a)
        d = cls(a.shape, a.dtype)
        L.check(L.lib().hh_memcpy(d.ptr, a.ctypes.data, d.nbytes, 0))
        return d

    def to_numpy(self):
        a = np.empty(self.shape, dtype=self.dtype, order="F")
        L.check(L.lib().hh_memcpy(a.ctypes.data, self.ptr, self.nbytes, 1))
        return a

    def empty_like(self):
        return DeviceArray(self.shape, self.dtype)

    def copy_from(self, other: "DeviceArray"):
        assert other.nbytes == self.nbytes
        L.check(L.lib().hh_memcpy(self.ptr, other.ptr, self.nbytes, 2))

    def __del__(self):
        p = getattr(self, "ptr", None)
        if p:
            try:
                L.lib().hh_device_free(p)
            except Exception:
                pass
            self.ptr = None
