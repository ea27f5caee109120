"""Device arrays over the library's own allocator (voxel_gpu_malloc / free,
memcpy_h2d / d2h): what the Python front ends use for HBM buffers, so that
no third-party runtime computes or allocates on the product path.

A DeviceArray is a flat, contiguous run of `n` elements of a numpy dtype.
`view(k)` is the first k elements without a copy (it keeps its owner
alive).  A caller's device buffer (for example a framework tensor) can be
wrapped in place with `DeviceArray.borrow(ptr, n, dtype, owner)`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


class DeviceArray:
    __slots__ = ("ptr", "n", "dtype", "_own", "_owner")

    def __init__(self, n: int, dtype):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        nbytes = max(self.n, 1) * self.dtype.itemsize
        p = lib.voxel_gpu_malloc(nbytes)
        if not p:
            check("voxel_gpu_malloc")
            raise MemoryError("voxel_gpu_malloc(%d) failed" % nbytes)
        self.ptr = int(p)
        self._own = True
        self._owner = None

    @classmethod
    def borrow(cls, ptr: int, n: int, dtype, owner=None) -> "DeviceArray":
        d = cls.__new__(cls)
        d.ptr, d.n, d.dtype, d._own, d._owner = int(ptr), int(n), np.dtype(dtype), False, owner
        return d

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size and lib.voxel_gpu_memcpy_h2d(C.c_void_p(d.ptr), a.ctypes.data, a.nbytes) != 0:
            check("voxel_gpu_memcpy_h2d")
        return d

    @classmethod
    def zeros(cls, n: int, dtype) -> "DeviceArray":
        return cls.from_numpy(np.zeros(n, dtype))

    def view(self, k: int) -> "DeviceArray":
        if k > self.n:
            raise ValueError("view of %d elements of %d" % (k, self.n))
        return DeviceArray.borrow(self.ptr, k, self.dtype, owner=self)

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.n and lib.voxel_gpu_memcpy_d2h(out.ctypes.data, C.c_void_p(self.ptr), out.nbytes) != 0:
            check("voxel_gpu_memcpy_d2h")
        return out

    # the interface the scan helpers read (voxel_amd.scan)
    def data_ptr(self) -> int:
        return self.ptr

    def numel(self) -> int:
        return self.n

    def __len__(self) -> int:
        return self.n

    def __del__(self, _lib=lib):          # the library outlives module teardown
        if getattr(self, "_own", False) and self.ptr:
            _lib.voxel_gpu_free(C.c_void_p(self.ptr))
            self.ptr = 0
