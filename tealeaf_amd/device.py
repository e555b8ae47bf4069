"""Device buffers for the suite, backed by the library's own allocator."""
from __future__ import annotations

import ctypes

import numpy as np

from .abi import lib


class DeviceArray:
    """A zero-filled device buffer of `n` elements of `dtype` (float64 default)."""

    def __init__(self, n: int, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        out = ctypes.c_void_p()
        lib().tlhip_alloc(ctypes.c_size_t(self.nbytes), ctypes.addressof(out))
        self.ptr = int(out.value)

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype).ravel()
        if a.size != self.n:
            raise ValueError(f"size mismatch: {a.size} != {self.n}")
        lib().tlhip_copy_to_device(self.ptr, a.ctypes.data, ctypes.c_size_t(self.nbytes))

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        lib().tlhip_copy_to_host(out.ctypes.data, self.ptr, ctypes.c_size_t(self.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            lib().tlhip_free(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __int__(self) -> int:
        return self.ptr
