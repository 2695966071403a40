"""Device memory through the library's own allocator entry points (cuda_malloc & co).

Mirrors ``DeviceBuffer`` of goldilocks-cuda-rs/src/memory.rs: an owned device allocation
of u64 words with host upload/download.
"""
import ctypes

import numpy as np

from ._lib import lib


class GpuError(RuntimeError):
    """A library entry point returned a nonzero status (the reference's CudaError)."""


def check(rc, what):
    if rc != 0:
        raise GpuError(f"{what} failed with status {rc}")


class DeviceBuffer:
    """Owned device allocation of ``n`` u64 words."""

    def __init__(self, n):
        self.n = int(n)
        p = ctypes.c_void_p()
        check(lib.cuda_malloc(ctypes.byref(p), max(self.n, 1) * 8), "cuda_malloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a, dtype=np.uint64).ravel()
        b = cls(a.size)
        if a.size:
            check(lib.cuda_memcpy_htod(b.ptr, a.ctypes.data, a.size * 8), "cuda_memcpy_htod")
        return b

    def to_numpy(self, n=None, offset=0):
        n = self.n - offset if n is None else int(n)
        out = np.empty(n, dtype=np.uint64)
        if n:
            check(lib.cuda_memcpy_dtoh(out.ctypes.data, self.ptr + 8 * offset, n * 8), "cuda_memcpy_dtoh")
        return out

    def free(self):
        if self.ptr:
            lib.cuda_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class PinnedArray:
    """Page-locked host array of ``n`` words of ``dtype`` (mam_host_alloc): device-to-host
    copies into it skip the staging copy that pageable memory takes. ``.a`` is the numpy
    view; the memory lives as long as this object."""

    def __init__(self, n, dtype=np.uint64):
        dt = np.dtype(dtype)
        self.n = int(n)
        p = ctypes.c_void_p()
        check(lib.mam_host_alloc(max(self.n, 1) * dt.itemsize, ctypes.byref(p)), "mam_host_alloc")
        self.ptr = p.value
        buf = (ctypes.c_char * (max(self.n, 1) * dt.itemsize)).from_address(self.ptr)
        self.a = np.frombuffer(buf, dtype=dt, count=self.n)
        self.a[:] = 0

    def free(self):
        if self.ptr:
            self.a = None
            lib.mam_host_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize():
    check(lib.cuda_device_synchronize(), "cuda_device_synchronize")


def device_count():
    c = ctypes.c_int(0)
    if lib.cuda_get_device_count(ctypes.byref(c)) != 0:
        return 0
    return c.value
