"""int8-matrix Ajtai commitment on INT8 MFMA: out = M z (mod p), M (rows x n, int8)
expanded from a 32-byte key with ChaCha8 (see include/morpheus_prover.h)."""
import ctypes

import numpy as np

from ._lib import lib
from .device import DeviceBuffer, check


def key_from_seed(seed):
    """32-byte key from a 64-bit seed: four SplitMix64 outputs, little-endian."""
    out = b""
    x = seed & (2**64 - 1)
    for _ in range(4):
        x = (x + 0x9E3779B97F4A7C15) & (2**64 - 1)
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2**64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2**64 - 1)
        z ^= z >> 31
        out += z.to_bytes(8, "little")
    return out


class AjtaiI8:
    def __init__(self, key32, rows, n):
        self.key, self.rows, self.n = bytes(key32), rows, n
        h = ctypes.c_void_p()
        check(lib.mam_ajtai_i8_setup(self.key, rows, n, ctypes.byref(h)), "mam_ajtai_i8_setup")
        self.h = h.value

    def commit_device(self, d_z, batch, d_out):
        check(lib.mam_ajtai_i8_commit(self.h, d_z.ptr, batch, d_out.ptr), "mam_ajtai_i8_commit")

    def commit(self, z, batch=1):
        z = np.ascontiguousarray(z, dtype=np.uint64)
        assert z.size == batch * self.n
        dz, do = DeviceBuffer.from_numpy(z), DeviceBuffer(batch * self.rows)
        self.commit_device(dz, batch, do)
        return do.to_numpy().reshape(batch, self.rows)

    def commit_prefix(self, z, n_used, batch=1):
        """Commit batch witnesses of n_used words against the first n_used columns."""
        z = np.ascontiguousarray(z, dtype=np.uint64)
        assert z.size == batch * n_used
        dz, do = DeviceBuffer.from_numpy(z), DeviceBuffer(batch * self.rows)
        check(lib.mam_ajtai_i8_commit_prefix(self.h, dz.ptr, batch, n_used, do.ptr), "mam_ajtai_i8_commit_prefix")
        return do.to_numpy().reshape(batch, self.rows)

    def commit_noalloc_device(self, d_z, batch, n_used, d_out):
        """Queue the allocation-free commit (mam_ajtai_i8_commit_noalloc) of batch witnesses of
        n_used words on the calling thread's stream; d_out is complete after a synchronize."""
        rc = lib.mam_ajtai_i8_commit_noalloc(self.h, d_z.ptr, batch, n_used, d_out.ptr)
        if rc == -2:
            raise ValueError("mam_ajtai_i8_commit_noalloc: invalid arguments, or the chunk partials "
                             f"of batch {batch} x {n_used} columns exceed its 16 MiB scratch")
        check(rc, "mam_ajtai_i8_commit_noalloc")

    def matrix_row(self, row, col0, count):
        out = np.zeros(count, dtype=np.int8)
        check(lib.mam_ajtai_i8_matrix_row(self.h, row, col0, count, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))),
              "mam_ajtai_i8_matrix_row")
        return out

    def free(self):
        if self.h:
            lib.mam_ajtai_i8_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
