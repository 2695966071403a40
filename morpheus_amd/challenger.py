"""Device Fiat-Shamir challengers (goldilocks-cuda-rs/src/challenger.rs: ChallengerBatch and
DuplexChallenger over the challenger_* entry points): n independent Poseidon2-w8 duplex sponges
in device memory, every call applying one operation to all of them."""
import ctypes
import struct

import numpy as np

from ._lib import lib
from .device import DeviceBuffer, check

STATE_BYTES = 136


class ChallengerBatch:
    """ChallengerBatch::new(count) (challenger.rs:55-69)."""

    def __init__(self, count):
        self.count = int(count)
        p = ctypes.c_void_p()
        check(lib.challenger_alloc_states(ctypes.byref(p), self.count), "challenger_alloc_states")
        self.ptr = p.value
        check(lib.challenger_init_states(self.ptr, self.count), "challenger_init_states")

    def _values(self, values, per):
        v = np.ascontiguousarray(values, dtype=np.uint64).ravel()
        if v.size != self.count * per:
            raise ValueError(f"values length must be count * {per}")
        return DeviceBuffer.from_numpy(v)

    def observe(self, values):
        d = self._values(values, 1)
        check(lib.challenger_observe_ffi(self.ptr, d.ptr, self.count), "challenger_observe_ffi")

    def observe_slice(self, values, slice_len):
        d = self._values(values, slice_len)
        check(lib.challenger_observe_slice_ffi(self.ptr, d.ptr, slice_len, self.count), "challenger_observe_slice_ffi")

    def observe_ext2(self, values):
        d = self._values(values, 2)
        check(lib.challenger_observe_ext2_ffi(self.ptr, d.ptr, self.count), "challenger_observe_ext2_ffi")

    def _sample(self, fn, per, *extra):
        d = DeviceBuffer(self.count * per)
        check(fn(self.ptr, d.ptr, *extra, self.count), fn.__name__)
        return d.to_numpy().reshape(self.count, per) if per > 1 else d.to_numpy()

    def sample(self):
        return self._sample(lib.challenger_sample_ffi, 1)

    def sample_array(self, per):
        return self._sample(lib.challenger_sample_array_ffi, per, per)

    def sample_ext2(self):
        return self._sample(lib.challenger_sample_ext2_ffi, 2)

    def sample_ext5(self):
        return self._sample(lib.challenger_sample_ext5_ffi, 5)

    def state(self, i):
        """(sponge[8], in[4], out[4], in_count, out_count) of challenger i."""
        buf = ctypes.create_string_buffer(STATE_BYTES)
        check(lib.challenger_copy_state_to_host(buf, self.ptr + STATE_BYTES * i), "challenger_copy_state_to_host")
        w = struct.unpack("<16Q2i", buf.raw)
        return np.array(w[:8], dtype=np.uint64), np.array(w[8:12], dtype=np.uint64), \
            np.array(w[12:16], dtype=np.uint64), w[16], w[17]

    def raw_state(self, i):
        buf = ctypes.create_string_buffer(STATE_BYTES)
        check(lib.challenger_copy_state_to_host(buf, self.ptr + STATE_BYTES * i), "challenger_copy_state_to_host")
        return buf.raw

    def load_state(self, i, raw):
        assert len(raw) == STATE_BYTES
        check(lib.challenger_copy_state_to_device(self.ptr + STATE_BYTES * i, raw), "challenger_copy_state_to_device")

    def free(self):
        if self.ptr:
            lib.cuda_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DuplexChallenger(ChallengerBatch):
    """Single challenger (challenger.rs:239-280)."""

    def __init__(self):
        super().__init__(1)

    def observe(self, value):
        super().observe([value])

    def observe_slice(self, values):
        values = list(values)
        super().observe_slice(values, len(values))

    def observe_ext2(self, value2):
        super().observe_ext2(value2)

    def sample(self):
        return int(super().sample()[0])

    def sample_array(self, count):
        return super().sample_array(count).ravel() if count > 1 else super().sample_array(count)

    def sample_ext2(self):
        return super().sample_ext2()[0]
