"""Ajtai commitment over R = F_q[X]/(X^64 + 1), q = 2^64 - 2^32 - 31, KAPPA = 42 rows.

Mirrors almost-goldilocks-cuda-rs/src/ajtai.rs:31-386: ``Seed``, ``ChunkSize``,
``RingCommitment`` and ``commit_batched`` / ``commit_batched_at`` / ``commit`` /
``commit_wide`` / ``commit_wide_device`` / ``commit_sparse``, with the same argument
meaning and errors (``GpuError`` for the reference's ``CudaError``). The matrix is never
materialized; it is regenerated from the key by ChaCha8 inside the kernels.
"""
import ctypes

import numpy as np

from ._lib import lib
from .device import DeviceBuffer, GpuError, check

RING_DIM = 64
KAPPA = 42
Q = 0xFFFFFFFEFFFFFFE1
CHUNK_SIZES = (64, 128, 256, 1024, 4096)
SUPPORTED_B = (1, 2, 4, 8, 16)


class Seed:
    """256-bit ChaCha8 key as 8 little-endian u32 words (ajtai.rs:52-71)."""

    def __init__(self, words):
        self.words = np.asarray(words, dtype=np.uint32).copy()
        assert self.words.size == 8

    @classmethod
    def from_bytes(cls, b):
        assert len(b) == 32
        return cls(np.frombuffer(bytes(b), dtype="<u4"))

    def to_bytes(self):
        return self.words.astype("<u4").tobytes()


class RingCommitment:
    """rows[i][r] = coefficient r of output row i, canonical in [0, q)."""

    def __init__(self, rows):
        self.rows = np.asarray(rows, dtype=np.uint64).reshape(KAPPA, RING_DIM)

    def __eq__(self, other):
        return isinstance(other, RingCommitment) and np.array_equal(self.rows, other.rows)


def pick_default_chunk(n, b):
    """Smallest supported chunk giving at most the target number of chunks (ajtai.rs:107-119)."""
    target = 1200 if b <= 2 else 200
    needed = (n + target - 1) // target
    for c in CHUNK_SIZES:
        if needed <= c:
            return c
    return 4096


def _key_buffer(seed):
    kw = np.zeros(4, dtype=np.uint64)
    kw.view(np.uint32)[:] = seed.words
    return DeviceBuffer.from_numpy(kw)


def _check(rc, what):
    if rc == -2:
        raise GpuError(f"{what}: batch size is not supported (must be in {SUPPORTED_B})")
    if rc == -3:
        raise GpuError(f"{what}: chunk size is not supported (must be in {CHUNK_SIZES})")
    check(rc, what)


def commit_batched_at(seed, witnesses, col_offset=0, chunk=None):
    """B binary witnesses (each N u64 bitmasks: bit l of word j = coefficient l of z_j)."""
    b = len(witnesses)
    if b not in SUPPORTED_B:
        raise GpuError(f"batch size {b} is not supported (must be in {SUPPORTED_B})")
    n = len(witnesses[0])
    if n == 0:
        raise GpuError("witness length must be > 0")
    if any(len(w) != n for w in witnesses):
        raise GpuError("all witnesses must have the same length")
    chunk = chunk or pick_default_chunk(n, b)
    dz = DeviceBuffer.from_numpy(np.concatenate([np.asarray(w, dtype=np.uint64) for w in witnesses]))
    dk, do = _key_buffer(seed), DeviceBuffer(b * KAPPA * RING_DIM)
    _check(lib.ajtai_commit_dense_batched_at_ffi(dk.ptr, dz.ptr, n, b, chunk, col_offset, do.ptr),
           "ajtai_commit_dense_batched_at_ffi")
    flat = do.to_numpy().reshape(b, KAPPA, RING_DIM)
    return [RingCommitment(flat[i]) for i in range(b)]


def commit_batched(seed, witnesses, chunk=None):
    return commit_batched_at(seed, witnesses, 0, chunk)


def commit(seed, z_bits, chunk=None):
    return commit_batched(seed, [z_bits], chunk)[0]


def commit_wide_device(seed, d_z_wide, col_offset=0, chunk=None):
    """Field-coefficient witness on the device, [n_ring][64] canonical words."""
    if d_z_wide.n == 0 or d_z_wide.n % RING_DIM:
        raise GpuError(f"d_z_wide length {d_z_wide.n} must be a nonzero multiple of {RING_DIM}")
    n_ring = d_z_wide.n // RING_DIM
    chunk = chunk or pick_default_chunk(n_ring, 1)
    dk, do = _key_buffer(seed), DeviceBuffer(KAPPA * RING_DIM)
    _check(lib.ajtai_commit_wide_ffi(dk.ptr, d_z_wide.ptr, n_ring, col_offset, chunk, do.ptr), "ajtai_commit_wide_ffi")
    return RingCommitment(do.to_numpy())


def commit_wide(seed, z_wide, col_offset=0, chunk=None):
    z = np.asarray(z_wide, dtype=np.uint64)
    if z.size == 0 or z.size % RING_DIM:
        raise GpuError(f"z_wide length {z.size} not divisible by {RING_DIM}")
    return commit_wide_device(seed, DeviceBuffer.from_numpy(z), col_offset, chunk)


def commit_sparse(seed, positions, chunk=None):
    """Binary witness given by its set positions p = 64 j + l (single witness)."""
    pos = np.asarray(positions, dtype=np.uint64)
    chunk = chunk or 1024
    dk, do = _key_buffer(seed), DeviceBuffer(KAPPA * RING_DIM)
    dp = DeviceBuffer.from_numpy(pos if pos.size else np.zeros(1, dtype=np.uint64))
    _check(lib.ajtai_commit_sparse_ffi(dk.ptr, dp.ptr, pos.size, chunk, do.ptr), "ajtai_commit_sparse_ffi")
    return RingCommitment(do.to_numpy())


def matrix_entry(seed, row, col):
    out = np.zeros(RING_DIM, dtype=np.uint64)
    check(lib.mam_ajtai_ring_matrix_entry(seed.to_bytes(), row, col,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))),
          "mam_ajtai_ring_matrix_entry")
    return out
