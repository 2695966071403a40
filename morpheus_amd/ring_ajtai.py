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


# ---- folds, splits and ternary commits (ajtai.rs:400-1500) ----

SPLIT_K_CHUNKS = 13
MAX_MULTIFOLD = 511          # the int16 output bound (ajtai.rs:590-594); 64 KB of LDS tables


class RingChallenge:
    """Ring element with 64 int8 coefficients (the reference draws them from {-1, 0, 1, 2})."""

    def __init__(self, coeffs):
        self.coeffs = np.asarray(coeffs, dtype=np.int8).copy()
        if self.coeffs.size != RING_DIM:
            raise GpuError(f"a ring challenge has {RING_DIM} coefficients, got {self.coeffs.size}")

    @classmethod
    def constant_one(cls):
        c = np.zeros(RING_DIM, dtype=np.int8)
        c[0] = 1
        return cls(c)


class TernaryChunks:
    """k_chunks ternary vectors pos_i - neg_i, each [n_ring] u64 bit words, chunk-major."""

    def __init__(self, n_ring, k_chunks, pos, neg):
        self.n_ring, self.k_chunks = int(n_ring), int(k_chunks)
        self.pos = np.asarray(pos, dtype=np.uint64)
        self.neg = np.asarray(neg, dtype=np.uint64)


class TernaryChunksDevice:
    """Device form of :class:`TernaryChunks` (pos, neg are DeviceBuffers of k_chunks * n_ring words)."""

    def __init__(self, n_ring, k_chunks, pos, neg):
        self.n_ring, self.k_chunks, self.pos, self.neg = int(n_ring), int(k_chunks), pos, neg

    @classmethod
    def from_host(cls, chunks):
        return cls(chunks.n_ring, chunks.k_chunks, DeviceBuffer.from_numpy(chunks.pos),
                   DeviceBuffer.from_numpy(chunks.neg))

    def to_host(self):
        return TernaryChunks(self.n_ring, self.k_chunks, self.pos.to_numpy(), self.neg.to_numpy())


def _bytes_buffer(a):
    """Device copy of a small int8 / int16 array (padded to whole u64 words)."""
    raw = np.ascontiguousarray(a).view(np.uint8).ravel()
    words = np.zeros((raw.size + 7) // 8, dtype=np.uint64)
    words.view(np.uint8)[:raw.size] = raw
    return DeviceBuffer.from_numpy(words)


def _i16_from(buf, n):
    return buf.to_numpy().view(np.int16)[:n].copy()


def _challenges_all(challenges, m):
    """[m][64] int8 with the implicit constant-1 challenge of instance 0 first (ajtai.rs:1100-1110)."""
    if m == 0:
        raise GpuError("need at least one instance")
    if len(challenges) + 1 != m:
        raise GpuError(f"expected {m - 1} challenges (instance 0 has implicit weight 1), got {len(challenges)}")
    return np.concatenate([RingChallenge.constant_one().coeffs] + [c.coeffs for c in challenges])


def fold_witness_device(d_z1, r, d_z2, d_dst):
    """d_dst = z1 + r z2 as [n_ring][64] canonical F_q words (ajtai.rs:432-465); z1, z2 binary."""
    n_ring = d_z1.n
    if d_z2.n != n_ring:
        raise GpuError("z1 and z2 must have the same length")
    if d_dst.n != n_ring * RING_DIM:
        raise GpuError(f"dst length {d_dst.n} != N_ring * 64 = {n_ring * RING_DIM}")
    rc = r.coeffs.ctypes.data_as(ctypes.c_void_p)
    _check(lib.ajtai_fold_witness_ffi(d_z1.ptr, d_z2.ptr, rc, n_ring, pick_default_chunk(n_ring, 1), d_dst.ptr),
           "ajtai_fold_witness_ffi")


def fold_witness(z1, r, z2):
    z1, z2 = np.asarray(z1, dtype=np.uint64), np.asarray(z2, dtype=np.uint64)
    if z1.size != z2.size:
        raise GpuError("z1 and z2 must have the same length")
    if z1.size == 0:
        return np.zeros(0, dtype=np.uint64)
    out = DeviceBuffer(z1.size * RING_DIM)
    fold_witness_device(DeviceBuffer.from_numpy(z1), r, DeviceBuffer.from_numpy(z2), out)
    return out.to_numpy()


def fold_commitment(c1, r, c2):
    """c1 + r c2 (ajtai.rs:493-530)."""
    d1, d2 = DeviceBuffer.from_numpy(c1.rows), DeviceBuffer.from_numpy(c2.rows)
    do = DeviceBuffer(KAPPA * RING_DIM)
    _check(lib.ajtai_fold_commitment_ffi(d1.ptr, d2.ptr, r.coeffs.ctypes.data_as(ctypes.c_void_p), do.ptr),
           "ajtai_fold_commitment_ffi")
    return RingCommitment(do.to_numpy())


def _multifold_common(m):
    if m > MAX_MULTIFOLD:
        raise GpuError(f"num_instances {m} exceeds {MAX_MULTIFOLD} per launch")


def multifold_witness(witnesses, challenges):
    """sum_m r_m z_m over binary witnesses, r_0 = 1, as int16 [n_ring * 64] (ajtai.rs:571-660)."""
    m = len(witnesses)
    r_all = _challenges_all(challenges, m)
    _multifold_common(m)
    n_ring = len(witnesses[0])
    if any(len(w) != n_ring for w in witnesses):
        raise GpuError("all witnesses must have the same length")
    if n_ring == 0:
        return np.zeros(0, dtype=np.int16)
    dz = DeviceBuffer.from_numpy(np.concatenate([np.asarray(w, dtype=np.uint64) for w in witnesses]))
    dr, do = _bytes_buffer(r_all), DeviceBuffer((n_ring * RING_DIM + 3) // 4)
    _check(lib.ajtai_multifold_witness_ffi(dz.ptr, dr.ptr, do.ptr, m, n_ring, 1024), "ajtai_multifold_witness_ffi")
    return _i16_from(do, n_ring * RING_DIM)


def _multifold_mixed(fn, binary_witnesses, ternary_chunks, challenges, chunk_arg):
    k_bin, k_tern = len(binary_witnesses), ternary_chunks.k_chunks
    m = k_bin + k_tern
    r_all = _challenges_all(challenges, m)
    _multifold_common(m)
    n_ring = ternary_chunks.n_ring
    if any(len(w) != n_ring for w in binary_witnesses):
        raise GpuError("binary witnesses and ternary chunks must have the same n_ring")
    if n_ring == 0:
        return np.zeros(0, dtype=np.int16)
    dz = DeviceBuffer.from_numpy(np.concatenate([np.asarray(w, dtype=np.uint64) for w in binary_witnesses])
                                 if k_bin else np.zeros(1, dtype=np.uint64))
    dr, do = _bytes_buffer(r_all), DeviceBuffer((n_ring * RING_DIM + 3) // 4)
    args = [dz.ptr if k_bin else None, ternary_chunks.pos.ptr if k_tern else None,
            ternary_chunks.neg.ptr if k_tern else None, dr.ptr, do.ptr, k_bin, k_tern, n_ring] + chunk_arg
    _check(fn(*args), fn.__name__)
    return _i16_from(do, n_ring * RING_DIM)


def multifold_mixed_witness(binary_witnesses, ternary_chunks, challenges):
    """Binary instances first, then the ternary chunks (device), r_0 = 1 (ajtai.rs:679-790)."""
    return _multifold_mixed(lib.ajtai_multifold_mixed_witness_ffi, binary_witnesses, ternary_chunks, challenges,
                            [1024])


def multifold_mixed_witness_tc(binary_witnesses, ternary_chunks, challenges):
    return _multifold_mixed(lib.ajtai_multifold_mixed_witness_tc_ffi, binary_witnesses, ternary_chunks,
                            challenges, [])


def multifold_mixed_witness_tc_fused(binary_witnesses, ternary_chunks, challenges):
    return _multifold_mixed(lib.ajtai_multifold_mixed_witness_tc_fused_ffi, binary_witnesses, ternary_chunks,
                            challenges, [])


def multifold_commitment(commitments, challenges):
    """c_0 + sum_{m >= 1} r_m c_m (ajtai.rs:1078-1130)."""
    m = len(commitments)
    r_all = _challenges_all(challenges, m)
    dc = DeviceBuffer.from_numpy(np.concatenate([c.rows.ravel() for c in commitments]))
    dr, do = _bytes_buffer(r_all), DeviceBuffer(KAPPA * RING_DIM)
    _check(lib.ajtai_multifold_commitment_ffi(dc.ptr, dr.ptr, m, do.ptr), "ajtai_multifold_commitment_ffi")
    return RingCommitment(do.to_numpy())


def split_witness_device(d_z_wide, n_ring):
    """13 ternary chunks of an int16 witness already on the device (packed in u64 words)."""
    if n_ring == 0:
        return TernaryChunksDevice(0, SPLIT_K_CHUNKS, DeviceBuffer(0), DeviceBuffer(0))
    dp, dn = DeviceBuffer(SPLIT_K_CHUNKS * n_ring), DeviceBuffer(SPLIT_K_CHUNKS * n_ring)
    _check(lib.ajtai_split_witness_ffi(d_z_wide.ptr, dp.ptr, dn.ptr, n_ring), "ajtai_split_witness_ffi")
    return TernaryChunksDevice(n_ring, SPLIT_K_CHUNKS, dp, dn)


def split_witness(z_wide):
    """v = sum_i 2^i (pos_i - neg_i) for every int16 coefficient, |v| < 2^13 (ajtai.rs:1261-1300)."""
    z = np.asarray(z_wide, dtype=np.int16)
    if z.size % RING_DIM:
        raise GpuError(f"z_wide length {z.size} not divisible by RING_DIM = {RING_DIM}")
    n_ring = z.size // RING_DIM
    if n_ring == 0:
        return TernaryChunks(0, SPLIT_K_CHUNKS, np.zeros(0, np.uint64), np.zeros(0, np.uint64))
    return split_witness_device(_bytes_buffer(z), n_ring).to_host()


class MaterializedM:
    """M[0..n][0..KAPPA][0..64] generated once from the seed (ajtai.rs:1400-1430)."""

    def __init__(self, seed, n):
        if n == 0:
            raise GpuError("n must be > 0")
        self.n = int(n)
        self.buf = DeviceBuffer(self.n * KAPPA * RING_DIM)
        dk = _key_buffer(seed)
        check(lib.ajtai_materialize_m_ffi(dk.ptr, self.buf.ptr, self.n), "ajtai_materialize_m_ffi")


def _ternary_out(do):
    flat = do.to_numpy().reshape(SPLIT_K_CHUNKS, KAPPA, RING_DIM)
    return [RingCommitment(flat[i]) for i in range(SPLIT_K_CHUNKS)]


def _ternary_check(chunks):
    if chunks.k_chunks != SPLIT_K_CHUNKS:
        raise GpuError(f"expected {SPLIT_K_CHUNKS} ternary chunks, got {chunks.k_chunks}")
    if chunks.n_ring == 0:
        raise GpuError("n_ring must be > 0")


def commit_ternary(seed, chunks, chunk=None):
    """The 13 commits of a split witness under one ChaCha8 matrix (ajtai.rs:1351-1395)."""
    _ternary_check(chunks)
    chunk = chunk or pick_default_chunk(chunks.n_ring, SPLIT_K_CHUNKS)
    dk, do = _key_buffer(seed), DeviceBuffer(SPLIT_K_CHUNKS * KAPPA * RING_DIM)
    _check(lib.ajtai_commit_ternary_ffi(dk.ptr, chunks.pos.ptr, chunks.neg.ptr, chunks.n_ring, chunk, do.ptr),
           "ajtai_commit_ternary_ffi")
    return _ternary_out(do)


def commit_ternary_premat(m, chunks, chunk=None):
    """commit_ternary against a materialized matrix; the same result (ajtai.rs:1441-1480)."""
    _ternary_check(chunks)
    if m.n != chunks.n_ring:
        raise GpuError(f"MaterializedM.n = {m.n} mismatches chunks.n_ring = {chunks.n_ring}")
    chunk = chunk or pick_default_chunk(chunks.n_ring, SPLIT_K_CHUNKS)
    do = DeviceBuffer(SPLIT_K_CHUNKS * KAPPA * RING_DIM)
    _check(lib.ajtai_commit_ternary_premat_ffi(m.buf.ptr, chunks.pos.ptr, chunks.neg.ptr, chunks.n_ring, chunk,
                                               do.ptr), "ajtai_commit_ternary_premat_ffi")
    return _ternary_out(do)
