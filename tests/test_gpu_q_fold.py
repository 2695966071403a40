"""GPU parity of the ring-Ajtai fold, multifold, split and ternary-commit entry points and the
batched selective add (morpheus_agl.h; almost-goldilocks-cuda-rs/src/ffi.rs:283-291, 435-492,
603-647) against the oracle, word for word, plus the homomorphisms the reference's
integration tests check (ajtai_integration.rs:551-582, 779-822, 873-1004, 1164-1233) at sizes
the oracle does not reach."""
import ctypes

import numpy as np
import pytest

import pyoracle as O

import morpheus_amd as M
from morpheus_amd import ring_ajtai as R

pytestmark = pytest.mark.gpu
Q = O.Q
lib = M.lib
KEY = bytes((11 * i + 5) % 256 for i in range(32))


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert lib.almost_goldilocks_cuda_init() == 0


def challenge(rng, wide=False):
    """{-1, 0, 1, 2} as the reference samples; wide: any int8, the extremes included"""
    if wide:
        c = rng.integers(-128, 128, 64, dtype=np.int8)
        c[:4] = [-128, 127, -2, 3]
        return R.RingChallenge(c)
    return R.RingChallenge(rng.integers(-1, 3, 64, dtype=np.int8))


def r_all_of(chs):
    return np.concatenate([R.RingChallenge.constant_one().coeffs] + [c.coeffs for c in chs])


def bits(rng, n):
    z = rng.integers(0, 1 << 64, n, dtype=np.uint64)
    if n >= 3:
        z[:3] = [0, (1 << 64) - 1, 1 << 63]
    return z


@pytest.mark.parametrize("n,wide", [(1, False), (5, False), (1000, True), (70001, False)])
def test_fold_witness(n, wide):
    print(f"fold_witness n={n} wide={wide}", flush=True)
    rng = np.random.default_rng(100 + n)
    z1, z2, r = bits(rng, n), bits(rng, n), challenge(rng, wide)
    got = R.fold_witness(z1, r, z2)
    assert np.array_equal(got, O.ring_fold_witness(z1, z2, r.coeffs))


def test_fold_witness_chunk_and_args():
    d = M.DeviceBuffer(64)
    rc = np.zeros(64, dtype=np.int8)
    assert lib.ajtai_fold_witness_ffi(d.ptr, d.ptr, rc.ctypes.data, 1, 100, d.ptr) == -3
    assert lib.ajtai_fold_witness_ffi(d.ptr, d.ptr, rc.ctypes.data, 0, 64, d.ptr) == 0
    with pytest.raises(M.GpuError):
        R.fold_witness(np.zeros(2, np.uint64), R.RingChallenge(rc), np.zeros(3, np.uint64))


def test_fold_commitment_is_commit_of_folded_witness():
    rng = np.random.default_rng(7)
    n = 300
    seed = R.Seed.from_bytes(KEY)
    z1, z2, r = bits(rng, n), bits(rng, n), challenge(rng)
    c1, c2 = R.commit(seed, z1), R.commit(seed, z2)
    cf = R.fold_commitment(c1, r, c2)
    assert np.array_equal(cf.rows.ravel(), O.ring_fold_commitment(c1.rows, c2.rows, r.coeffs))
    assert cf == R.commit_wide(seed, R.fold_witness(z1, r, z2))
    cw = challenge(rng, wide=True)
    assert np.array_equal(R.fold_commitment(c1, cw, c2).rows.ravel(),
                          O.ring_fold_commitment(c1.rows, c2.rows, cw.coeffs))


@pytest.mark.parametrize("m,n,wide", [(1, 1, False), (3, 100, True), (64, 3000, False), (511, 2, False)])
def test_multifold_witness(m, n, wide):
    print(f"multifold_witness m={m} n={n} wide={wide}", flush=True)
    rng = np.random.default_rng(200 + m)
    ws = [bits(rng, n) for _ in range(m)]
    chs = [challenge(rng, wide) for _ in range(m - 1)]
    got = R.multifold_witness(ws, chs)
    want = O.ring_multifold_mixed_witness(np.concatenate(ws), None, None, r_all_of(chs), m, 0, n)
    assert np.array_equal(got, want)


def test_multifold_limits():
    d = M.DeviceBuffer(8 * 512)
    r = M.DeviceBuffer(512 * 8)
    assert lib.ajtai_multifold_witness_ffi(d.ptr, r.ptr, d.ptr, 512, 1, 1024) == -1
    assert lib.ajtai_multifold_witness_ffi(d.ptr, r.ptr, d.ptr, 2, 1, 0) == -1
    assert lib.ajtai_multifold_witness_ffi(d.ptr, r.ptr, d.ptr, 0, 1, 1024) == -1
    with pytest.raises(M.GpuError):
        R.multifold_witness([np.zeros(1, np.uint64)] * 2, [])


def split_chunks(rng, n):
    v = rng.integers(-8191, 8192, n * 64).astype(np.int16)
    v[:6] = [0, 1, -1, 8191, -8191, -4096][:min(6, v.size)]
    pos, neg = O.ring_split_witness(v, n)
    return v, R.TernaryChunksDevice.from_host(R.TernaryChunks(n, 13, pos, neg)), pos, neg


@pytest.mark.parametrize("k,n", [(2, 1), (0, 100), (50, 1000)])
@pytest.mark.parametrize("form", ["plain", "tc", "tc_fused"])
def test_multifold_mixed_witness(k, n, form):
    print(f"multifold_mixed k={k} n={n} form={form}", flush=True)
    rng = np.random.default_rng(300 + k + n)
    _, tern, pos, neg = split_chunks(rng, n)
    ws = [bits(rng, n) for _ in range(k)]
    chs = [challenge(rng) for _ in range(k + 13 - 1)]
    if k == 0:
        chs[3] = challenge(rng, wide=True)
    fn = {"plain": R.multifold_mixed_witness, "tc": R.multifold_mixed_witness_tc,
          "tc_fused": R.multifold_mixed_witness_tc_fused}[form]
    got = fn(ws, tern, chs)
    want = O.ring_multifold_mixed_witness(np.concatenate(ws) if k else None, pos, neg, r_all_of(chs), k, 13, n)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("m", [1, 5, 63])
def test_multifold_commitment(m):
    rng = np.random.default_rng(400 + m)
    cs = [R.RingCommitment(rng.integers(0, Q, 42 * 64, dtype=np.uint64)) for _ in range(m)]
    cs[0].rows[0, :2] = [0, Q - 1]
    chs = [challenge(rng, wide=(m == 5)) for _ in range(m - 1)]
    got = R.multifold_commitment(cs, chs)
    want = O.ring_multifold_commitment(np.concatenate([c.rows.ravel() for c in cs]), r_all_of(chs), m)
    assert np.array_equal(got.rows.ravel(), want)


@pytest.mark.parametrize("n", [1, 77, 70001])
def test_split_witness(n):
    print(f"split_witness n={n}", flush=True)
    rng = np.random.default_rng(500 + n)
    v = rng.integers(-8191, 8192, n * 64).astype(np.int16)
    v[:6] = [0, 1, -1, 8191, -8191, -4096][:min(6, v.size)]
    got = R.split_witness(v)
    pos, neg = O.ring_split_witness(v, n)
    assert got.k_chunks == 13 and got.n_ring == n
    assert np.array_equal(got.pos, pos) and np.array_equal(got.neg, neg)


def test_split_empty_and_bad_length():
    assert R.split_witness(np.zeros(0, np.int16)).n_ring == 0
    with pytest.raises(M.GpuError):
        R.split_witness(np.zeros(65, np.int16))


@pytest.mark.parametrize("n,chunk", [(1, None), (100, 64), (400, 4096), (333, 256)])
def test_commit_ternary_and_premat(n, chunk):
    print(f"commit_ternary n={n} chunk={chunk}", flush=True)
    rng = np.random.default_rng(600 + n)
    seed = R.Seed.from_bytes(KEY)
    _, tern, pos, neg = split_chunks(rng, n)
    want = O.ring_commit_ternary(KEY, pos, neg, n)
    got = R.commit_ternary(seed, tern, chunk)
    assert all(np.array_equal(got[i].rows, want[i]) for i in range(13))
    mm = R.MaterializedM(seed, n)
    got2 = R.commit_ternary_premat(mm, tern, chunk)
    assert all(np.array_equal(got2[i].rows, want[i]) for i in range(13))


@pytest.mark.parametrize("n", [1, 3, 50])
def test_materialize(n):
    mm = R.MaterializedM(R.Seed.from_bytes(KEY), n)
    assert np.array_equal(mm.buf.to_numpy(), O.ring_materialize(KEY, n))


def test_commit_ternary_bad_args():
    d = M.DeviceBuffer(13 * 42 * 64)
    assert lib.ajtai_commit_ternary_ffi(d.ptr, d.ptr, d.ptr, 1, 100, d.ptr) == -3
    assert lib.ajtai_commit_ternary_premat_ffi(d.ptr, d.ptr, d.ptr, 1, 100, d.ptr) == -3
    assert lib.ajtai_commit_ternary_ffi(d.ptr, d.ptr, d.ptr, 0, 64, d.ptr) == -1
    with pytest.raises(M.GpuError):
        R.commit_ternary_premat(R.MaterializedM(R.Seed.from_bytes(KEY), 2),
                                R.TernaryChunksDevice(3, 13, d, d))


def test_fold_split_commit_pipeline_homomorphism():
    """At a size past the oracle: the mixed multifold of 50 binary witnesses and a split
    witness, split again, ternary-committed, recombines as sum 2^i c_i to the wide commit of
    the folded witness, and that equals the multifold of the commitments."""
    rng = np.random.default_rng(77)
    n, k = 1 << 14, 50
    seed = R.Seed.from_bytes(KEY)
    v0 = rng.integers(-3, 4, n * 64).astype(np.int16)
    t0 = R.split_witness_device(R._bytes_buffer(v0), n)
    ws = [bits(rng, n) for _ in range(k)]
    chs = [challenge(rng) for _ in range(k + 13 - 1)]
    folded = R.multifold_mixed_witness(ws, t0, chs)
    assert np.abs(folded.astype(np.int64)).max() < 8192
    ct = R.commit_ternary(seed, R.split_witness_device(R._bytes_buffer(folded), n))
    acc = np.zeros(42 * 64, dtype=object)
    for i in range(13):
        acc = (acc + (1 << i) * ct[i].rows.ravel().astype(object)) % Q
    vf = np.array([int(x) % Q for x in folded.astype(np.int64)], dtype=np.uint64)
    wide = R.commit_wide(seed, vf)
    assert [int(x) for x in acc] == [int(x) for x in wide.rows.ravel()]
    # commitments of the 50 binary witnesses and of the 13 input chunks, multifolded
    c_bin = R.commit_batched(seed, ws[:16]) + R.commit_batched(seed, ws[16:32]) + \
        R.commit_batched(seed, ws[32:48]) + R.commit_batched(seed, ws[48:50])
    c_t0 = R.commit_ternary(seed, t0)
    assert R.multifold_commitment(c_bin + c_t0, chs) == wide


@pytest.mark.parametrize("total,n_planes,bx", [(1, 1, 1), (300, 3, 7), (5000, 4, 64), (1 << 16, 2, 256)])
def test_selective_add_batched_planes(total, n_planes, bx):
    print(f"selective_add total={total} planes={n_planes} bx={bx}", flush=True)
    rng = np.random.default_rng(800 + total)
    packed = (total + 63) // 64
    eq = rng.integers(0, Q, 2 * total, dtype=np.uint64)
    planes = rng.integers(0, 1 << 64, n_planes * packed, dtype=np.uint64)
    planes[0] = (1 << 64) - 1
    de, dp = M.DeviceBuffer.from_numpy(eq), M.DeviceBuffer.from_numpy(planes)
    part = M.DeviceBuffer(2 * n_planes * bx)
    assert lib.aext2_selective_add_batched_planes_ffi(de.ptr, dp.ptr, part.ptr, total, n_planes, packed, bx) == 0
    p = part.to_numpy().reshape(bx, n_planes, 2)
    got = [[sum(int(x) for x in p[:, pl, c]) % Q for c in range(2)] for pl in range(n_planes)]
    want = O.q_selective_add_planes(eq, planes, total, n_planes, packed)
    assert got == [[int(x) for x in w] for w in want]
