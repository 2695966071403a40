"""GPU parity of the device challengers (challenger_* entry points) with the oracle's duplex
transcript used with empty labels (an observe is an append of one value, a sample is a
challenge with no label), plus the bit permutation and Ext2 scale-accumulate entry points."""
import numpy as np
import pytest

import pyoracle as O

import morpheus_amd as M
from morpheus_amd.challenger import ChallengerBatch, DuplexChallenger

pytestmark = pytest.mark.gpu
P = O.P


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.goldilocks_cuda_init() == 0


OPS = [("observe_slice", 3), ("sample", 1), ("observe", 1), ("sample_array", 5), ("observe_ext2", 2),
       ("sample_ext2", 2), ("observe_slice", 4), ("sample_ext5", 5), ("observe_slice", 7), ("sample_array", 9),
       ("observe", 1), ("observe", 1), ("sample", 1), ("sample", 1), ("observe_slice", 1), ("sample_array", 4)]


def run_ops(ch, ts, rng, ops):
    n = ch.count
    for op, k in ops:
        if op.startswith("observe"):
            vals = rng.integers(0, 1 << 64, (n, k), dtype=np.uint64)
            vals[0, 0] = np.uint64(P + 5)  # a non-canonical word is absorbed as its residue
            if op == "observe":
                ch.observe(vals[:, 0])
            elif op == "observe_ext2":
                ch.observe_ext2(vals)
            else:
                ch.observe_slice(vals, k)
            for i, t in enumerate(ts):
                t.append_scalars(b"", vals[i])
        else:
            got = {"sample": ch.sample, "sample_ext2": ch.sample_ext2, "sample_ext5": ch.sample_ext5}[op]() \
                if op != "sample_array" else ch.sample_array(k)
            got = np.asarray(got).reshape(n, -1)
            for i, t in enumerate(ts):
                want = [t.challenge_scalar(b"") for _ in range(k)]
                assert [int(x) for x in got[i]] == want, (op, i)


def test_challenger_batch_matches_oracle_duplex():
    print("challenger batch", flush=True)
    n = 300
    rng = np.random.default_rng(3)
    ch = ChallengerBatch(n)
    ts = [O.Transcript(b"") for _ in range(n)]
    run_ops(ch, ts, rng, OPS)
    # the states themselves: sponge, counts and the live buffer entries
    for i in (0, 7, n - 1):
        sp, inb, outb, ni, no = ch.state(i)
        t = ts[i].t
        assert np.array_equal(sp, np.array(t.state, dtype=np.uint64))
        assert (ni, no) == (t.in_count, t.out_count)
        assert np.array_equal(inb[:ni], np.array(t.in_buf, dtype=np.uint64)[:ni])
        assert np.array_equal(outb[:no], np.array(t.out_buf, dtype=np.uint64)[:no])
    ch.free()


def test_state_round_trip_and_single_challenger():
    print("challenger state copy", flush=True)
    rng = np.random.default_rng(4)
    a, b = ChallengerBatch(4), ChallengerBatch(4)
    a.observe_slice(rng.integers(0, P, 4 * 6, dtype=np.uint64), 6)
    a.sample()
    raw = a.raw_state(2)
    b.load_state(1, raw)
    assert b.raw_state(1) == raw
    vals = rng.integers(0, P, 4 * 3, dtype=np.uint64)
    a.observe_slice(vals, 3)
    b.observe_slice(vals.reshape(4, 3)[[0, 2, 2, 3]].ravel(), 3)  # b[1] replays a[2]'s input
    assert np.array_equal(a.sample_array(6)[2], b.sample_array(6)[1])
    # the single-challenger mirror equals a fresh oracle transcript
    d, t = DuplexChallenger(), O.Transcript(b"")
    for v in (1, 2, 3, P - 1, 1 << 40):
        d.observe(v)
        t.append_u64(b"", v)
    assert d.sample() == t.challenge_scalar(b"")
    assert [int(x) for x in d.sample_ext2()] == [t.challenge_scalar(b""), t.challenge_scalar(b"")]


@pytest.mark.parametrize("n_bits,seed", [(1, 1), (5, 2), (12, 3), (20, 4)])
def test_bit_permute(n_bits, seed):
    print("bit_permute", n_bits, flush=True)
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 1 << 64, 1 << n_bits, dtype=np.uint64)
    inv = rng.permutation(n_bits).astype(np.int32)
    d_map = M.DeviceBuffer.from_numpy(np.concatenate([inv, np.zeros(n_bits % 2, np.int32)]).view(np.uint64))
    d_in, d_out = M.DeviceBuffer.from_numpy(x), M.DeviceBuffer(1 << n_bits)
    assert M.lib.bit_permute_gl_ffi(d_in.ptr, d_out.ptr, d_map.ptr, n_bits, 1 << n_bits) == 0
    idx = np.arange(1 << n_bits)
    src = np.zeros_like(idx)
    for b in range(n_bits):
        src |= ((idx >> b) & 1) << int(inv[b])
    assert np.array_equal(d_out.to_numpy(), x[src])


def test_ext2_scale_accumulate():
    print("ext2_scale_accumulate", flush=True)
    rng = np.random.default_rng(5)
    n = 1000
    src = rng.integers(0, P, 2 * n, dtype=np.uint64)
    acc = rng.integers(0, P, 2 * n, dtype=np.uint64)
    s = [int(x) for x in rng.integers(0, P, 2, dtype=np.uint64)]
    d_src, d_acc = M.DeviceBuffer.from_numpy(src), M.DeviceBuffer.from_numpy(acc)
    assert M.lib.ext2_scale_accumulate_ffi(s[0], s[1], d_src.ptr, d_acc.ptr, n) == 0
    got = d_acc.to_numpy().reshape(n, 2)
    for i in range(0, n, 37):
        a0, a1 = int(src[2 * i]), int(src[2 * i + 1])
        prod = [(s[0] * a0 + 7 * s[1] * a1) % P, (s[0] * a1 + s[1] * a0) % P]
        assert [int(got[i][0]), int(got[i][1])] == [(int(acc[2 * i]) + prod[0]) % P, (int(acc[2 * i + 1]) + prod[1]) % P]
