"""The zk-torch-3 transcript on the device (mam_dtranscript_run) against the host transcript
and the oracle, on random labelled sequences (transcript.rs:167-330 semantics: new with raw
label bytes, length-prefixed labels, empty labels skipped, overwrite duplex, reversed output
buffer), and the lane-parallel Poseidon2-w8 against the scalar host permutation."""
import ctypes

import numpy as np
import pytest

import pyoracle as O

import morpheus_amd as M
from morpheus_amd.prover import DeviceTranscript, transcript_program

pytestmark = pytest.mark.gpu
P = M.P


def test_wide_permutation_matches_scalar():
    rng = np.random.default_rng(5)
    n = 1000
    st = rng.integers(0, 1 << 64, 8 * n, dtype=np.uint64)
    st[:8] = np.arange(8, dtype=np.uint64)
    st[8:16] = [P, P + 1, (1 << 64) - 1, 0, 1, P - 1, 2, 3]
    d = M.DeviceBuffer.from_numpy(st)
    assert M.lib.mam_poseidon2_wide_permute(d.ptr, n) == 0
    got = d.to_numpy().reshape(n, 8)
    for i in list(range(20)) + [n - 1]:
        s = (ctypes.c_uint64 * 8)(*[int(v) for v in st[8 * i:8 * i + 8]])
        M.lib.mam_poseidon2_permute_host(s)
        assert [int(v) for v in got[i]] == list(s), i


def random_ops(rng, count):
    ops = []
    for _ in range(count):
        label = bytes(rng.integers(0, 256, int(rng.integers(0, 20)), dtype=np.uint8))
        if rng.integers(0, 3) == 0:
            ops.append(("challenge", label, int(rng.integers(1, 7))))
        else:
            vals = rng.integers(0, 1 << 64, int(rng.integers(0, 9)), dtype=np.uint64)
            ops.append(("append", label, vals))
    return ops


def host_run_exact(t, ops):
    """challenge with n draws = label once, then n squeezes (append_scalars / challenge_ext2)"""
    out = []
    for kind, label, arg in ops:
        if kind == "append":
            t.append_scalars(label, arg)
        else:
            out.append(t.challenge_scalar(label))
            for _ in range(arg - 1):
                out.append(t.challenge_scalar(b""))
    return out


@pytest.mark.parametrize("seed", range(6))
def test_device_transcript_matches_host(seed):
    rng = np.random.default_rng(100 + seed)
    init = bytes(rng.integers(0, 256, int(rng.integers(0, 30)), dtype=np.uint8))
    ops = random_ops(rng, 40)
    host = M.Transcript(init)
    dev = DeviceTranscript(M.Transcript(init))
    want = host_run_exact(host, ops)
    got = dev.run(ops)
    assert [int(v) for v in got] == want
    assert np.array_equal(dev.to_host().state(), host.state())
    assert bytes(dev.to_host().t) == bytes(host.t)
    # a second program continues from the stored state, and the oracle agrees
    ops2 = random_ops(rng, 10)
    ref = O.Transcript(init)
    host_run_exact_oracle(ref, ops + ops2)
    dev.run(ops2)
    assert np.array_equal(dev.to_host().state(), np.array(ref.t.state, dtype=np.uint64))


def host_run_exact_oracle(t, ops):
    for kind, label, arg in ops:
        if kind == "append":
            t.append_scalars(label, arg)
        else:
            t.challenge_scalar(label)
            for _ in range(arg - 1):
                t.challenge_scalar(b"")


def test_reference_label_cases():
    """empty labels absorb nothing; Ext2 appends are two values after one label; a challenge
    right after a duplex pops state[0] first"""
    for ops in ([("append", b"", [5]), ("challenge", b"", 2)],
                [("append", b"round_message", [1, 2]), ("challenge", b"challenge", 2)],
                [("challenge", b"x" * 300, 9)]):
        host = M.Transcript(b"zk")
        want = host_run_exact(host, ops)
        dev = DeviceTranscript(M.Transcript(b"zk"))
        assert [int(v) for v in dev.run(ops)] == want
        assert bytes(dev.to_host().t) == bytes(host.t)


def test_malformed_program_is_reported():
    dev = DeviceTranscript()
    prog, src = transcript_program([("append", b"abc", [1, 2, 3])])
    dp, ds, dd, err = (M.DeviceBuffer.from_numpy(prog), M.DeviceBuffer.from_numpy(src), M.DeviceBuffer(4),
                       M.DeviceBuffer(1))
    # src shorter than the program needs
    assert M.lib.mam_dtranscript_run(dev.state.ptr, dp.ptr, prog.size, ds.ptr, 2, dd.ptr, 0, err.ptr) == 0
    assert int(err.to_numpy()[0]) & 0xFFFFFFFF == 1
    # truncated program
    assert M.lib.mam_dtranscript_run(dev.state.ptr, dp.ptr, 3, ds.ptr, 3, dd.ptr, 0, err.ptr) == 0
    assert int(err.to_numpy()[0]) & 0xFFFFFFFF == 1
    assert M.lib.mam_dtranscript_run(None, dp.ptr, 3, ds.ptr, 3, dd.ptr, 0, err.ptr) == -2
