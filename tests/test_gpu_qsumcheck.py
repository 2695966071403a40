"""GPU parity of the zk-torch-4 linear sumcheck over the q Ext2 (mam_q_sumcheck_prove,
zk-torch-4/src/sumcheck/gpu_prover.rs:26-112) against the oracle's CPU prover
(cpu_ext2_prover.rs:23-100): round messages, challenges, final values and the transcript state
afterwards, word for word; at 2^20 and 2^22 elements, acceptance by the host verifier, the
claim from the oracle's round-0 sum and the final values as full multilinear evaluations."""
import numpy as np
import pytest

import pyoracle as O

import morpheus_amd as M
from morpheus_amd.qprover import QTranscript, q_sumcheck_prove, q_sumcheck_verify

pytestmark = pytest.mark.gpu
Q = O.Q


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.almost_goldilocks_cuda_init() == 0


def inputs(rng, np_, nv, base):
    n = 1 << nv
    raw = rng.integers(0, 1 << 64, np_ * n * (1 if base else 2), dtype=np.uint64)
    edge = np.array([Q, (1 << 64) - 1, Q - 1], dtype=np.uint64)   # non-canonical words, the top residue
    raw[:min(3, raw.size)] = edge[:min(3, raw.size)]
    return raw


@pytest.mark.parametrize("np_,nv,base", [(1, 1, True), (2, 1, False), (3, 5, True), (8, 4, False), (2, 12, True),
                                         (4, 10, False), (5, 14, True), (3, 16, False), (1, 17, True)])
def test_q_sumcheck_matches_oracle(np_, nv, base):
    print(f"q_sumcheck np={np_} nv={nv} base={base}", flush=True)
    rng = np.random.default_rng(1000 * np_ + nv)
    raw = inputs(rng, np_, nv, base)
    label = b"zk-torch-4-sumcheck"
    t_ref = O.QTranscript(label)
    msgs, chs, fin, fe = O.q_sumcheck_prove(t_ref, raw, np_, nv, not base)
    t = QTranscript(label)
    got = q_sumcheck_prove(t, M.DeviceBuffer.from_numpy(raw), np_, nv, base)
    assert np.array_equal(got["round_msgs"], msgs)
    assert np.array_equal(got["challenges"], chs)
    assert np.array_equal(got["final_evals"], fin)
    assert np.array_equal(got["final_eval"], fe)
    assert np.array_equal(t.get_state(), t_ref.state())
    # the transcript continues identically
    assert np.array_equal(t.challenge_vector(b"next", 3), t_ref.challenge(b"next", 3))


@pytest.mark.parametrize("np_,nv,base", [(2, 20, True), (3, 22, False)])
def test_q_sumcheck_full_size_properties(np_, nv, base):
    print(f"q_sumcheck full np={np_} nv={nv} base={base}", flush=True)
    rng = np.random.default_rng(77 + nv)
    raw = inputs(rng, np_, nv, base)
    n = 1 << nv
    d = M.DeviceBuffer.from_numpy(raw)
    got = q_sumcheck_prove(QTranscript(b"full"), d, np_, nv, base)
    m0 = O.q_sumcheck_msg(0 if base else 1, raw, np_, n, n // 2)
    claim = ((int(m0[0]) + int(m0[2])) % Q, (int(m0[1]) + int(m0[3])) % Q)
    ok, chs = q_sumcheck_verify(QTranscript(b"full"), claim, np_, nv, got["round_msgs"], got["final_eval"])
    assert ok
    assert np.array_equal(np.array(chs, dtype=np.uint64), got["challenges"])
    point = got["challenges"].ravel()
    for i in range(np_):
        poly = raw[i * n:(i + 1) * n] if base else raw[2 * i * n:2 * (i + 1) * n]
        v = O.q_partial_eval(poly, 0 if base else 1, nv, point, 1, nv)
        assert np.array_equal(v, got["final_evals"][i])


def test_q_sumcheck_rejects_bad_arguments():
    d = M.DeviceBuffer(16)
    out = np.zeros(64, dtype=np.uint64)
    p = out.ctypes.data_as(M.qprover._u64p)
    t = QTranscript(b"x")
    assert M.lib.mam_q_sumcheck_prove(M.qprover.ctypes.byref(t.t), d.ptr, 9, 2, 1, p, p, p, p) == -2
    assert M.lib.mam_q_sumcheck_prove(M.qprover.ctypes.byref(t.t), d.ptr, 1, 0, 1, p, p, p, p) == -2
    assert M.lib.mam_q_sumcheck_prove(M.qprover.ctypes.byref(t.t), None, 1, 2, 1, p, p, p, p) == -2
