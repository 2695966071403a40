"""The C1 proof on the GPU serializes to the same bincode bytes (and fingerprint) as the
oracle's proof of the same statement: bit-identical transcript bytes."""
import numpy as np
import pytest

import pyoracle as O

import morpheus_amd as M
from morpheus_amd.prover import fingerprint, sumcheck_proof_bincode

pytestmark = pytest.mark.gpu
P = O.P


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.goldilocks_cuda_init() == 0


@pytest.mark.parametrize("nv,seed", [(10, 1), (14, 2)])
def test_sumcheck_proof_bytes(nv, seed):
    print("wire", nv, flush=True)
    rng = np.random.default_rng(seed)
    fg = rng.integers(0, P, 2 << nv, dtype=np.uint64)
    got = M.sumcheck_prove(M.Transcript(b"c1-wire"), M.DeviceBuffer.from_numpy(fg), 2, nv, base=True)
    ext = np.zeros((2, 1 << nv, 2), dtype=np.uint64)
    ext[:, :, 0] = fg.reshape(2, -1)
    want = O.sumcheck_prove_ext2(O.Transcript(b"c1-wire"), ext, 2, nv)
    b = sumcheck_proof_bincode(got)
    assert b == O.sumcheck_proof_bincode(want["final_eval"], want["round_msgs"])
    assert fingerprint(b) == fingerprint(O.sumcheck_proof_bincode(want["final_eval"], want["round_msgs"]))
