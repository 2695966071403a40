"""GPU tests of batched proving (mam_prove_batch): every instance's proof equals the oracle's
(product sumcheck followed by a Basefold opening of f at the sumcheck challenges on one
transcript), and equals itself proved alone (one worker) and in the mix (several workers)."""
import numpy as np
import pytest

import pyoracle as O
from refmath import rand_field

import morpheus_amd as M
from morpheus_amd.batch import Instance, prove_batch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.goldilocks_cuda_init() == 0


def oracle_instance(f, g, label, nv, table, lr, hash_kind, nq):
    ext = np.zeros((2, 1 << nv, 2), dtype=np.uint64)
    ext[0, :, 0], ext[1, :, 0] = f, g
    t = O.Transcript(label)
    sc = O.sumcheck_prove_ext2(t, ext, 2, nv)
    cw, bh, tree = O.basefold_commit(f, nv, lr, hash_kind)
    proof = O.basefold_open_ext2(cw, bh, tree[-4:], nv, lr, hash_kind, sc["challenges"].ravel(), table.entries,
                                 table.num_vars, t, nq)
    return sc, tree[-4:], proof


def make(nvs, seed):
    rng = np.random.default_rng(seed)
    data = [(rand_field(rng, 1 << nv), rand_field(rng, 1 << nv)) for nv in nvs]
    return data, [Instance(f, g, b"inst-%d" % i) for i, (f, g) in enumerate(data)]


def test_batch_matches_oracle_and_sequential():
    print("batch oracle", flush=True)
    nvs = [3, 7, 5, 9, 4, 6]
    lr, nq, tnv = 1, 12, 9
    table = M.BasefoldTable.generate(tnv, lr, tnv, 42)
    data, insts = make(nvs, 1)
    seq = prove_batch(insts, table, lr, 0, nq, workers=1)
    mix = prove_batch(insts, table, lr, 0, nq, workers=3)
    for i, nv in enumerate(nvs):
        sc, root, proof = oracle_instance(data[i][0], data[i][1], b"inst-%d" % i, nv, table, lr, 0, nq)
        for k in ("round_msgs", "challenges", "final_evals"):
            assert np.array_equal(seq[i][k], sc[k]), (i, k)
        assert np.array_equal(seq[i]["root"], root)
        assert np.array_equal(seq[i]["proof"], proof)
        # f(r) from the sumcheck equals the Basefold opening's evaluation claim
        assert np.array_equal(seq[i]["final_evals"][0], proof[:2])
        for k in ("round_msgs", "challenges", "final_evals", "root", "proof"):
            assert np.array_equal(mix[i][k], seq[i][k]), (i, k)


def test_batch_into_pinned_buffers():
    """Proofs written into caller-owned page-locked buffers (PinnedArray, larger than needed)
    equal the library-allocated ones; the words past the proof stay untouched."""
    print("batch pinned", flush=True)
    nvs = [6, 9, 4]
    lr, nq, tnv = 1, 10, 9
    table = M.BasefoldTable.generate(tnv, lr, tnv, 42)
    _, insts = make(nvs, 3)
    ref = prove_batch(insts, table, lr, 0, nq, workers=2)
    bufs = [M.PinnedArray(M.proof_len(nv, lr, nq) + 5) for nv in nvs]
    for b in bufs:
        b.a[:] = 0xABCD
    got = prove_batch(insts, table, lr, 0, nq, workers=2, proof_bufs=[b.a for b in bufs])
    for i, nv in enumerate(nvs):
        plen = M.proof_len(nv, lr, nq)
        assert np.array_equal(got[i]["proof"], ref[i]["proof"]), i
        assert np.array_equal(bufs[i].a[:plen], ref[i]["proof"]), i
        assert np.all(bufs[i].a[plen:] == 0xABCD), i
    for b in bufs:
        b.free()


def test_batch_uneven_sizes_many_workers():
    print("batch uneven", flush=True)
    nvs = [12, 2, 10, 2, 11, 8, 12, 3, 9, 6, 7, 5]
    lr, nq, tnv = 1, 20, 12
    table = M.BasefoldTable.generate(tnv, lr, tnv, 7)
    data, insts = make(nvs, 2)
    seq = prove_batch(insts, table, lr, 0, nq, workers=1)
    for w in (4, 8):
        mix = prove_batch(insts, table, lr, 0, nq, workers=w)
        for i in range(len(nvs)):
            for k in ("round_msgs", "challenges", "final_evals", "root", "proof"):
                assert np.array_equal(mix[i][k], seq[i][k]), (w, i, k)
    # every opening verifies against the transcript state the sumcheck leaves behind
    for i, nv in enumerate(nvs):
        ext = np.zeros((2, 1 << nv, 2), dtype=np.uint64)
        ext[0, :, 0], ext[1, :, 0] = data[i]
        ot = O.Transcript(b"inst-%d" % i)
        O.sumcheck_prove_ext2(ot, ext, 2, nv)
        assert O.basefold_verify_ext2(seq[i]["root"], seq[i]["challenges"].ravel(), nv, lr, 0, seq[i]["proof"], nq,
                                      table.entries, tnv, ot)


def test_batch_rejects_bad_arguments():
    table = M.BasefoldTable.generate(4, 1, 4, 1)
    _, insts = make([5], 3)
    with pytest.raises(M.GpuError):
        prove_batch(insts, table, 1, 0, 4, workers=1)  # num_vars above the table's
    _, insts = make([3], 3)
    with pytest.raises(M.GpuError):
        prove_batch(insts, table, 1, 0, 4, workers=0)
    _, insts = make([1], 3)  # an opening needs at least two variables
    with pytest.raises(M.GpuError):
        prove_batch(insts, table, 1, 0, 4, workers=1)
