"""GPU parity of the batch opening (mam_basefold_batch_open_ext2) with the oracle's
restatement of batch_open_ext2, and of the host batch verifier with the oracle's; at larger
sizes the proof is checked by both verifiers and tampering is rejected."""
import numpy as np
import pytest

import pyoracle as O

import morpheus_amd as M
from morpheus_amd.prover import EvaluationExt2, batch_open_ext2, batch_verify_ext2

pytestmark = pytest.mark.gpu
P = O.P


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.goldilocks_cuda_init() == 0


def mle_ext2(f, pt, nv):
    e = np.zeros((f.size, 2), dtype=np.uint64)
    e[:, 0] = f
    return O.partial_eval_ext2(e, nv, np.asarray(pt, dtype=np.uint64).ravel(), nv).ravel()[:2]


def make(nv, lr, npoly, npts, spec, hash_kind, seed):
    rng = np.random.default_rng(seed)
    polys = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(npoly)]
    pts = rng.integers(0, P, (npts, nv, 2), dtype=np.uint64)
    evals = [EvaluationExt2(pi, qi, mle_ext2(polys[pi], pts[qi], nv)) for pi, qi in spec]
    comms = [M.BasefoldCommitment.commit(f, nv, lr, hash_kind) for f in polys]
    return polys, pts, evals, comms


CASES = [
    (4, 1, 3, 2, [(0, 0), (1, 0), (2, 1), (0, 1)], M.HASH_MONOLITH, 7),
    (9, 1, 4, 3, [(0, 0), (1, 1), (2, 2), (3, 0), (1, 2), (0, 1)], M.HASH_MONOLITH, 9),
    (7, 2, 2, 1, [(0, 0), (1, 0)], M.HASH_POSEIDON2, 5),
    (10, 1, 1, 1, [(0, 0)], M.HASH_MONOLITH, 3),
    (6, 1, 20, 2, [(i, i % 2) for i in range(20)], M.HASH_MONOLITH, 4),  # more claims than one launch takes
]


@pytest.mark.parametrize("nv,lr,npoly,npts,spec,hash_kind,nq", CASES)
def test_batch_open_matches_oracle(nv, lr, npoly, npts, spec, hash_kind, nq):
    print("batch_open", nv, lr, npoly, npts, len(spec), hash_kind, flush=True)
    polys, pts, evals, comms = make(nv, lr, npoly, npts, spec, hash_kind, nv + npoly)
    table = M.BasefoldTable.generate(nv, lr, nv, 11)
    proof = batch_open_ext2(comms, pts, evals, table, M.Transcript(b"bo"), nq)
    oc = []
    for f in polys:
        cw, bh, tree = O.basefold_commit(f, nv, lr, hash_kind)
        oc.append((cw, bh, tree[-4:]))
    want = O.basefold_batch_open_ext2(oc, nv, lr, hash_kind, pts, [(e.poly, e.point, e.value) for e in evals],
                                      table.entries, nv, O.Transcript(b"bo"), nq)
    assert proof.size == want.size
    assert np.array_equal(proof, want)
    roots = np.array([c.root for c in comms])
    assert batch_verify_ext2(roots, pts, evals, proof, table, M.Transcript(b"bo"), nv, lr, nq, hash_kind)
    assert O.basefold_batch_verify_ext2(roots, nv, lr, hash_kind, pts, [(e.poly, e.point, e.value) for e in evals],
                                        proof, nq, table.entries, nv, O.Transcript(b"bo"))
    for c in comms:
        c.free()


def test_batch_open_large_verifies_and_rejects_tampering():
    nv, lr, nq = 18, 1, 40
    print("batch_open large", nv, flush=True)
    spec = [(0, 0), (1, 0), (2, 1), (1, 1), (0, 2)]
    polys, pts, evals, comms = make(nv, lr, 3, 3, spec, M.HASH_MONOLITH, 21)
    table = M.BasefoldTable.generate(nv, lr, nv, 12)
    proof = batch_open_ext2(comms, pts, evals, table, M.Transcript(b"big"), nq)
    roots = np.array([c.root for c in comms])

    def ver(pf, ev=evals):
        return batch_verify_ext2(roots, pts, ev, pf, table, M.Transcript(b"big"), nv, lr, nq)

    assert ver(proof)
    # the same commitments opened twice give the same proof (arena reuse)
    assert np.array_equal(proof, batch_open_ext2(comms, pts, evals, table, M.Transcript(b"big"), nq))
    bad_ev = list(evals)
    bad_ev[2] = EvaluationExt2(evals[2].poly, evals[2].point, [(evals[2].value[0] + 1) % P, evals[2].value[1]])
    assert not ver(proof, bad_ev)
    for pos in (3, 6 * nv + 1, proof.size // 2, proof.size - 5):
        bad = proof.copy()
        bad[pos] ^= np.uint64(1 << 33)
        assert not ver(bad), pos
    for c in comms:
        c.free()
