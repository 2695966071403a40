"""Parity of the batched-proving and proving-service workloads at their benchmark sizes.

Batched proving: the 64 seeded instances of log-uniform size 2^12..2^22 (the bench's own
inputs, morpheus_amd.workloads). Every instance of at most 2^16 entries equals the oracle's
proof word for word; every instance's sumcheck and Basefold opening are accepted by the
oracle's verifiers on a replayed transcript, its claim g(0) + g(1) equals sum f g computed
in numpy mod p, f(r) and g(r) equal the oracle's MLEs at the challenges, and f(r) is the
opening's evaluation. One worker, the autotuned worker count and the graph form give identical proofs.

Proving service: requests m x k x n = 256^3, 256 x 4096 x 256, 4096^3 and 1024 x 512 x 2048
(2^24-entry matrices, a 12-round degree-3 sumcheck, the int8 Ajtai prefix commit against a
2^24-column set-up). Each request's result is the same proved alone and in the service
under a budget that admits one request at a time and under one that mixes them; the 256^3
proof equals the oracle's word for word; for every request the oracle's sumcheck and
Basefold verifiers accept the replayed transcript, the claim equals C(r) and the final
evaluations equal A(r_l, r_i) and B(r_j, r_l) (oracle MLEs of the integer matrices), and
sampled Ajtai rows equal the exact integer row products mod p.
"""
import sys

import numpy as np
import pytest

import pyoracle as O
from refmath import P, e2mul, gl_dot_np

import morpheus_amd as M
from morpheus_amd import workloads as W
from morpheus_amd.ajtai import AjtaiI8, key_from_seed
from morpheus_amd.batch import Instance, autotune_workers, prove_batch, prove_batch_graph
from morpheus_amd.matmul import MatmulJob, lift_int, prove_matmul, serve

pytestmark = pytest.mark.gpu
LR, NQ = 1, 100


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.goldilocks_cuda_init() == 0


def lift(ev):
    e = np.zeros((ev.size, 2), dtype=np.uint64)
    e[:, 0] = ev
    return e


def mle(ev, point2):
    point2 = np.ascontiguousarray(point2, dtype=np.uint64).ravel()
    nv = point2.size // 2
    return O.partial_eval_ext2(lift(np.ascontiguousarray(ev, dtype=np.uint64)), nv, point2, nv).reshape(-1)[:2]


def pair(x):
    return (int(x[0]), int(x[1]))


def words(x):
    return np.array(x, dtype=np.uint64)


def test_batch_64_instances_at_config_scale():
    data = W.batch_instance_data(64)
    insts = [Instance(f, g, label) for f, g, label in data]
    tnv = max(i.num_vars for i in insts)
    table = M.BasefoldTable.generate(tnv, LR, tnv, 42)
    say("c4: 64 instances, one worker; sizes", sum(1 << i.num_vars for i in insts))
    seq = prove_batch(insts, table, LR, 0, NQ, workers=1)
    best, times = autotune_workers(insts, table, LR, 0, NQ)
    say("c4: autotuned workers", best, {k: round(v, 4) for k, v in times.items()})
    for w in sorted({best, 8}):
        mix = prove_batch(insts, table, LR, 0, NQ, workers=w)
        for i in range(len(insts)):
            for k in ("round_msgs", "challenges", "final_evals", "root", "proof"):
                assert np.array_equal(mix[i][k], seq[i][k]), (w, i, k)
    # the graph form (device transcripts, one graph launch for the sumchecks and commitments)
    for b, m in ((2, 256), (0, 64)):
        say("c4: graph form, branches", b, "max_wg", m)
        gr = prove_batch_graph(insts, table, LR, 0, NQ, workers=8, branches=b, max_wg=m)
        for i in range(len(insts)):
            for k in ("round_msgs", "challenges", "final_evals", "root", "proof"):
                assert np.array_equal(gr[i][k], seq[i][k]), ("graph", b, m, i, k)
    assert M.lib.mam_batch_graph_release() == 0
    say("c4: verifying every instance")
    for i, (f, g, label) in enumerate(data):
        nv, o = insts[i].num_vars, seq[i]
        claim = gl_dot_np(f, g)
        m0 = o["round_msgs"][0]
        assert ((int(m0[0][0]) + int(m0[1][0])) % P, (int(m0[0][1]) + int(m0[1][1])) % P) == (claim, 0), i
        r = o["challenges"].ravel()
        fr, gr = mle(f, r), mle(g, r)
        assert np.array_equal(o["final_evals"][0], fr), i
        assert np.array_equal(o["final_evals"][1], gr), i
        assert np.array_equal(o["proof"][:2], fr), i
        t = O.Transcript(label)
        assert O.sumcheck_verify_ext2(t, words([claim, 0]), 2, nv, o["round_msgs"], words(e2mul(pair(fr), pair(gr)))), i
        assert O.basefold_verify_ext2(o["root"], r, nv, LR, 0, o["proof"], NQ, table.entries, tnv, t), i
        if nv <= 16:
            ext = np.zeros((2, 1 << nv, 2), dtype=np.uint64)
            ext[0, :, 0], ext[1, :, 0] = f, g
            ot = O.Transcript(label)
            sc = O.sumcheck_prove_ext2(ot, ext, 2, nv)
            cw, bh, tree = O.basefold_commit(f, nv, LR, O.MONOLITH)
            proof = O.basefold_open_ext2(cw, bh, tree[-4:], nv, LR, O.MONOLITH, sc["challenges"].ravel(), table.entries,
                                         tnv, ot, NQ)
            for k in ("round_msgs", "challenges", "final_evals"):
                assert np.array_equal(o[k], sc[k]), (i, k)
            assert np.array_equal(o["root"], tree[-4:]), i
            assert np.array_equal(o["proof"], proof), i


def verify_matmul(a, b, label, o, table_entries, tnv, key):
    """Checks one matmul proof with the oracle's verifiers and the integer matrices."""
    c = (a.astype(np.float64) @ b.astype(np.float64)).astype(np.int64)  # exact: |c| < 2^53
    m, k = a.shape
    n = b.shape[1]
    lm, lk, ln = (m - 1).bit_length(), (k - 1).bit_length(), (n - 1).bit_length()
    A, B, C = lift_int(a).ravel(), lift_int(b).ravel(), lift_int(c).ravel()
    t = O.Transcript(label)
    for root in o["roots"]:
        t.append_scalars(b"matmul-root", root)
    for row in (0, 1, 100, 255):
        mrow = O.ajtai_i8_row(key, row, C.size).astype(np.int64)
        assert int(o["ajtai"][row]) == int(mrow @ c.ravel()) % P, row  # |row . c| < 2^57
    t.append_scalars(b"matmul-ajtai", o["ajtai"])
    point = np.array([t.challenge_ext2(b"matmul-point") for _ in range(ln + lm)], dtype=np.uint64)
    assert np.array_equal(point, o["point"])
    claim = mle(C, point)
    assert np.array_equal(o["claim"], claim)
    t.append_ext2(b"matmul-claim", claim)
    r_j, r_i, r_l = point[:ln], point[ln:], o["challenges"]
    a_at = mle(A, np.concatenate([r_l, r_i]))
    b_at = mle(B, np.concatenate([r_j, r_l]))
    assert np.array_equal(o["final_evals"][0], a_at)
    assert np.array_equal(o["final_evals"][1], b_at)
    assert [int(x) for x in o["final_evals"][2]] == [1, 0]
    assert O.sumcheck_verify_ext2(t, claim, 3, lk, o["round_msgs"], words(e2mul(pair(a_at), pair(b_at))))
    pts = (np.concatenate([r_l, r_i]), np.concatenate([r_j, r_l]), point)
    for name, root, nv, pt in zip(("proof_a", "proof_b", "proof_c"), o["roots"], (lm + lk, lk + ln, lm + ln), pts):
        assert O.basefold_verify_ext2(root, pt.ravel(), nv, LR, 0, o[name], NQ, table_entries, tnv, t), name
    assert np.array_equal(o["proof_a"][:2], a_at)
    assert np.array_equal(o["proof_b"][:2], b_at)
    assert np.array_equal(o["proof_c"][:2], claim)


def test_matmul_service_at_config_scale():
    dims = [(256, 256, 256), (256, 4096, 256), (4096, 4096, 4096), (1024, 512, 2048)]
    ops = W.matmul_operands(dims, seed=31)
    labels = [b"c5-scale-%d" % i for i in range(len(dims))]
    say("c5: uploading", dims)
    jobs = [MatmulJob.from_int8(a, b, lab) for (a, b), lab in zip(ops, labels)]
    tnv = max(max(j.lm + j.lk, j.lk + j.ln, j.lm + j.ln) for j in jobs)
    table = M.BasefoldTable.generate(tnv, LR, tnv, 42)
    key = key_from_seed(2024)
    aj = AjtaiI8(key, 256, 1 << 24)
    say("c5: alone")
    alone = [prove_matmul(j, table, LR, 0, NQ, aj) for j in jobs]
    zero = np.zeros(len(jobs))
    for workers, budget in ((3, 1 << 20), (4, 12 << 30)):
        say("c5: served", workers, budget)
        outs, admit, done = serve(jobs, zero, table, LR, 0, NQ, aj, workers, budget)
        for i in range(len(jobs)):
            for k in alone[i]:
                assert np.array_equal(outs[i][k], alone[i][k]), (workers, i, k)
        assert np.all(done >= admit)
        if budget < (1 << 30):
            # every request is above the budget, so each runs alone: admissions queue
            order = np.argsort(admit)
            assert np.all(admit[order[1:]] >= done[order[:-1]] - 1e-3)
    say("c5: oracle, word for word at 256^3")
    want = O.matmul_prove(ops[0][0], ops[0][1], labels[0], table.entries, tnv, LR, NQ, key, 256)
    for k in ("roots", "ajtai", "point", "claim", "round_msgs", "challenges", "final_evals", "proof_a", "proof_b",
              "proof_c"):
        assert np.array_equal(alone[0][k], want[k]), k
    for i, ((a, b), lab) in enumerate(zip(ops, labels)):
        say("c5: verifying", dims[i])
        verify_matmul(a, b, lab, alone[i], table.entries, tnv, key)
    aj.free()
