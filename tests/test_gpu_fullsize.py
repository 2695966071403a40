"""Parity at the benchmark sizes, through the C-ABI.

- Product sumcheck, 2 base polynomials of 2^20: the whole proof equals the oracle's.
- Basefold commit + open, 2^22 evaluations, rate 1/2, Monolith, 100 queries (oracle open
  at this size takes ~40 s, so the checks are size-independent properties): the oracle's
  verifier accepts the proof on a replayed transcript and rejects it after one changed word
  in an oracle, a root and a query path; the claimed evaluation equals the oracle's MLE of
  the evaluations at the point; bh is the bit-reversed evaluations; the codeword matches
  the closed form codeword[s 2^nv + u] = MLE(evals) at y_{nv-1-i} = (-1)^(bit i of u) at
  sampled positions. A 2^18 opening is compared word for word with the oracle's.
- int8 Ajtai, witness 2^22, 256 rows: equal to the oracle's exact matvec.
- Ring Ajtai, 2^22 binary coefficients (2^16 ring columns): additive over disjoint
  witnesses, and a single set bit in the last column equals the oracle's X^l M_{i,j}.
"""
import sys

import numpy as np
import pytest

import pyoracle as O
from refmath import P, rand_field

import morpheus_amd as M
from morpheus_amd import DeviceBuffer as DB
from morpheus_amd import ring_ajtai as R
from morpheus_amd.ajtai import AjtaiI8, key_from_seed

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.goldilocks_cuda_init() == 0
    assert M.lib.monolith_cuda_init() == 0
    assert M.lib.almost_goldilocks_cuda_init() == 0


def test_sumcheck_2p20_equals_oracle():
    np_, nv = 2, 20
    say("sumcheck", np_, nv)
    rng = np.random.default_rng(2020)
    vals = rand_field(rng, np_ << nv)
    ext = np.zeros((np_, 1 << nv, 2), dtype=np.uint64)
    ext[:, :, 0] = vals.reshape(np_, -1)
    d = DB.from_numpy(vals)
    t_gpu, t_cpu = M.Transcript(b"sumcheck-c1"), O.Transcript(b"sumcheck-c1")
    got = M.sumcheck_prove(t_gpu, d, np_, nv, base=True)
    want = O.sumcheck_prove_ext2(t_cpu, ext, np_, nv)
    for k in ("round_msgs", "challenges", "final_evals", "final_eval"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(t_gpu.state(), t_cpu.state())
    d.free()


def lift(ev):
    e = np.zeros((ev.size, 2), dtype=np.uint64)
    e[:, 0] = ev
    return e


def mle(ev, point2):
    nv = point2.size // 2
    return O.partial_eval_ext2(lift(ev), nv, point2, nv).reshape(-1)[:2]


def bitrev_perm(nv):
    idx = np.arange(1 << nv, dtype=np.int64)
    rev = np.zeros_like(idx)
    for i in range(nv):
        rev |= ((idx >> i) & 1) << (nv - 1 - i)
    return rev


def test_basefold_2p22_properties():
    nv, lr, nq, hk = 22, 1, 100, 0
    say("basefold", nv, lr, nq)
    rng = np.random.default_rng(2222)
    ev = rand_field(rng, 1 << nv)
    table = M.BasefoldTable.generate(nv, lr, nv, 42)
    c = M.BasefoldCommitment.commit(ev, nv, lr, hk)
    pt = rand_field(rng, 2 * nv)
    proof = c.open_ext2(pt, table, M.Transcript(b"basefold-c2"), nq)

    def verify(pf):
        return O.basefold_verify_ext2(c.root, pt, nv, lr, hk, pf, nq, table.entries, nv, O.Transcript(b"basefold-c2"))

    assert verify(proof)
    assert np.array_equal(proof[:2], mle(ev, pt))
    # one changed word in: the first oracle, the first folded root, the last query path
    for pos in (2, 2 + 6 * nv, proof.size - 1):
        bad = proof.copy()
        bad[pos] ^= np.uint64(1 << 7)
        assert not verify(bad), pos
    cw, bh = c.to_host_cache()
    assert np.array_equal(bh, ev[bitrev_perm(nv)])
    for u in (0, 1, (1 << nv) - 1, int(rng.integers(0, 1 << nv))):
        y = np.zeros(2 * nv, dtype=np.uint64)
        for i in range(nv):
            y[2 * (nv - 1 - i)] = (P - 1) if (u >> i) & 1 else 1
        want = int(mle(ev, y)[0])
        for s in range(1 << lr):
            assert int(cw[s * (1 << nv) + u]) == want, (u, s)
    c.free()


def test_basefold_2p18_equals_oracle():
    nv, lr, nq, hk = 18, 1, 100, 0
    say("basefold", nv, lr, nq)
    rng = np.random.default_rng(1818)
    ev = rand_field(rng, 1 << nv)
    table = M.BasefoldTable.generate(nv, lr, nv, 42)
    c = M.BasefoldCommitment.commit(ev, nv, lr, hk)
    pt = rand_field(rng, 2 * nv)
    proof = c.open_ext2(pt, table, M.Transcript(b"open"), nq)
    cw_o, bh_o, tree_o = O.basefold_commit(ev, nv, lr, hk)
    assert np.array_equal(c.root, tree_o[-4:])
    want = O.basefold_open_ext2(cw_o, bh_o, tree_o[-4:], nv, lr, hk, pt, table.entries, nv, O.Transcript(b"open"), nq)
    assert np.array_equal(proof, want)
    c.free()


def test_ajtai_i8_2p22_equals_oracle():
    rows, n = 256, 1 << 22
    say("ajtai", rows, n)
    key = key_from_seed(2022)
    rng = np.random.default_rng(4)
    z = rng.integers(0, 1 << 64, n, dtype=np.uint64)
    A = AjtaiI8(key, rows, n)
    got = A.commit(z, 1)
    want = O.ajtai_i8_commit(key, rows, n, z, 1).reshape(1, rows)
    assert np.array_equal(got, want)
    A.free()


KEY = bytes((11 * i + 5) % 256 for i in range(32))


def test_ring_commit_2p22_additive_and_last_column():
    N = 1 << 16  # 2^22 binary coefficients
    say("ring", N)
    seed = R.Seed.from_bytes(KEY)
    rng = np.random.default_rng(16)
    z = rng.integers(0, 1 << 64, N, dtype=np.uint64)
    mask = rng.integers(0, 1 << 64, N, dtype=np.uint64)
    z1, z2 = z & mask, z & ~mask
    c, c1, c2 = (R.commit(seed, w).rows for w in (z, z1, z2))
    s = (c1.astype(object) + c2.astype(object)) % O.Q
    assert np.array_equal(c.astype(object), s)
    last = np.zeros(N, dtype=np.uint64)
    last[N - 1] = np.uint64(1 << 63)
    assert np.array_equal(R.commit(seed, last).rows, O.ring_commit_sparse(KEY, np.array([64 * N - 1], dtype=np.uint64)))
