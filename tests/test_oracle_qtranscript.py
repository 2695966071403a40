"""The zk-torch-4 transcript (Monolith-w12 over q) and the q linear sumcheck.

The oracle is pinned by the reference's golden vectors (zk-torch-4/src/crypto/monolith.rs:
373-415: the permutation of [0..11] and RC[1][0..3]; transcript.rs:359-373: the golden
squeeze), by the reference's transcript properties (transcript.rs:232-355), and the sumcheck
by the reference's prove/verify cases (cpu_ext2_prover.rs:120-207) checked with a verifier and
a multilinear evaluation written here from the definitions. The library's host transcript is
then compared with the oracle over random operation sequences."""
import numpy as np
import pytest

import pyoracle as O

import morpheus_amd as M

Q = O.Q
GOLDEN_PERM = [0xed82a66d4453e6b3, 0x752b191c23268621, 0x590fab26245cc4e9, 0xa63b2125e2a94a88,
               0x6ab5bea637404df2, 0x6dbeafbbb35f0b77, 0xfc9cdfae47529f13, 0xe57947afed67d70f,
               0x59da37a113e7987e, 0x95662ed82cec7bb2, 0x7429836455d95e39, 0x91861bd52d8c67f2]


# ---- Ext2 over q (X^2 = 3) on Python integers, written from the definition ----
def add(a, b):
    return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)


def mul(a, b):
    return ((a[0] * b[0] + 3 * a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def inv(a):
    n = (a[0] * a[0] - 3 * a[1] * a[1]) % Q
    ni = pow(n, Q - 2, Q)
    return (a[0] * ni % Q, (-a[1]) * ni % Q)


def lagrange(evals, x):
    res = (0, 0)
    for i in range(len(evals)):
        b = (1, 0)
        for j in range(len(evals)):
            if j != i:
                b = mul(b, mul(((x[0] - j) % Q, x[1]), inv(((i - j) % Q, 0))))
        res = add(res, mul(b, evals[i]))
    return res


def mle(evals, point):
    """sum_x p(x) prod_k (r_k if bit k of x else 1 - r_k): variable k = bit k (LSB first)."""
    total = (0, 0)
    for x, v in enumerate(evals):
        w = (1, 0)
        for k, r in enumerate(point):
            w = mul(w, r if (x >> k) & 1 else ((1 - r[0]) % Q, (-r[1]) % Q))
        total = add(total, mul(v, w))
    return total


def verify(label, proof, claim, nv, np_):
    """SumcheckVerifier::verify (verifier.rs:15-86) on the oracle transcript."""
    msgs, chs, finals, fe = proof
    t = O.QTranscript(label)
    t.append_u64s(b"num_var", [nv])
    t.append_u64s(b"num_poly", [np_])
    cur, got = claim, []
    for r in range(nv):
        ev = [(int(a), int(b)) for a, b in msgs[r]]
        if add(ev[0], ev[1]) != cur:
            return False, got
        for e in msgs[r]:
            t.append_ext2(b"round_message", e)
        ch = t.challenge(b"challenge", 2)
        got.append((int(ch[0]), int(ch[1])))
        cur = lagrange(ev, got[-1])
    return cur == (int(fe[0]), int(fe[1])), got


# ---- the permutation ----
def test_round_constants_golden_and_distinct():
    rc = O.MONOQ_RC.reshape(5, 12)
    assert [int(v) for v in rc[0, :4]] == [0xdc5869754490c2c9, 0x8df3ae2a26b03334, 0x6531a374c7fc9b9b,
                                           0x0758fb4f5d864daf]
    assert len({int(v) for v in rc.ravel()}) == 60 and int(rc.max()) < Q


def test_permutation_golden_and_library_twin():
    assert [int(v) for v in O.monoq_permute(np.arange(12))] == GOLDEN_PERM
    assert [int(v) for v in M.monolith_q_permute(np.arange(12))] == GOLDEN_PERM
    rng = np.random.default_rng(3)
    for _ in range(20):
        s = rng.integers(0, 1 << 64, 12, dtype=np.uint64)       # non-canonical words included
        assert np.array_equal(O.monoq_permute(s), M.monolith_q_permute(s))


def test_permutation_diffuses_and_is_canonical():
    a = np.array([(i * 0x123456789ABCDEF) % (1 << 64) for i in range(12)], dtype=np.uint64)
    b = a.copy()
    b[0] += np.uint64(1)
    pa, pb = O.monoq_permute(a), O.monoq_permute(b)
    assert all(pa[i] != pb[i] for i in range(12))
    assert int(pa.max()) < Q


# ---- the transcript ----
def squeeze_golden(cls):
    t = cls(b"zk-torch-4-test")
    t.append_u64s(b"a", [1]) if cls is O.QTranscript else t.append_scalar(b"a", 1)
    t.append_u64s(b"b", [2]) if cls is O.QTranscript else t.append_scalar(b"b", 2)
    t.append_u64s(b"c", [0xDEADBEEF]) if cls is O.QTranscript else t.append_u64(b"c", 0xDEADBEEF)
    return int(t.challenge(b"out", 1)[0]) if cls is O.QTranscript else t.challenge_scalar(b"out")


def test_golden_squeeze():
    assert squeeze_golden(O.QTranscript) == 0x5df33cda688d0bce
    assert squeeze_golden(M.QTranscript) == 0x5df33cda688d0bce


class _Lib:
    """the library transcript behind the oracle's method names"""

    def __init__(self, label=b""):
        self.t = M.QTranscript(label)

    def append_u64s(self, label, v):
        self.t.append_scalars(label, v)

    def append_ext2(self, label, v):
        self.t.append_ext2(label, v)

    def challenge(self, label, n):
        return self.t.challenge_vector(label, n)

    def fork(self, fam, i):
        c = _Lib.__new__(_Lib)
        c.t = self.t.fork(fam, i)
        return c

    def clone(self):
        c = _Lib.__new__(_Lib)
        c.t = self.t.clone()
        return c

    def state(self):
        return self.t.get_state()


@pytest.mark.parametrize("cls", [O.QTranscript, _Lib])
def test_reference_transcript_properties(cls):
    """transcript.rs:232-355: determinism, input and label sensitivity, the label length
    prefix, fresh squeezes, duplexing boundary, absorb invalidation, forks by id and family."""
    def c(t, label=b"c"):
        return int(t.challenge(label, 1)[0])
    t1, t2 = cls(b"test"), cls(b"test")
    t1.append_u64s(b"val", [42])
    t2.append_u64s(b"val", [42])
    assert c(t1) == c(t2)
    t1, t2 = cls(b"test"), cls(b"test")
    t1.append_u64s(b"val", [42])
    t2.append_u64s(b"val", [43])
    assert c(t1) != c(t2)
    assert c(cls(b"a")) != c(cls(b"b"))
    t1, t2 = cls(b"x"), cls(b"x")
    t1.append_u64s(b"ab", [0])
    t1.append_u64s(b"c", [0])
    t2.append_u64s(b"a", [0])
    t2.append_u64s(b"bc", [0])
    assert c(t1, b"") != c(t2, b"")
    t = cls(b"test")
    assert c(t, b"x") != c(t, b"x")
    vals = cls(b"test").challenge(b"x", 24)
    assert len({int(v) for v in vals}) == 24
    t1 = cls(b"test")
    c(t1, b"a")
    t1.append_u64s(b"b", [7])
    t3 = cls(b"test")
    c(t3, b"a")
    assert c(t1) != c(t3)
    root = cls(b"root")
    assert c(root.fork(b"fam", 0)) != c(root.fork(b"fam", 1))
    assert c(root.fork(b"fam", 0)) == c(root.fork(b"fam", 0))
    assert c(root.fork(b"dag_partition", 3)) != c(root.fork(b"open_reducer", 3))
    assert c(root.fork(b"ft_bucket", 0)) != c(root.fork(b"ft_bucket0", 0))
    e = cls(b"test").challenge(b"x", 2)
    assert e[0] != e[1]


def test_library_transcript_matches_oracle():
    rng = np.random.default_rng(11)
    for trial in range(30):
        label = bytes(rng.integers(0, 256, int(rng.integers(0, 20)), dtype=np.uint8))
        a, b = O.QTranscript(label), M.QTranscript(label)
        for _ in range(int(rng.integers(1, 25))):
            op = int(rng.integers(0, 4))
            lab = bytes(rng.integers(0, 256, int(rng.integers(0, 12)), dtype=np.uint8))
            if op == 0:
                v = rng.integers(0, 1 << 64, int(rng.integers(0, 30)), dtype=np.uint64)
                a.append_u64s(lab, v)
                b.append_scalars(lab, v)
            elif op == 1:
                v = rng.integers(0, 1 << 64, 2, dtype=np.uint64)
                a.append_ext2(lab, v)
                b.append_ext2(lab, v)
            elif op == 2:
                n = int(rng.integers(1, 20))
                assert np.array_equal(a.challenge(lab, n), b.challenge_vector(lab, n))
            else:
                i = int(rng.integers(0, 1000))
                a, b = a.fork(lab, i), b.fork(lab, i)
            assert np.array_equal(a.state(), b.get_state())
        h = (int(b.t.in_count) * 100 + int(b.t.out_count)) % (1 << 64)
        for v in b.get_state():
            h = (h * 0x9e3779b97f4a7c15 + int(v)) % (1 << 64)
        assert b.fingerprint() == h


# ---- the sumcheck ----
def lift(vals):
    return np.array(vals, dtype=np.uint64)


@pytest.mark.parametrize("label,polys,claim", [
    (b"cpu-ext2-roundtrip", [list(range(1, 5)), list(range(5, 9))], 70),
    (b"three-poly", [[1, 2, 3, 4], [2, 3, 5, 7], [11, 13, 17, 19]], 887),
    (b"larger", [[i * 3 + 1 for i in range(16)], [i * 5 + 7 for i in range(16)]],
     sum((i * 3 + 1) * (i * 5 + 7) for i in range(16))),
])
def test_reference_sumcheck_cases(label, polys, claim):
    """cpu_ext2_prover.rs:120-207: the proofs verify, the verifier's challenges are the
    prover's, the first message sums to the claim, a tampered message fails."""
    np_, nv = len(polys), len(polys[0]).bit_length() - 1
    t = O.QTranscript(label)
    proof = O.q_sumcheck_prove(t, lift(sum(polys, [])), np_, nv, False)
    ok, chs = verify(label, proof, (claim % Q, 0), nv, np_)
    assert ok
    assert chs == [(int(a), int(b)) for a, b in proof[1]]
    assert add(tuple(int(x) for x in proof[0][0][0]), tuple(int(x) for x in proof[0][0][1])) == (claim % Q, 0)
    bad = (proof[0].copy(),) + proof[1:]
    bad[0][0, 0, 0] = (int(bad[0][0, 0, 0]) + 1) % Q
    assert not verify(label, bad, (claim % Q, 0), nv, np_)[0]


@pytest.mark.parametrize("np_,nv,ext", [(1, 1, True), (3, 5, True), (2, 4, False), (4, 3, True)])
def test_sumcheck_final_values_are_the_multilinear_evaluations(np_, nv, ext):
    rng = np.random.default_rng(100 * np_ + nv)
    n = 1 << nv
    raw = rng.integers(0, 1 << 64, np_ * n * (2 if ext else 1), dtype=np.uint64)
    t = O.QTranscript(b"mle")
    msgs, chs, finals, fe = O.q_sumcheck_prove(t, raw, np_, nv, ext)
    point = [(int(a), int(b)) for a, b in chs]
    prod = (1, 0)
    claim = (0, 0)
    tables = []
    for i in range(np_):
        if ext:
            ev = [(int(raw[2 * (i * n + x)]) % Q, int(raw[2 * (i * n + x) + 1]) % Q) for x in range(n)]
        else:
            ev = [(int(raw[i * n + x]) % Q, 0) for x in range(n)]
        tables.append(ev)
        v = mle(ev, point)
        assert (int(finals[i][0]), int(finals[i][1])) == v
        prod = mul(prod, v)
    assert (int(fe[0]), int(fe[1])) == prod
    for x in range(n):
        term = (1, 0)
        for ev in tables:
            term = mul(term, ev[x])
        claim = add(claim, term)
    assert verify(b"mle", (msgs, chs, finals, fe), claim, nv, np_)[0]
    # the library's host verifier accepts the same proof with the same challenges
    ok, lchs = M.q_sumcheck_verify(M.QTranscript(b"mle"), claim, np_, nv, msgs, fe)
    assert ok and [(int(a), int(b)) for a, b in lchs] == point
