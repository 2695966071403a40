"""Pins of the almost-Goldilocks oracle (oracle/q.c): the reference's golden vectors,
field identities against Python integers, closed forms for eq / partial evaluation /
sumcheck, and the ring-Ajtai identities (negacyclic shifts, linearity, agreement of the
binary, wide and sparse forms, column offsets, the ChaCha8 matrix stream)."""
import numpy as np
import pytest

import pyoracle as O

Q = O.Q
C = (1 << 32) + 31


def qm(a, b):
    return int(O.L.or_q_mul(a, b))


def q2m(a, b):
    out = np.zeros(2, dtype=np.uint64)
    O.L.or_q2_mul(O._p(O.arr(a)), O._p(O.arr(b)), O._p(out))
    return [int(x) for x in out]


def test_reference_golden_vectors():
    # cuda_almost_goldilocks/almost_field_test.cu:143-180
    assert Q == 0xFFFFFFFEFFFFFFE1 and Q + C == 1 << 64
    half = 0x7FFFFFFF7FFFFFF1
    assert (2 * half) % Q == 1
    for a, b, want in [(0, 0, 0), (1, 2, 3), (Q - 1, 1, 0), (Q - 5, 7, 2), ((1 << 64) - 1, 1, C)]:
        assert O.L.or_q_add(a, b) == want
    for a, b, want in [(0, 12345, 0), (1, 12345, 12345), (12345, 1, 12345), (1 << 32, 1 << 32, C),
                       (Q - 1, Q - 1, 1), (2, half, 1)]:
        assert qm(a, b) == want
    # almost_extension_test.cu:107-126
    assert q2m([1, 2], [3, 4]) == [27, 10]
    assert q2m([2, 3], [2, 3]) == [31, 12]
    assert q2m([0, 1], [0, 1]) == [3, 0]


def test_field_against_python_ints():
    rng = np.random.default_rng(1)
    vals = [0, 1, 2, Q - 1, Q, Q + 1, (1 << 64) - 1, 1 << 63, C] + [int(x) for x in rng.integers(0, 1 << 64, 40,
                                                                                               dtype=np.uint64)]
    for a in vals:
        for b in vals[:12]:
            assert O.L.or_q_add(a, b) == (a + b) % Q
            assert O.L.or_q_sub(a, b) == (a - b) % Q
            assert qm(a, b) == (a * b) % Q
        if a % Q:
            assert qm(a, O.L.or_q_inv(a)) == 1
    # 3 is a quadratic non-residue mod q, so X^2 - 3 is irreducible
    assert pow(3, (Q - 1) // 2, Q) == Q - 1
    for _ in range(20):
        a = [int(x) % Q for x in rng.integers(0, 1 << 64, 2, dtype=np.uint64)]
        inv = np.zeros(2, dtype=np.uint64)
        O.L.or_q2_inv(O._p(O.arr(a)), O._p(inv))
        assert q2m(a, inv) == [1, 0]


def test_eq_table_closed_form():
    rng = np.random.default_rng(2)
    for nv in (0, 1, 3, 6):
        r = [int(x) % Q for x in rng.integers(0, 1 << 64, 2 * nv, dtype=np.uint64)]
        t = O.q_eq_table(r, nv, ext=True).reshape(-1, 2)
        # sum over the cube is 1; each entry is the product of its factors
        s0 = sum(int(v[0]) for v in t) % Q
        s1 = sum(int(v[1]) for v in t) % Q
        assert (s0, s1) == (1, 0)
        for x in range(1 << nv):
            acc = [1, 0]
            for i in range(nv):
                f = [r[2 * i], r[2 * i + 1]] if (x >> i) & 1 else [(1 - r[2 * i]) % Q, (-r[2 * i + 1]) % Q]
                acc = [(acc[0] * f[0] + 3 * acc[1] * f[1]) % Q, (acc[0] * f[1] + acc[1] * f[0]) % Q]
            assert [int(t[x][0]), int(t[x][1])] == acc


def test_partial_eval_is_mle_evaluation():
    rng = np.random.default_rng(3)
    nv = 5
    f = rng.integers(0, 1 << 64, 1 << nv, dtype=np.uint64)
    r = rng.integers(0, 1 << 64, 2 * nv, dtype=np.uint64)
    full = O.q_partial_eval(f, 0, nv, r, 1, nv)
    eq = O.q_eq_table([int(x) % Q for x in r], nv, ext=True).reshape(-1, 2)
    s = [0, 0]
    for x in range(1 << nv):
        fx = int(f[x]) % Q
        s = [(s[0] + fx * int(eq[x][0])) % Q, (s[1] + fx * int(eq[x][1])) % Q]
    assert [int(full[0]), int(full[1])] == s


def test_sumcheck_message_and_fold_consistency():
    rng = np.random.default_rng(4)
    d, n = 3, 16
    polys = rng.integers(0, 1 << 64, 2 * d * n, dtype=np.uint64)
    msg = O.q_sumcheck_msg(1, polys, d, n, n // 2).reshape(-1, 2)
    # g(0) + g(1) = sum over all x of prod_i p_i(x)
    tot = [0, 0]
    for x in range(n):
        prod = [1, 0]
        for i in range(d):
            v = [int(polys[2 * (i * n + x)]) % Q, int(polys[2 * (i * n + x) + 1]) % Q]
            prod = [(prod[0] * v[0] + 3 * prod[1] * v[1]) % Q, (prod[0] * v[1] + prod[1] * v[0]) % Q]
        tot = [(tot[0] + prod[0]) % Q, (tot[1] + prod[1]) % Q]
    assert [(int(msg[0][k]) + int(msg[1][k])) % Q for k in range(2)] == tot
    # folding at c reproduces the message value at c (c = 2)
    folded = O.q_sumcheck_fold(1, polys, [2, 0], d, n, n // 2)
    g2 = [0, 0]
    for y in range(n // 2):
        prod = [1, 0]
        for i in range(d):
            v = [int(folded[2 * (i * n + y)]), int(folded[2 * (i * n + y) + 1])]
            prod = [(prod[0] * v[0] + 3 * prod[1] * v[1]) % Q, (prod[0] * v[1] + prod[1] * v[0]) % Q]
        g2 = [(g2[0] + prod[0]) % Q, (g2[1] + prod[1]) % Q]
    assert [int(msg[2][0]), int(msg[2][1])] == g2


KEY = bytes(range(32))


def test_ring_matrix_is_the_chacha8_stream():
    kw = O.ring_key_words(KEY)
    for row, col in [(0, 0), (41, 7), (5, (1 << 32) + 3)]:
        m = O.ring_matrix_entry(KEY, row, col)
        assert all(int(v) < Q for v in m)
        for blk in range(8):
            out = np.zeros(16, dtype=np.uint32)
            nonce = np.array([row, col & 0xFFFFFFFF, col >> 32], dtype=np.uint32)
            O.L.or_chacha_block(8, O._u32p(kw), blk, O._u32p(nonce), O._u32p(out))
            words = [(int(out[2 * k + 1]) << 32) | int(out[2 * k]) for k in range(8)]
            if all(w < Q for w in words):  # no rejection in this block
                assert [int(v) for v in m[8 * blk:8 * blk + 8]] == words


# (row, column, block) whose first ChaCha8 block holds a word >= q for REJ_KEY: the only
# matrix entries on which the rejection path runs (a search of the first 8 M columns).
REJ_KEY = bytes((11 * i + 5) % 256 for i in range(32))
REJECTS = [(40, 2108465, 3), (31, 2660596, 5), (21, 6196264, 3), (0, 7120001, 0)]


def stream_words(kw, row, col, counter):
    out = np.zeros(16, dtype=np.uint32)
    nonce = np.array([row, col & 0xFFFFFFFF, col >> 32], dtype=np.uint32)
    O.L.or_chacha_block(8, O._u32p(kw), counter, O._u32p(nonce), O._u32p(out))
    return [(int(out[2 * k + 1]) << 32) | int(out[2 * k]) for k in range(8)]


def test_ring_matrix_rejection_path():
    """A rejected word is skipped and the block's coefficients continue with the accepted
    words of the blocks at counters blk + 8, blk + 16, ... (ajtai_chacha8.cuh:115-178)."""
    kw = O.ring_key_words(REJ_KEY)
    for row, col, blk in REJECTS:
        words = stream_words(kw, row, col, blk)
        assert any(w >= Q for w in words), (row, col, blk)
        accepted = [w for w in words + stream_words(kw, row, col, blk + 8) if w < Q][:8]
        m = O.ring_matrix_entry(REJ_KEY, row, col)
        assert [int(v) for v in m[8 * blk:8 * blk + 8]] == accepted
        assert all(int(v) < Q for v in m)
        for b in range(8):
            if b != blk:
                assert [int(v) for v in m[8 * b:8 * b + 8]] == stream_words(kw, row, col, b)


def negacyclic_shift(a, l):
    out = [0] * 64
    for r in range(64):
        out[r] = int(a[r - l]) if r >= l else (-int(a[r - l + 64])) % Q
    return out


def test_ring_commit_single_bits_are_negacyclic_shifts():
    for j, l in [(0, 0), (1, 5), (2, 63)]:
        z = np.zeros(3, dtype=np.uint64)
        z[j] = np.uint64(1) << np.uint64(l)
        c = O.ring_commit_binary(KEY, z, 3, 1)[0]
        for i in (0, 17, 41):
            assert [int(v) for v in c[i]] == negacyclic_shift(O.ring_matrix_entry(KEY, i, j), l)


def test_ring_commit_forms_agree_and_are_linear():
    rng = np.random.default_rng(5)
    N = 4
    z1 = rng.integers(0, 1 << 64, N, dtype=np.uint64)
    z2 = rng.integers(0, 1 << 64, N, dtype=np.uint64) & ~z1  # disjoint bits
    c1, c2 = O.ring_commit_binary(KEY, z1, N, 1)[0], O.ring_commit_binary(KEY, z2, N, 1)[0]
    c12 = O.ring_commit_binary(KEY, z1 | z2, N, 1)[0]
    assert np.array_equal(c12, (c1.astype(object) + c2.astype(object)) % Q)
    # wide form with 0/1 coefficients, sparse form with the set positions
    wide = np.array([(int(z1[j]) >> l) & 1 for j in range(N) for l in range(64)], dtype=np.uint64)
    assert np.array_equal(O.ring_commit_wide(KEY, wide, N), c1)
    pos = [64 * j + l for j in range(N) for l in range(64) if (int(z1[j]) >> l) & 1]
    rng.shuffle(pos)
    assert np.array_equal(O.ring_commit_sparse(KEY, np.array(pos, dtype=np.uint64)), c1)
    # batched = per-witness; col_offset = zero columns in front
    cb = O.ring_commit_binary(KEY, np.concatenate([z1, z2]), N, 2)
    assert np.array_equal(cb[0], c1) and np.array_equal(cb[1], c2)
    padded = np.concatenate([np.zeros(3, dtype=np.uint64), z1])
    assert np.array_equal(O.ring_commit_binary(KEY, z1, N, 1, col_offset=3)[0],
                          O.ring_commit_binary(KEY, padded, N + 3, 1)[0])


def test_fused_permute_peval_reduces_to_partial_eval():
    # identity permutation: the fused kernel's oracle equals the pinned partial evaluation
    rng = np.random.default_rng(6)
    n, m = 7, 3
    f = rng.integers(0, 1 << 64, 1 << n, dtype=np.uint64)
    r = rng.integers(0, 1 << 64, 2 * m, dtype=np.uint64)
    eq = O.q_eq_table(r, m, ext=True)
    lo, hi = O.split_luts(n, list(range(n)), n // 2)
    got = O.fused_permute_peval(Q, f, eq, lo, hi, n, m, n // 2, 1 << (n - m))
    assert np.array_equal(got, O.q_partial_eval(f, 0, n, r, 1, m))
    # a permutation moves variables: folding permuted variables = folding a permuted table
    perm = [4, 0, 6, 1, 3, 5, 2]  # bit p of the new index is bit perm[p] of the stored index
    lo, hi = O.split_luts(n, perm, 3)
    permuted = np.array([f[sum(((x >> p) & 1) << perm[p] for p in range(n))] for x in range(1 << n)],
                        dtype=np.uint64)
    got = O.fused_permute_peval(Q, f, eq, lo, hi, n, m, 3, 1 << (n - m))
    assert np.array_equal(got, O.q_partial_eval(permuted, 0, n, r, 1, m))
