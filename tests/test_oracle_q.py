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


# ---- folding, splitting and ternary commits (ajtai.cuh:700-1280), pinned by the identities the
# reference's own integration tests check (almost-goldilocks-cuda-rs/tests/ajtai_integration.rs) ----
RKEY = bytes((11 * i + 5) % 256 for i in range(32))


def rand_challenge(rng):
    return rng.integers(-1, 3, 64, dtype=np.int8)


def test_fold_homomorphism_and_trivial_challenges():
    """c1 + r c2 = M (z1 + r z2) word for word (ajtai_integration.rs:551-582); r = 0 keeps z1
    and c1; r = 1 adds z2 and c2."""
    rng = np.random.default_rng(61)
    N = 3
    z1 = rng.integers(0, 1 << 64, N, dtype=np.uint64)
    z2 = rng.integers(0, 1 << 64, N, dtype=np.uint64)
    r = rand_challenge(rng)
    c1, c2 = O.ring_commit_binary(RKEY, z1, N, 1)[0], O.ring_commit_binary(RKEY, z2, N, 1)[0]
    w = O.ring_fold_witness(z1, z2, r)
    assert np.array_equal(O.ring_commit_wide(RKEY, w, N).ravel(), O.ring_fold_commitment(c1, c2, r))
    zero, one = np.zeros(64, dtype=np.int8), np.zeros(64, dtype=np.int8)
    one[0] = 1
    bits = lambda z: np.array([(int(z[j]) >> k) & 1 for j in range(N) for k in range(64)], dtype=np.uint64)
    assert np.array_equal(O.ring_fold_witness(z1, z2, zero), bits(z1))
    assert np.array_equal(O.ring_fold_witness(z1, z2, one), bits(z1) + bits(z2))
    assert np.array_equal(O.ring_fold_commitment(c1, c2, zero), c1.ravel())
    assert [int(x) for x in O.ring_fold_commitment(c1, c2, one)] == \
        [(int(a) + int(b)) % Q for a, b in zip(c1.ravel(), c2.ravel())]
    # the commitment fold takes challenge coefficients in [-2, 2] only (ajtai.cuh:808-826);
    # the witness fold takes any int8 (ajtai.cuh:738-748), so -128 on a wrap adds +128
    big = r.copy()
    big[5], big[9] = 3, -128
    cut = big.copy()
    cut[5] = cut[9] = 0
    assert np.array_equal(O.ring_fold_commitment(c1, c2, big), O.ring_fold_commitment(c1, c2, cut))
    e = np.zeros(64, dtype=np.int8)
    e[0] = -128
    top = np.array([1 << 63], dtype=np.uint64)                   # X^63 * (-128) = 128 X^63
    one_bit = np.array([2], dtype=np.uint64)                     # X * (-128 X^63) = +128
    e63 = np.zeros(64, dtype=np.int8)
    e63[63] = -128
    assert int(O.ring_fold_witness(np.zeros(1, np.uint64), top, e)[63]) == Q - 128
    assert int(O.ring_fold_witness(np.zeros(1, np.uint64), one_bit, e63)[0]) == 128
    # X * z: challenge X shifts negacyclically (top coefficient wraps with a sign change)
    x = np.zeros(64, dtype=np.int8)
    x[1] = 1
    shifted = O.ring_fold_witness(np.zeros(N, dtype=np.uint64), z2, x).reshape(N, 64)
    for j in range(N):
        b = [(int(z2[j]) >> k) & 1 for k in range(64)]
        assert [int(v) for v in shifted[j]] == [(Q - b[63]) % Q] + b[:63]


def test_multifold_equals_repeated_folds_and_commitment_homomorphism():
    """multifold of [z1, z2] with [1, r] is fold_witness (as integers); the multifold of
    commitments is the commitment of the multifold witness (ajtai_integration.rs:779-822)."""
    rng = np.random.default_rng(62)
    N, M = 2, 5
    z = rng.integers(0, 1 << 64, M * N, dtype=np.uint64)
    r_all = np.concatenate([np.eye(1, 64, 0, dtype=np.int8)[0]] + [rand_challenge(rng) for _ in range(M - 1)])
    w = O.ring_multifold_mixed_witness(z, None, None, r_all, M, 0, N)
    f = O.ring_fold_witness(z[:N], z[N:2 * N], r_all[64:128])
    w2 = O.ring_multifold_mixed_witness(z[:2 * N], None, None, r_all[:128], 2, 0, N)
    assert [int(v) % Q for v in w2] == [int(v) for v in f]
    cs = np.concatenate([O.ring_commit_binary(RKEY, z[m * N:(m + 1) * N], N, 1).ravel() for m in range(M)])
    wf = np.array([int(v) % Q for v in w], dtype=np.uint64)
    assert np.array_equal(O.ring_commit_wide(RKEY, wf, N).ravel(), O.ring_multifold_commitment(cs, r_all, M))
    assert np.abs(w.astype(np.int64)).max() <= M * 128


def test_split_reconstructs_and_ternary_commit_homomorphism():
    """split: v = sum 2^i (pos_i - neg_i), pos & neg = 0 (ajtai_integration.rs:873-1004); a
    pos-only ternary commit is the binary commit, and sum 2^i c_i commits to v (:1164-1233);
    mixed multifold is binary multifold plus the signed ternary terms (:1361-1458)."""
    rng = np.random.default_rng(63)
    N = 2
    v = rng.integers(-8191, 8192, N * 64).astype(np.int16)
    v[:6] = [0, 1, -1, 8191, -8191, -4096]
    pos, neg = O.ring_split_witness(v, N)
    assert not np.any(pos & neg)
    rec = np.zeros(N * 64, dtype=np.int64)
    for i in range(13):
        for j in range(N):
            for k in range(64):
                rec[j * 64 + k] += (1 << i) * (((int(pos[i * N + j]) >> k) & 1) - ((int(neg[i * N + j]) >> k) & 1))
    assert np.array_equal(rec, v.astype(np.int64))
    c = O.ring_commit_ternary(RKEY, pos, neg, N)
    acc = [0] * (42 * 64)
    for i in range(13):
        acc = [(a + (1 << i) * int(x)) % Q for a, x in zip(acc, c[i].ravel())]
    vf = np.array([int(x) % Q for x in v.astype(np.int64)], dtype=np.uint64)
    assert acc == [int(x) for x in O.ring_commit_wide(RKEY, vf, N).ravel()]
    zero = np.zeros(13 * N, dtype=np.uint64)
    cp = O.ring_commit_ternary(RKEY, pos, zero, N)
    assert np.array_equal(cp, O.ring_commit_binary(RKEY, pos, N, 13))
    # mixed multifold: 2 binary instances + the 13 chunks
    zb = rng.integers(0, 1 << 64, 2 * N, dtype=np.uint64)
    r_all = np.concatenate([rand_challenge(rng) for _ in range(15)])
    mixed = O.ring_multifold_mixed_witness(zb, pos, neg, r_all, 2, 13, N).astype(np.int64)
    plus = O.ring_multifold_mixed_witness(np.concatenate([zb, pos]), None, None, r_all, 15, 0, N).astype(np.int64)
    minus = O.ring_multifold_mixed_witness(np.concatenate([np.zeros(2 * N, dtype=np.uint64), neg]), None, None,
                                           r_all, 15, 0, N).astype(np.int64)
    assert np.array_equal(mixed, plus - minus)


def test_materialized_matrix_and_selective_add():
    m = O.ring_materialize(RKEY, 3).reshape(3, 42, 64)
    for j, i in ((0, 0), (2, 41), (1, 17)):
        assert np.array_equal(m[j, i], O.ring_matrix_entry(RKEY, i, j))
    rng = np.random.default_rng(64)
    total, n_planes = 300, 3
    packed = (total + 63) // 64
    eq = rng.integers(0, Q, 2 * total, dtype=np.uint64)
    planes = rng.integers(0, 1 << 64, n_planes * packed, dtype=np.uint64)
    got = O.q_selective_add_planes(eq, planes, total, n_planes, packed)
    for p in range(n_planes):
        s0 = sum(int(eq[2 * i]) for i in range(total) if (int(planes[p * packed + i // 64]) >> (i % 64)) & 1) % Q
        s1 = sum(int(eq[2 * i + 1]) for i in range(total) if (int(planes[p * packed + i // 64]) >> (i % 64)) & 1) % Q
        assert [int(x) for x in got[p]] == [s0, s1]
