"""GPU parity: every HIP entry point against the oracle or the field definition, bit-exact
on canonical residues, through the C-ABI."""
import ctypes
import sys

import numpy as np
import pytest

import pyoracle as O
from refmath import GOLDEN, P, e2add, e2mul, e2sub, pairs, rand_field

import morpheus_amd as M
from morpheus_amd import DeviceBuffer as DB

pytestmark = pytest.mark.gpu
U = ctypes.POINTER(ctypes.c_uint64)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.goldilocks_cuda_init() == 0, "no usable GPU"
    assert M.lib.monolith_cuda_init() == 0


def up(a):
    return DB.from_numpy(np.asarray(a, dtype=np.uint64))


def ok(rc):
    assert rc == 0, rc


def test_field_batches():
    rng = np.random.default_rng(100)
    n = 3000
    a = rng.integers(0, 1 << 64, n, dtype=np.uint64)
    b = rng.integers(0, 1 << 64, n, dtype=np.uint64)
    a[:4] = [0, P - 1, P, (1 << 64) - 1]
    da, db, dr = up(a), up(b), DB(2 * n)
    A = [int(x) % P for x in a]
    B = [int(x) % P for x in b]
    for name, f in [("gl_batch_add", lambda x, y: (x + y) % P), ("gl_batch_sub", lambda x, y: (x - y) % P),
                    ("gl_batch_mul", lambda x, y: x * y % P)]:
        ok(getattr(M.lib, name)(da.ptr, db.ptr, dr.ptr, n))
        got = dr.to_numpy(n)
        assert [int(x) for x in got] == [f(x, y) for x, y in zip(A, B)], name
    ok(M.lib.gl_batch_inverse(da.ptr, dr.ptr, n))
    assert [int(x) for x in dr.to_numpy(n)] == [pow(x, P - 2, P) if x else 0 for x in A]
    ok(M.lib.gl_batch_exp(da.ptr, 12345, dr.ptr, n))
    assert [int(x) for x in dr.to_numpy(n)] == [pow(x, 12345, P) for x in A]
    # Ext2: n/2 elements
    m = n // 2
    ok(M.lib.ext2_batch_mul_ffi(da.ptr, db.ptr, dr.ptr, m))
    got = pairs(dr.to_numpy(2 * m))
    pa, pb = pairs(np.array(A, dtype=np.uint64)), pairs(np.array(B, dtype=np.uint64))
    assert got == [e2mul(x, y) for x, y in zip(pa, pb)]
    ok(M.lib.ext2_batch_sub_ffi(da.ptr, db.ptr, dr.ptr, m))
    assert pairs(dr.to_numpy(2 * m)) == [e2sub(x, y) for x, y in zip(pa, pb)]
    ok(M.lib.ext2_batch_inverse_ffi(da.ptr, dr.ptr, m))
    inv = pairs(dr.to_numpy(2 * m))
    for x, y in zip(pa[4:], inv[4:]):
        assert e2mul(x, y) == (1, 0)


def test_monolith_and_poseidon2_permutations():
    g = GOLDEN["monolith_w12_permute_0_to_11"]
    inp = np.array(g["input"], dtype=np.uint64)
    out = np.zeros(12, dtype=np.uint64)
    ok(M.lib.monolith_permute_test_ffi(inp.ctypes.data_as(U), out.ctypes.data_as(U)))
    assert [int(x) for x in out] == g["output"]
    edge = np.array([8, P - 8] + [0] * 10, dtype=np.uint64)
    ok(M.lib.monolith_permute_test_ffi(edge.ctypes.data_as(U), out.ctypes.data_as(U)))
    assert np.array_equal(out, O.monolith_permute(edge))
    rng = np.random.default_rng(101)
    st = rng.integers(0, 1 << 64, 12 * 777, dtype=np.uint64)
    d = up(st)
    ok(M.lib.mam_monolith_permute_batch(d.ptr, 777))
    got = d.to_numpy().reshape(-1, 12)
    for i in range(0, 777, 37):
        assert np.array_equal(got[i], O.monolith_permute(st[12 * i:12 * i + 12]))
    s8 = rng.integers(0, 1 << 64, 8 * 300, dtype=np.uint64)
    d8, o8 = up(s8), DB(8 * 300)
    ok(M.lib.poseidon2_hash_batch_ffi(d8.ptr, o8.ptr, 300))
    got = o8.to_numpy().reshape(-1, 8)
    for i in range(0, 300, 13):
        assert np.array_equal(got[i], O.poseidon2_permute(s8[8 * i:8 * i + 8], 0))


@pytest.mark.parametrize("hash_kind", [0, 1])
@pytest.mark.parametrize("log_leaves", [0, 1, 3, 8, 9, 10, 13])
def test_merkle_trees(hash_kind, log_leaves):
    say("merkle", hash_kind, log_leaves)
    nl = 1 << log_leaves
    rng = np.random.default_rng(200 + log_leaves)
    for ext2 in (False, True):
        cw = rng.integers(0, 1 << 64, (4 if ext2 else 2) * nl, dtype=np.uint64)
        d, t = up(cw), DB((2 * nl - 1) * 4)
        f = {(0, False): M.lib.monolith_merkle_tree_gl_ffi, (0, True): M.lib.monolith_merkle_tree_ext2_ffi,
             (1, False): M.lib.poseidon2_merkle_tree_gl_ffi, (1, True): M.lib.poseidon2_merkle_tree_ext2_ffi}
        ok(f[(hash_kind, ext2)](d.ptr, t.ptr, nl))
        want = O.merkle_tree(cw, nl, hash_kind, ext2)
        assert np.array_equal(t.to_numpy(), want)


def test_merkle_tree_large_mixed_kernels():
    # 2^17 leaves: bottom levels with thread-per-permutation kernels, top with lane groups
    nl = 1 << 17
    rng = np.random.default_rng(250)
    cw = rng.integers(0, 1 << 64, 2 * nl, dtype=np.uint64)
    d, t = up(cw), DB((2 * nl - 1) * 4)
    ok(M.lib.monolith_merkle_tree_gl_ffi(d.ptr, t.ptr, nl))
    assert np.array_equal(t.to_numpy(), O.merkle_tree(cw, nl, 0, False))


def test_merkle_specific_codeword():
    cw = np.array(GOLDEN["monolith_specific_codeword"]["codeword"], dtype=np.uint64)
    d, t = up(cw), DB(7 * 4)
    ok(M.lib.monolith_merkle_tree_gl_ffi(d.ptr, t.ptr, 4))
    assert np.array_equal(t.to_numpy(), O.merkle_tree(cw, 4, 0))


@pytest.mark.parametrize("nv", [0, 1, 5, 10, 11, 14])
def test_eq_tables(nv):
    rng = np.random.default_rng(300 + nv)
    r = rand_field(rng, max(2 * nv, 1))
    dr, a, b = up(r), DB(2 << nv), DB(2 << nv)
    res = ctypes.c_void_p()
    ok(M.lib.ext2_eq_dp_all_ffi(dr.ptr, a.ptr, b.ptr, nv, ctypes.byref(res)))
    M.synchronize()
    assert res.value == a.ptr
    assert np.array_equal(a.to_numpy(2 << nv), O.eq_table_ext2(r, nv))
    rb = r[0::2].copy() if nv else r
    drb = up(rb)
    ok(M.lib.eq_dp_all_ffi(drb.ptr, a.ptr, b.ptr, nv, ctypes.byref(res)))
    M.synchronize()
    assert np.array_equal(a.to_numpy(1 << nv), O.eq_table(rb, nv))


@pytest.mark.parametrize("log_n,m", [(4, 1), (6, 3), (12, 12), (13, 5)])
def test_partial_eval(log_n, m):
    rng = np.random.default_rng(400 + log_n)
    base = rand_field(rng, 1 << log_n)
    r2 = rand_field(rng, 2 * m)
    ext_in = np.zeros(2 << log_n, dtype=np.uint64)
    ext_in[0::2] = base
    d_in, d_out, d_r = up(base), DB(2 << log_n), up(r2)
    ok(M.lib.partial_eval_ext2_from_gl_ffi(d_in.ptr, d_out.ptr, d_r.ptr, log_n, m))
    assert np.array_equal(d_out.to_numpy(2 << (log_n - m)), O.partial_eval_ext2(ext_in, log_n, r2, m))
    rb = r2[0::2].copy()
    d_data, d_rb = up(base), up(rb)
    ok(M.lib.partial_eval_gl_ffi(d_data.ptr, d_rb.ptr, log_n, m))
    rbe = np.zeros(2 * m, dtype=np.uint64)
    rbe[0::2] = rb
    want = O.partial_eval_ext2(ext_in, log_n, rbe, m)[0::2]
    assert np.array_equal(d_data.to_numpy(1 << (log_n - m)), want)


@pytest.mark.parametrize("nv,lr", [(1, 1), (4, 1), (10, 1), (12, 2), (13, 1), (16, 3), (17, 1), (20, 1), (22, 1)])
def test_commit_pieces(nv, lr):
    say("commit pieces", nv, lr)
    rng = np.random.default_rng(500 + nv)
    ev = rng.integers(0, 1 << 64, 1 << nv, dtype=np.uint64)
    cw_o, bh_o, _ = O.basefold_commit(ev, nv, lr, with_tree=False)
    d_ev, d_co, d_bh, d_cw = up(ev), DB(1 << nv), DB(1 << nv), DB(1 << (nv + lr))
    ok(M.lib.basefold_bhc_interpolate_ffi(d_ev.ptr, d_co.ptr, d_bh.ptr, nv))
    assert np.array_equal(d_bh.to_numpy(), bh_o)
    ok(M.lib.basefold_encode_ffi(d_co.ptr, d_cw.ptr, nv, lr))
    assert np.array_equal(d_cw.to_numpy(), cw_o)
    # in-place bit reversal round trip
    ok(M.lib.basefold_bit_reverse_gl_ffi(d_bh.ptr, nv))
    M.synchronize()
    assert np.array_equal(d_bh.to_numpy(), np.array([int(x) % P for x in ev], dtype=np.uint64))


# The commitment hashes one copy of its replicated codeword and stores every digest at all
# 2^lr positions: the whole tree is compared with the oracle's full hash of the codeword.
# Sizes reach every tree kernel: lane-parallel only (small), the one-lane leaf and node
# launches with the lane-parallel top (nv 17, 18), and log rates 0 to 3.
@pytest.mark.parametrize("nv,lr,hash_kind", [(1, 1, 0), (2, 1, 0), (5, 2, 1), (9, 1, 0), (12, 3, 0), (14, 1, 1),
                                             (13, 0, 0), (17, 1, 0), (18, 2, 0), (16, 1, 1)])
def test_commit_root(nv, lr, hash_kind):
    say("commit", nv, lr, hash_kind)
    rng = np.random.default_rng(600 + nv)
    ev = rand_field(rng, 1 << nv)
    cw_o, bh_o, tree_o = O.basefold_commit(ev, nv, lr, hash_kind)
    c = M.BasefoldCommitment.commit(ev, nv, lr, hash_kind)
    assert np.array_equal(c.root, tree_o[-4:])
    cw, bh = c.to_host_cache()
    assert np.array_equal(cw, cw_o) and np.array_equal(bh, bh_o)
    _, _, d_tree = c.buffers()
    tree = np.empty(tree_o.size, dtype=np.uint64)
    ok(M.lib.cuda_memcpy_dtoh(tree.ctypes.data, d_tree, tree.size * 8))
    assert np.array_equal(tree, tree_o)


def test_codeword_folds_and_sumcheck_pieces():
    rng = np.random.default_rng(700)
    pairs_n = 1000
    cw = rand_field(rng, 2 * pairs_n)
    cw2 = rand_field(rng, 4 * pairs_n)
    table = O.basefold_table(11, 0, 1, 42)[:2 * pairs_n]
    ch = rand_field(rng, 2)
    d_cw, d_cw2, d_t, d_o = up(cw), up(cw2), up(table), DB(2 * pairs_n)
    ok(M.lib.basefold_fold_mixed_ffi(d_cw.ptr, d_t.ptr, int(ch[0]), int(ch[1]), d_o.ptr, pairs_n))
    got = pairs(d_o.to_numpy())
    c = (int(ch[0]), int(ch[1]))
    for i in range(0, pairs_n, 17):
        v0, v1 = (int(cw[2 * i]), 0), (int(cw[2 * i + 1]), 0)
        x0, w = int(table[2 * i]), int(table[2 * i + 1])
        want = e2add(v0, e2mul(e2mul(e2sub(c, (x0, 0)), e2sub(v1, v0)), (w, 0)))
        assert got[i] == want
    ok(M.lib.basefold_fold_ext2_ffi(d_cw2.ptr, d_t.ptr, int(ch[0]), int(ch[1]), d_o.ptr, pairs_n))
    got = pairs(d_o.to_numpy())
    p2 = pairs(cw2)
    for i in range(0, pairs_n, 19):
        x0, w = int(table[2 * i]), int(table[2 * i + 1])
        want = e2add(p2[2 * i], e2mul(e2mul(e2sub(c, (x0, 0)), e2sub(p2[2 * i + 1], p2[2 * i])), (w, 0)))
        assert got[i] == want
    ok(M.lib.basefold_fold_gl_ffi(d_cw.ptr, d_t.ptr, int(ch[0]), d_o.ptr, pairs_n))
    got = d_o.to_numpy(pairs_n)
    for i in range(0, pairs_n, 23):
        v0, v1, x0, w = int(cw[2 * i]), int(cw[2 * i + 1]), int(table[2 * i]), int(table[2 * i + 1])
        assert int(got[i]) == (v0 + (int(ch[0]) - x0) * (v1 - v0) * w) % P
    # product / fused round / dot: compare the summed partials with the definitions
    eq = rand_field(rng, 8 * pairs_n)
    bh = rand_field(rng, 8 * pairs_n)
    d_eq, d_bh = up(eq), up(bh)
    nb = 7
    pc = [DB(2 * nb) for _ in range(3)]
    ok(M.lib.basefold_sumcheck_product_ext2_ffi(d_eq.ptr, d_bh.ptr, pc[0].ptr, pc[1].ptr, pc[2].ptr, pairs_n, nb))
    sums = []
    for k in range(3):
        s = (0, 0)
        for v in pairs(pc[k].to_numpy()):
            s = e2add(s, v)
        sums.append(s)
    E, B = pairs(eq), pairs(bh)
    c0 = c1 = c2 = (0, 0)
    for i in range(pairs_n):
        e0, e1, b0, b1 = E[2 * i], E[2 * i + 1], B[2 * i], B[2 * i + 1]
        c0 = e2add(c0, e2mul(e0, b0))
        c1 = e2add(c1, e2add(e2mul(e0, b1), e2mul(e1, b0)))
        c2 = e2add(c2, e2mul(e1, b1))
    assert sums == [c0, c1, c2]
    eo, bo = DB(4 * pairs_n), DB(4 * pairs_n)
    ok(M.lib.fused_sumcheck_round_ext2_ffi(d_eq.ptr, d_bh.ptr, int(ch[0]), int(ch[1]), eo.ptr, bo.ptr, pc[0].ptr,
                                           pc[1].ptr, pc[2].ptr, pairs_n, nb))
    # reference semantics: eval (c + ch d) of 4 -> 2, then interp (value, diff)
    ea = e2add(E[0], e2mul(c, E[1]))
    eb = e2add(E[2], e2mul(c, E[3]))
    assert pairs(eo.to_numpy(4))[:2] == [ea, e2sub(eb, ea)]
    dp = DB(2 * nb)
    ok(M.lib.basefold_dot_product_mixed_ffi(d_cw.ptr, d_eq.ptr, dp.ptr, 2 * pairs_n, nb))
    s = (0, 0)
    for v in pairs(dp.to_numpy()):
        s = e2add(s, v)
    want = (0, 0)
    for i in range(2 * pairs_n):
        want = e2add(want, e2mul((int(cw[i]), 0), E[i]))
    assert s == want


@pytest.mark.parametrize("np_,nv,base", [(1, 1, True), (2, 2, True), (2, 9, True), (3, 8, False), (4, 6, True),
                                         (2, 14, False), (8, 5, True), (1, 6, False), (1, 7, True), (2, 3, False),
                                         (3, 1, False), (3, 2, True), (3, 7, True), (3, 11, False), (2, 17, True)])
def test_sumcheck_prove_matches_oracle(np_, nv, base):
    say("sumcheck", np_, nv, base)
    rng = np.random.default_rng(800 + 10 * np_ + nv)
    if base:
        vals = rand_field(rng, np_ << nv)
        ext = np.zeros((np_, 1 << nv, 2), dtype=np.uint64)
        ext[:, :, 0] = vals.reshape(np_, -1)
    else:
        ext = rand_field(rng, np_ * (2 << nv)).reshape(np_, 1 << nv, 2)
        vals = ext.ravel()
    d = up(vals)
    t_gpu, t_cpu = M.Transcript(b"sumcheck"), O.Transcript(b"sumcheck")
    got = M.sumcheck_prove(t_gpu, d, np_, nv, base=base)
    want = O.sumcheck_prove_ext2(t_cpu, ext, np_, nv)
    for k in ("round_msgs", "challenges", "final_evals", "final_eval"):
        assert np.array_equal(got[k], want[k]), k
    assert np.array_equal(t_gpu.state(), t_cpu.state())
    # the input is left untouched
    assert np.array_equal(d.to_numpy(), vals)


def test_sumcheck_reference_ffi_loop_and_state():
    # drive the reference-named round-message / fold entry points like GpuSumcheckStateExt2
    np_, nv = 3, 7
    rng = np.random.default_rng(900)
    ext = rand_field(rng, np_ * (2 << nv)).reshape(np_, 1 << nv, 2)
    want = O.sumcheck_prove_ext2(O.Transcript(b"s"), ext, np_, nv)
    n = 1 << nv
    d, s = up(ext), DB(np_ * 2 * n)
    t = M.Transcript(b"s")
    t.append_u64(b"num_var", nv)
    t.append_u64(b"num_poly", np_)
    nb = 5
    part = DB(nb * (np_ + 1) * 2)
    cur, other = d, s
    for r in range(nv):
        half = (n >> r) // 2
        ok(M.lib.sumcheck_round_message_ext2_ffi(cur.ptr, part.ptr, np_, n, half, nb))
        pp = pairs(part.to_numpy())
        msg = []
        for c in range(np_ + 1):
            acc = (0, 0)
            for b in range(nb):
                acc = e2add(acc, pp[b * (np_ + 1) + c])
            msg.append(acc)
        assert [list(m) for m in msg] == [[int(x) for x in v] for v in want["round_msgs"][r]]
        for m_ in msg:
            t.append_ext2(b"round_message", m_)
        ch = t.challenge_ext2(b"challenge")
        ok(M.lib.sumcheck_fold_ext2_ffi(cur.ptr, other.ptr, int(ch[0]), int(ch[1]), np_, n, half))
        cur, other = other, cur
    fin = cur.to_numpy().reshape(np_, n, 2)[:, 0, :]
    assert np.array_equal(fin, want["final_evals"])
    # stateful interface
    d = up(ext)  # the loop above ping-ponged through the first upload
    st = M.GpuSumcheckStateExt2(d, np_, nv, base=False)
    t2 = M.Transcript(b"s")
    t2.append_u64(b"num_var", nv)
    t2.append_u64(b"num_poly", np_)
    for r in range(nv):
        msg = st.compute_round_message()
        assert np.array_equal(msg, want["round_msgs"][r])
        for m_ in msg:
            t2.append_ext2(b"round_message", m_)
        st.fold(t2.challenge_ext2(b"challenge"))
    assert np.array_equal(st.final_evaluations(), want["final_evals"])
    st.free()


# Poseidon2 trees run every opening round in product form (k_open_u): nv 16 and 17 give the
# eq factor 3 sub-table groups; Monolith at nv 20 runs its product-form rounds with 4 groups.
@pytest.mark.parametrize("nv,lr,nq,hash_kind", [(2, 1, 3, 0), (3, 1, 10, 0), (6, 3, 34, 0), (9, 1, 20, 1),
                                                 (12, 1, 100, 0), (14, 2, 30, 0), (16, 1, 20, 1), (17, 1, 16, 1),
                                                 (20, 1, 8, 0)])
def test_basefold_open_matches_oracle(nv, lr, nq, hash_kind):
    say("open", nv, lr, nq, hash_kind)
    rng = np.random.default_rng(1000 + nv)
    ev = rand_field(rng, 1 << nv)
    table = M.BasefoldTable.generate(nv, lr, nv, 42)
    c = M.BasefoldCommitment.commit(ev, nv, lr, hash_kind)
    pt = rand_field(rng, 2 * nv)
    proof = c.open_ext2(pt, table, M.Transcript(b"open"), nq)
    cw_o, bh_o, tree_o = O.basefold_commit(ev, nv, lr, hash_kind)
    want = O.basefold_open_ext2(cw_o, bh_o, tree_o[-4:], nv, lr, hash_kind, pt, table.entries, nv,
                                O.Transcript(b"open"), nq)
    assert np.array_equal(proof, want)
    assert M.verify_ext2(c.root, pt, nv, lr, hash_kind, proof, nq, table, M.Transcript(b"open"))


def test_basefold_open_with_larger_table():
    # zk-torch-3 shares one table generated for the largest polynomial
    nv, lr, tnv, nq = 6, 1, 9, 12
    rng = np.random.default_rng(1100)
    ev = rand_field(rng, 1 << nv)
    table = M.BasefoldTable.generate(tnv, lr, tnv, 42)
    c = M.BasefoldCommitment.commit(ev, nv, lr, 0)
    pt = rand_field(rng, 2 * nv)
    proof = c.open_ext2(pt, table, M.Transcript(b"t"), nq)
    cw_o, bh_o, tree_o = O.basefold_commit(ev, nv, lr, 0)
    want = O.basefold_open_ext2(cw_o, bh_o, tree_o[-4:], nv, lr, 0, pt, table.entries, tnv, O.Transcript(b"t"), nq)
    assert np.array_equal(proof, want)
