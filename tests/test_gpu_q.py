"""GPU parity of the almost-Goldilocks entry points (morpheus_agl.h) against the oracle:
field and Ext2 batch ops, eq tables (plain, stream, batched), partial evaluation,
sumcheck round messages and folds (plain and batched), and the ring Ajtai commitments."""
import ctypes
import gc

import numpy as np
import pytest

import pyoracle as O

import morpheus_amd as M
from morpheus_amd import ring_ajtai as R

pytestmark = pytest.mark.gpu
Q = O.Q
lib = M.lib


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert lib.almost_goldilocks_cuda_init() == 0


def dev(a):
    return M.DeviceBuffer.from_numpy(np.asarray(a, dtype=np.uint64))


def words(rng, n):
    """uniform words incl. non-canonical ones and edge values"""
    x = rng.integers(0, 1 << 64, n, dtype=np.uint64)
    edges = np.array([0, 1, Q - 1, Q, Q + 1, (1 << 64) - 1], dtype=np.uint64)
    x[:min(n, 6)] = edges[:min(n, 6)]
    return x


def to_int(a):
    return [int(v) for v in a]


def test_base_batch_ops():
    print("q base batch ops", flush=True)
    rng = np.random.default_rng(1)
    n = 1000
    a, b = words(rng, n), words(rng, n)[::-1].copy()
    da, db, dr = dev(a), dev(b), M.DeviceBuffer(n)
    A, B = [int(v) % Q for v in a], [int(v) % Q for v in b]
    for name, f in [("add", lambda x, y: (x + y) % Q), ("sub", lambda x, y: (x - y) % Q),
                    ("mul", lambda x, y: x * y % Q), ("div", lambda x, y: x * pow(y, Q - 2, Q) % Q)]:
        assert getattr(lib, f"agl_batch_{name}_ffi")(da.ptr, db.ptr, dr.ptr, n) == 0
        assert to_int(dr.to_numpy()) == [f(x, y) for x, y in zip(A, B)], name
    for name, f in [("neg", lambda x: (-x) % Q), ("square", lambda x: x * x % Q), ("double", lambda x: 2 * x % Q),
                    ("inverse", lambda x: pow(x, Q - 2, Q))]:
        assert getattr(lib, f"agl_batch_{name}_ffi")(da.ptr, dr.ptr, n) == 0
        assert to_int(dr.to_numpy()) == [f(x) for x in A], name
    assert lib.agl_batch_exp_ffi(da.ptr, 12345, dr.ptr, n) == 0
    assert to_int(dr.to_numpy()) == [pow(x, 12345, Q) for x in A]
    assert lib.agl_batch_mul_scalar_ffi(Q + 7, da.ptr, dr.ptr, n) == 0
    assert to_int(dr.to_numpy()) == [7 * x % Q for x in A]


def q2mul(a, b):
    return ((a[0] * b[0] + 3 * a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def test_ext2_batch_ops():
    print("q ext2 batch ops", flush=True)
    rng = np.random.default_rng(2)
    n = 500
    a, b = words(rng, 2 * n), words(rng, 2 * n)[::-1].copy()
    da, db, dr = dev(a), dev(b), M.DeviceBuffer(2 * n)
    A = [(int(a[2 * i]) % Q, int(a[2 * i + 1]) % Q) for i in range(n)]
    Bv = [(int(b[2 * i]) % Q, int(b[2 * i + 1]) % Q) for i in range(n)]

    def got():
        r = dr.to_numpy()
        return [(int(r[2 * i]), int(r[2 * i + 1])) for i in range(n)]

    assert lib.aext2_batch_mul_ffi(da.ptr, db.ptr, dr.ptr, n) == 0
    assert got() == [q2mul(x, y) for x, y in zip(A, Bv)]
    assert lib.aext2_batch_sub_ffi(da.ptr, db.ptr, dr.ptr, n) == 0
    assert got() == [((x[0] - y[0]) % Q, (x[1] - y[1]) % Q) for x, y in zip(A, Bv)]
    assert lib.aext2_batch_square_ffi(da.ptr, dr.ptr, n) == 0
    assert got() == [q2mul(x, x) for x in A]
    assert lib.aext2_batch_frobenius_ffi(da.ptr, dr.ptr, n) == 0
    assert got() == [(x[0], (-x[1]) % Q) for x in A]
    assert lib.aext2_batch_inverse_ffi(da.ptr, dr.ptr, n) == 0
    for x, y in zip(A, got()):
        assert q2mul(x, y) == ((1, 0) if x != (0, 0) else (0, 0))
    assert lib.aext2_batch_exp_ffi(da.ptr, 1000, dr.ptr, n) == 0
    for x, y in zip(A[:50], got()[:50]):
        e = (1, 0)
        for _ in range(1000):
            e = q2mul(e, x)
        assert y == e
    acc = dev(b)
    assert lib.aext2_scale_accumulate_ffi(5, 9, da.ptr, acc.ptr, n) == 0
    r = acc.to_numpy()
    for i in range(n):
        p = q2mul((5, 9), A[i])
        assert (int(r[2 * i]), int(r[2 * i + 1])) == ((Bv[i][0] + p[0]) % Q, (Bv[i][1] + p[1]) % Q)
    assert lib.agl_to_aext2_batch_ffi(da.ptr, dr.ptr, n) == 0
    assert got() == [(int(a[i]) % Q, 0) for i in range(n)]
    assert lib.aext2_to_agl_batch_ffi(da.ptr, dr.ptr, n) == 0
    assert to_int(dr.to_numpy(n)) == [x[0] for x in A]


@pytest.mark.parametrize("log_n", [0, 1, 5, 10, 13])
def test_eq_tables(log_n):
    print("q eq", log_n, flush=True)
    rng = np.random.default_rng(10 + log_n)
    n = 1 << log_n
    r = words(rng, 2 * max(log_n, 1))
    want = O.q_eq_table(r, log_n, ext=True)
    want_base = O.q_eq_table(r[:max(log_n, 1)], log_n, ext=False)
    dr, a, b = dev(r), M.DeviceBuffer(2 * n), M.DeviceBuffer(2 * n)
    res = ctypes.c_void_p()
    assert lib.aext2_eq_dp_all_ffi(dr.ptr, a.ptr, b.ptr, log_n, ctypes.byref(res)) == 0
    tgt = b if log_n & 1 else a
    assert res.value == tgt.ptr and np.array_equal(tgt.to_numpy(2 * n), want)
    assert lib.aext2_eq_dp_all_stream_ffi(dr.ptr, a.ptr, b.ptr, log_n, ctypes.byref(res), None) == 0
    assert lib.cuda_device_synchronize() == 0
    assert np.array_equal(tgt.to_numpy(2 * n), want)
    drb = dev(r[:max(log_n, 1)])
    assert lib.agl_eq_dp_all_ffi(drb.ptr, a.ptr, b.ptr, log_n, ctypes.byref(res)) == 0
    assert np.array_equal(tgt.to_numpy(n), want_base)


def test_eq_batched():
    print("q eq batched", flush=True)
    rng = np.random.default_rng(20)
    log_n, leaves, stride = 7, 5, (1 << 7) + 3
    r = words(rng, 2 * log_n * leaves)
    a, b = M.DeviceBuffer(2 * stride * leaves), M.DeviceBuffer(2 * stride * leaves)
    res = ctypes.c_void_p()
    assert lib.aext2_eq_dp_all_batched_ffi(dev(r).ptr, a.ptr, b.ptr, log_n, leaves, stride, ctypes.byref(res),
                                           None) == 0
    assert lib.cuda_device_synchronize() == 0
    out = b.to_numpy()
    for l in range(leaves):
        want = O.q_eq_table(r[2 * log_n * l:2 * log_n * (l + 1)], log_n, ext=True)
        assert np.array_equal(out[2 * stride * l:2 * stride * l + 2 * (1 << log_n)], want)


@pytest.mark.parametrize("log_n,m", [(6, 1), (9, 4), (12, 12)])
def test_partial_eval(log_n, m):
    print("q peval", log_n, m, flush=True)
    rng = np.random.default_rng(30 + m)
    f = words(rng, 1 << log_n)
    r_base, r_ext = words(rng, log_n), words(rng, 2 * log_n)
    d = dev(f)
    assert lib.agl_partial_eval_ffi(d.ptr, dev(r_base).ptr, log_n, m) == 0
    want = O.q_partial_eval(f, 0, log_n, r_base, 0, m)[0::2]
    assert np.array_equal(d.to_numpy(1 << (log_n - m)), want)
    out = M.DeviceBuffer(2 << (log_n - 1))
    assert lib.agl_partial_eval_ext2_from_base_ffi(dev(f).ptr, out.ptr, dev(r_ext).ptr, log_n, m) == 0
    assert np.array_equal(out.to_numpy(2 << (log_n - m)), O.q_partial_eval(f, 0, log_n, r_ext, 1, m))


def reduce_partials(part, blocks, width, ext):
    part = part.reshape(blocks, width, 2 if ext else 1)
    s = [[0, 0] for _ in range(width)]
    for blk in range(blocks):
        for c in range(width):
            s[c][0] = (s[c][0] + int(part[blk, c, 0])) % Q
            if ext:
                s[c][1] = (s[c][1] + int(part[blk, c, 1])) % Q
    return s


@pytest.mark.parametrize("ext", [False, True])
@pytest.mark.parametrize("d,n,blocks", [(1, 8, 1), (2, 1 << 12, 7), (3, 1 << 10, 64)])
def test_sumcheck_ref(ext, d, n, blocks):
    print("q sumcheck", ext, d, n, flush=True)
    rng = np.random.default_rng(40 + d)
    w = 2 if ext else 1
    polys = words(rng, w * d * n)
    dp, part = dev(polys), M.DeviceBuffer(blocks * (d + 1) * w)
    fn = lib.aext2_sumcheck_round_message_ffi if ext else lib.agl_sumcheck_round_message_ffi
    assert fn(dp.ptr, part.ptr, d, n, n // 2, blocks) == 0
    want = O.q_sumcheck_msg(1 if ext else 0, polys, d, n, n // 2).reshape(-1, 2)
    got = reduce_partials(part.to_numpy(), blocks, d + 1, ext)
    assert got == [[int(v[0]), int(v[1]) if ext else 0] for v in want]
    ch = [int(x) for x in words(rng, 2)]
    out = M.DeviceBuffer(w * d * n)
    if ext:
        assert lib.aext2_sumcheck_fold_ffi(dp.ptr, out.ptr, ch[0], ch[1], d, n, n // 2) == 0
    else:
        assert lib.agl_sumcheck_fold_ffi(dp.ptr, out.ptr, ch[0], d, n, n // 2) == 0
    folded = O.q_sumcheck_fold(1 if ext else 0, polys, ch if ext else [ch[0], 0], d, n, n // 2)
    got = out.to_numpy()
    stride = w * n
    for i in range(d):
        assert np.array_equal(got[i * stride:i * stride + w * n // 2], folded[i * stride:i * stride + w * n // 2])


def test_sumcheck_batched():
    print("q sumcheck batched", flush=True)
    rng = np.random.default_rng(50)
    n, leaves, bx = 1 << 9, 6, 3
    polys = words(rng, leaves * 2 * 2 * n)
    dp, part = dev(polys), M.DeviceBuffer(bx * leaves * 3 * 2)
    assert lib.aext2_sumcheck_batched_round_message_ffi(dp.ptr, part.ptr, n, n // 2, leaves, bx) == 0
    p = part.to_numpy().reshape(bx, leaves, 3, 2)
    ch = [int(x) for x in words(rng, 2)]
    out = M.DeviceBuffer(leaves * 2 * 2 * n)
    assert lib.aext2_sumcheck_batched_fold_ffi(dp.ptr, out.ptr, ch[0], ch[1], n, n // 2, leaves) == 0
    fo = out.to_numpy()
    for l in range(leaves):
        leaf = polys[l * 4 * n:(l + 1) * 4 * n]
        want = O.q_sumcheck_msg(1, leaf, 2, n, n // 2).reshape(-1, 2)
        for c in range(3):
            got = [sum(int(p[b, l, c, k]) for b in range(bx)) % Q for k in range(2)]
            assert got == [int(want[c][0]), int(want[c][1])]
        fw = O.q_sumcheck_fold(1, leaf, ch, 2, n, n // 2)
        for i in range(2):
            s = l * 4 * n + i * 2 * n
            assert np.array_equal(fo[s:s + n], fw[i * 2 * n:i * 2 * n + n])


KEY = bytes((7 * i + 3) % 256 for i in range(32))


@pytest.mark.parametrize("N,B,chunk,col_offset", [(1, 1, 64, 0), (37, 1, 64, 0), (8, 2, 128, 5), (13, 4, 4096, 0),
                                                  (6, 8, 256, (1 << 32) + 1), (5, 16, 1024, 2)])
def test_ring_commit_dense(N, B, chunk, col_offset):
    print("ring dense", N, B, chunk, col_offset, flush=True)
    rng = np.random.default_rng(N * 131 + B)
    z = rng.integers(0, 1 << 64, B * N, dtype=np.uint64)
    z[0] = 0
    if B * N > 1:
        z[1] = (1 << 64) - 1
    seed = R.Seed.from_bytes(KEY)
    got = R.commit_batched_at(seed, [z[b * N:(b + 1) * N] for b in range(B)], col_offset, chunk)
    want = O.ring_commit_binary(KEY, z, N, B, col_offset)
    for b in range(B):
        assert np.array_equal(got[b].rows, want[b]), b
    assert np.array_equal(R.matrix_entry(seed, 41, col_offset + N - 1),
                          O.ring_matrix_entry(KEY, 41, col_offset + N - 1))


def test_ring_commit_errors_and_default_chunk():
    seed = R.Seed.from_bytes(KEY)
    dk = R._key_buffer(seed)
    dz, do = dev(np.ones(12, dtype=np.uint64)), M.DeviceBuffer(3 * 42 * 64)
    assert lib.ajtai_commit_dense_batched_ffi(dk.ptr, dz.ptr, 4, 3, 64, do.ptr) == -2
    assert lib.ajtai_commit_dense_batched_ffi(dk.ptr, dz.ptr, 4, 1, 100, do.ptr) == -3
    assert lib.ajtai_commit_wide_ffi(dk.ptr, dz.ptr, 1, 0, 65, do.ptr) == -3
    assert lib.ajtai_commit_sparse_ffi(dk.ptr, dz.ptr, 1, 0, do.ptr) == -3
    z = np.arange(1, 21, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    assert R.commit(seed, z) == R.RingCommitment(O.ring_commit_binary(KEY, z, 20, 1)[0])


@pytest.mark.parametrize("N,col_offset,chunk", [(1, 0, 64), (9, 3, 128), (24, 0, 4096)])
def test_ring_commit_wide(N, col_offset, chunk):
    print("ring wide", N, col_offset, flush=True)
    rng = np.random.default_rng(60 + N)
    zw = words(rng, 64 * N)
    got = R.commit_wide(R.Seed.from_bytes(KEY), zw, col_offset, chunk)
    assert np.array_equal(got.rows, O.ring_commit_wide(KEY, zw, N, col_offset))


def test_ring_commit_sparse():
    print("ring sparse", flush=True)
    rng = np.random.default_rng(70)
    pos = rng.integers(0, 64 * 30, 300, dtype=np.uint64)
    pos[:10] = 5  # duplicates add
    pos = np.concatenate([pos, np.sort(rng.integers(0, 64 * 30, 200, dtype=np.uint64))])
    seed = R.Seed.from_bytes(KEY)
    for chunk in (64, 4096):
        assert np.array_equal(R.commit_sparse(seed, pos, chunk).rows, O.ring_commit_sparse(KEY, pos))
    assert np.array_equal(R.commit_sparse(seed, []).rows, np.zeros((42, 64), dtype=np.uint64))


# (row, column, block) whose first ChaCha8 block holds a word >= q for REJ_KEY (a search of
# the first 8 M columns): the matrix entries on which the rejection path runs.
REJ_KEY = bytes((11 * i + 5) % 256 for i in range(32))
REJECTS = [(40, 2108465, 3), (31, 2660596, 5), (21, 6196264, 3), (0, 7120001, 0)]


def test_ring_commit_rejected_samples():
    print("ring rejection path", flush=True)
    seed = R.Seed.from_bytes(REJ_KEY)
    rng = np.random.default_rng(71)
    for row, col, blk in REJECTS:
        assert np.array_equal(R.matrix_entry(seed, row, col), O.ring_matrix_entry(REJ_KEY, row, col))
        off, N = col - 2, 5
        z = rng.integers(0, 1 << 64, 2 * N, dtype=np.uint64)
        got = R.commit_batched_at(seed, [z[:N], z[N:]], off, 64)
        want = O.ring_commit_binary(REJ_KEY, z, N, 2, off)
        for b in range(2):
            assert np.array_equal(got[b].rows, want[b]), (row, col, b)
        zw = words(rng, 64 * N)
        assert np.array_equal(R.commit_wide(seed, zw, off, 128).rows, O.ring_commit_wide(REJ_KEY, zw, N, off))
        pos = np.array([64 * col + 8 * blk, 64 * col + 63, 64 * (col - 1) + 5], dtype=np.uint64)
        assert np.array_equal(R.commit_sparse(seed, pos, 64).rows, O.ring_commit_sparse(REJ_KEY, pos))


def lifted(polys, bits, n, leaves, pw):
    """the batched layout with each leaf's f slot replaced by its bits as Ext2 (c1 = 0)"""
    out = polys.copy()
    for l in range(leaves):
        b = bits[l * pw:(l + 1) * pw]
        f = np.array([(int(b[i >> 6]) >> (i & 63)) & 1 for i in range(n)], dtype=np.uint64)
        s = l * 4 * n + 2 * n
        out[s:s + 2 * n:2] = f
        out[s + 1:s + 2 * n:2] = 0
    return out


def test_sumcheck_batched_round0_binary_and_lifts():
    print("q sumcheck batched round 0 binary", flush=True)
    rng = np.random.default_rng(51)
    n, leaves, bx = 1 << 9, 5, 3
    pw = n // 64 + 1  # a leaf's packed block may be longer than its bits
    polys = words(rng, leaves * 2 * 2 * n)  # the f slots hold garbage the round-0 kernels must ignore
    bits = rng.integers(0, 1 << 64, leaves * pw, dtype=np.uint64)
    want_layout = lifted(polys, bits, n, leaves, pw)
    dp, db = dev(polys), dev(bits)
    part = M.DeviceBuffer(bx * leaves * 3 * 2)
    assert lib.aext2_sumcheck_batched_round0_binary_msg_ffi(dp.ptr, db.ptr, part.ptr, n, n // 2, leaves, bx, pw) == 0
    p = part.to_numpy().reshape(bx, leaves, 3, 2)
    # the same per-block partials as the plain kernel on the lifted layout
    dl, part2 = dev(want_layout), M.DeviceBuffer(bx * leaves * 3 * 2)
    assert lib.aext2_sumcheck_batched_round_message_ffi(dl.ptr, part2.ptr, n, n // 2, leaves, bx) == 0
    assert np.array_equal(part.to_numpy(), part2.to_numpy())
    ch = [int(x) for x in words(rng, 2)]
    out = M.DeviceBuffer(leaves * 2 * 2 * n)
    assert lib.aext2_sumcheck_batched_round0_binary_fold_ffi(dp.ptr, db.ptr, out.ptr, ch[0], ch[1], n, n // 2, leaves,
                                                             pw) == 0
    fo = out.to_numpy()
    for l in range(leaves):
        leaf = want_layout[l * 4 * n:(l + 1) * 4 * n]
        want = O.q_sumcheck_msg(1, leaf, 2, n, n // 2).reshape(-1, 2)
        for c in range(3):
            got = [sum(int(p[b, l, c, k]) for b in range(bx)) % Q for k in range(2)]
            assert got == [int(want[c][0]), int(want[c][1])], (l, c)
        fw = O.q_sumcheck_fold(1, leaf, ch, 2, n, n // 2)
        for i in range(2):
            s = l * 4 * n + i * 2 * n
            assert np.array_equal(fo[s:s + n], fw[i * 2 * n:i * 2 * n + n]), (l, i)
    # binary lift writes the f slots only
    assert lib.aext2_batched_lift_binary_ffi(db.ptr, dp.ptr, n, leaves, pw) == 0
    assert np.array_equal(dp.to_numpy(), want_layout)
    # ternary lift: pos - neg, a bit set in both gives 0
    neg = rng.integers(0, 1 << 64, leaves * pw, dtype=np.uint64)
    dn = dev(neg)
    assert lib.aext2_batched_lift_ternary_single_ffi(db.ptr, dn.ptr, dp.ptr, n, leaves, pw) == 0
    got = dp.to_numpy()
    for l in range(leaves):
        s = l * 4 * n + 2 * n
        for i in range(0, n, 7):
            pb = (int(bits[l * pw + (i >> 6)]) >> (i & 63)) & 1
            nb = (int(neg[l * pw + (i >> 6)]) >> (i & 63)) & 1
            assert int(got[s + 2 * i]) == (pb - nb) % Q and int(got[s + 2 * i + 1]) == 0, (l, i)
        assert np.array_equal(got[l * 4 * n:l * 4 * n + 2 * n], polys[l * 4 * n:l * 4 * n + 2 * n])
    # a packed block shorter than the bits is rejected
    assert lib.aext2_batched_lift_binary_ffi(db.ptr, dp.ptr, n, leaves, n // 64 - 1) == -1


def test_bit_permute_twin_and_memory_entry_points():
    print("q bit permute, memory", flush=True)
    rng = np.random.default_rng(52)
    nb = 10
    x = words(rng, 1 << nb)
    inv = np.array(rng.permutation(nb), dtype=np.int32)
    dx, dy = dev(x), M.DeviceBuffer(1 << nb)
    dm = M.DeviceBuffer(nb)
    assert lib.cuda_memcpy_htod(dm.ptr, inv.ctypes.data, inv.nbytes) == 0
    assert lib.agl_bit_permute_ffi(dx.ptr, dy.ptr, dm.ptr, nb, 1 << nb) == 0
    y = dy.to_numpy()
    for i in range(0, 1 << nb, 37):
        src = sum(((i >> b) & 1) << int(inv[b]) for b in range(nb))
        assert y[i] == x[src]
    # caching allocator: a freed block is reused for a request of the same size and trimmed on demand
    gc.collect()
    assert lib.cuda_pool_trim(0) == 0
    p = ctypes.c_void_p()
    assert lib.cuda_malloc(ctypes.byref(p), 3 << 20) == 0
    first = p.value
    assert lib.cuda_free(first) == 0
    assert lib.agl_cached_bytes() >= 3 << 20
    assert lib.cuda_malloc(ctypes.byref(p), 3 << 20) == 0
    assert p.value == first
    # memset, zero buffer, staged copies, async copies on a caller stream
    h = np.arange(1 << 16, dtype=np.uint64)
    assert lib.agl_h2d_staged(p.value, h.ctypes.data, h.nbytes) == 0
    back = np.zeros_like(h)
    assert lib.agl_d2h_staged(back.ctypes.data, p.value, h.nbytes) == 0
    assert np.array_equal(back, h)
    assert lib.cuda_memset(p.value, 0xAB, 64) == 0
    assert lib.agl_d2h_staged(back.ctypes.data, p.value, h.nbytes) == 0
    assert all(int(v) == 0xABABABABABABABAB for v in back[:8]) and int(back[8]) == 8
    assert lib.agl_zero_buffer_ffi(p.value, 1 << 16) == 0
    assert lib.agl_d2h_staged(back.ctypes.data, p.value, h.nbytes) == 0
    assert not back.any()
    s = ctypes.c_void_p()
    assert lib.cuda_stream_create(ctypes.byref(s)) == 0
    q = ctypes.c_void_p()
    assert lib.cuda_malloc(ctypes.byref(q), h.nbytes) == 0
    assert lib.cuda_memcpy_htod_async(p.value, h.ctypes.data, h.nbytes, s.value) == 0
    assert lib.cuda_memcpy_dtod_async(q.value, p.value, h.nbytes, s.value) == 0
    assert lib.cuda_stream_synchronize(s.value) == 0
    assert lib.cuda_stream_destroy(s.value) == 0
    assert lib.cuda_memcpy_peer(p.value, 0, q.value, 0, h.nbytes) == 0
    assert lib.cuda_sync_mine() == 0
    assert lib.cuda_memcpy_dtoh(back.ctypes.data, p.value, h.nbytes) == 0
    assert np.array_equal(back, h)
    assert isinstance(lib.agl_cuda_error_string(0), bytes)
    assert lib.cuda_free(q.value) == 0 and lib.cuda_free(p.value) == 0
    gc.collect()
    assert lib.cuda_pool_trim(0) == 0
    assert lib.agl_cached_bytes() == 0
