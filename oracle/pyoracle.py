"""ctypes wrapper of the oracle (liboracle.so, built from oracle/*.c by oracle/Makefile).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker. The product (morpheus_amd) never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} missing: run make -C oracle")
L = ctypes.CDLL(LIB_PATH)

u64 = ctypes.c_uint64
u64p = ctypes.POINTER(u64)
ci = ctypes.c_int
cl = ctypes.c_long
P = 0xFFFFFFFF00000001
MONOLITH = 0
POSEIDON2 = 1


class Transcript_t(ctypes.Structure):
    _fields_ = [("state", u64 * 8), ("in_buf", u64 * 4), ("out_buf", u64 * 4),
                ("in_count", ctypes.c_int64), ("out_count", ctypes.c_int64)]


trp = ctypes.POINTER(Transcript_t)
_sigs = {
    "or_gl_add": (u64, [u64, u64]), "or_gl_sub": (u64, [u64, u64]), "or_gl_mul": (u64, [u64, u64]),
    "or_gl_inv": (u64, [u64]), "or_ext2_mul": (None, [u64p, u64p, u64p]), "or_ext2_inv": (None, [u64p, u64p]),
    "or_poseidon2_permute": (None, [u64p, ci]), "or_monolith_permute": (None, [u64p]),
    "or_monolith_hash_gl_leaf": (None, [u64, u64, u64p]), "or_monolith_compress": (None, [u64p, u64p, u64p]),
    "or_merkle_tree_gl": (None, [u64p, cl, ci, u64p]), "or_merkle_tree_ext2": (None, [u64p, cl, ci, u64p]),
    "or_tr_new": (None, [trp, ctypes.c_char_p, cl]), "or_tr_append_u64": (None, [trp, ctypes.c_char_p, cl, u64]),
    "or_tr_append_scalars": (None, [trp, ctypes.c_char_p, cl, u64p, cl]),
    "or_tr_append_ext2": (None, [trp, ctypes.c_char_p, cl, u64p]),
    "or_tr_challenge_scalar": (u64, [trp, ctypes.c_char_p, cl]),
    "or_tr_challenge_ext2": (None, [trp, ctypes.c_char_p, cl, u64p]),
    "or_eq_table": (None, [u64p, ci, u64p]), "or_eq_table_ext2": (None, [u64p, ci, u64p]),
    "or_partial_eval_ext2": (None, [u64p, ci, u64p, ci, u64p]),
    "or_sumcheck_prove_ext2": (None, [trp, u64p, ci, ci, u64p, u64p, u64p, u64p]),
    "or_sumcheck_verify_ext2": (ci, [trp, u64p, ci, ci, u64p, u64p]),
    "or_basefold_table_len": (cl, [ci, ci, ci]), "or_basefold_table_generate": (None, [ci, ci, ci, u64, u64p]),
    "or_basefold_commit": (None, [u64p, ci, ci, ci, u64p, u64p, u64p]),
    "or_basefold_proof_len": (cl, [ci, ci, ci]),
    "or_basefold_open_ext2": (None, [u64p, u64p, u64p, ci, ci, ci, u64p, u64p, ci, trp, ci, u64p]),
    "or_basefold_open_ext2_tree": (None, [u64p, u64p, u64p, u64p, ci, ci, ci, u64p, u64p, ci, trp, ci, u64p]),
    "or_set_threads": (None, [ci]),
    "or_get_threads": (ci, []),
    "or_basefold_verify_ext2": (ci, [u64p, u64p, ci, ci, ci, u64p, ci, u64p, ci, trp]),
    "or_basefold_batch_proof_len": (cl, [ci, ci, ci, ci]),
    "or_basefold_batch_open_ext2": (None, [ci, ctypes.POINTER(u64p), ctypes.POINTER(u64p), u64p, ci, ci, ci, ci, u64p,
                                           ci, ctypes.POINTER(ci), ctypes.POINTER(ci), u64p, u64p, ci, trp, ci, u64p]),
    "or_basefold_batch_verify_ext2": (ci, [ci, u64p, ci, ci, ci, ci, u64p, ci, ctypes.POINTER(ci), ctypes.POINTER(ci),
                                           u64p, u64p, ci, u64p, ci, trp]),
    "or_chacha_block": (None, [ci, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                               ctypes.POINTER(ctypes.c_uint32)]),
    "or_ajtai_i8_row": (None, [ctypes.c_char_p, ctypes.c_uint32, u64, ctypes.POINTER(ctypes.c_int8)]),
    "or_ajtai_i8_commit": (None, [ctypes.c_char_p, ci, u64, u64p, ci, u64p]),
    "or_q_add": (u64, [u64, u64]), "or_q_sub": (u64, [u64, u64]), "or_q_mul": (u64, [u64, u64]),
    "or_q_inv": (u64, [u64]), "or_q2_mul": (None, [u64p, u64p, u64p]), "or_q2_inv": (None, [u64p, u64p]),
    "or_q_eq_table": (None, [ci, u64p, ci, u64p]),
    "or_q_partial_eval": (None, [u64p, ci, ci, u64p, ci, ci, u64p]),
    "or_q_sumcheck_msg": (None, [ci, u64p, ci, u64, u64, u64p]),
    "or_q_sumcheck_fold": (None, [ci, u64p, u64p, u64p, ci, u64, u64]),
    "or_ring_matrix_entry": (None, [ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, u64, u64p]),
    "or_ring_commit_binary": (None, [ctypes.POINTER(ctypes.c_uint32), u64p, u64, ci, u64, u64p]),
    "or_ring_commit_wide": (None, [ctypes.POINTER(ctypes.c_uint32), u64p, u64, u64, u64p]),
    "or_ring_commit_sparse": (None, [ctypes.POINTER(ctypes.c_uint32), u64p, u64, u64p]),
    "or_fused_permute_peval": (None, [u64, u64p, u64p, ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32), ci, ci, ci, cl, u64p]),
}
for _n, (_r, _a) in _sigs.items():
    _f = getattr(L, _n)
    _f.restype = _r
    _f.argtypes = _a


def _p(a):
    return a.ctypes.data_as(u64p)


def arr(x):
    return np.ascontiguousarray(x, dtype=np.uint64)


class Transcript:
    def __init__(self, label=b""):
        self.t = Transcript_t()
        L.or_tr_new(ctypes.byref(self.t), label, len(label))

    def clone(self):
        c = Transcript.__new__(Transcript)
        c.t = Transcript_t()
        ctypes.pointer(c.t)[0] = self.t
        return c

    def append_u64(self, label, v):
        L.or_tr_append_u64(ctypes.byref(self.t), label, len(label), int(v))

    def append_scalars(self, label, v):
        a = arr(v)
        L.or_tr_append_scalars(ctypes.byref(self.t), label, len(label), _p(a), a.size)

    def append_ext2(self, label, v2):
        a = arr(v2)
        L.or_tr_append_ext2(ctypes.byref(self.t), label, len(label), _p(a))

    def challenge_scalar(self, label=b""):
        return int(L.or_tr_challenge_scalar(ctypes.byref(self.t), label, len(label)))

    def challenge_ext2(self, label=b""):
        o = np.zeros(2, dtype=np.uint64)
        L.or_tr_challenge_ext2(ctypes.byref(self.t), label, len(label), _p(o))
        return o

    def state(self):
        return np.array(self.t.state, dtype=np.uint64)


def poseidon2_permute(state8, initial_linear_layer=0):
    s = arr(state8).copy()
    L.or_poseidon2_permute(_p(s), initial_linear_layer)
    return s


def monolith_permute(state12):
    s = arr(state12).copy()
    L.or_monolith_permute(_p(s))
    return s


def eq_table_ext2(point2, nv):
    out = np.zeros(2 << nv, dtype=np.uint64)
    L.or_eq_table_ext2(_p(arr(point2)), nv, _p(out))
    return out


def eq_table(r, nv):
    out = np.zeros(1 << nv, dtype=np.uint64)
    L.or_eq_table(_p(arr(r)), nv, _p(out))
    return out


def partial_eval_ext2(in2, nv, r2, m):
    out = np.zeros(2 << (nv - m), dtype=np.uint64)
    L.or_partial_eval_ext2(_p(arr(in2)), nv, _p(arr(r2)), m, _p(out))
    return out


def sumcheck_prove_ext2(t, polys_ext2, np_, nv):
    """polys_ext2: [np][2^nv][2] words (copied). Returns dict like the library's."""
    p = arr(polys_ext2).copy()
    msgs = np.zeros((nv, np_ + 1, 2), dtype=np.uint64)
    chs = np.zeros((nv, 2), dtype=np.uint64)
    fin = np.zeros((np_, 2), dtype=np.uint64)
    fe = np.zeros(2, dtype=np.uint64)
    L.or_sumcheck_prove_ext2(ctypes.byref(t.t), _p(p), np_, nv, _p(msgs), _p(chs), _p(fin), _p(fe))
    return {"round_msgs": msgs, "challenges": chs, "final_evals": fin, "final_eval": fe}


def sumcheck_verify_ext2(t, claim2, np_, nv, round_msgs, final_eval):
    return bool(L.or_sumcheck_verify_ext2(ctypes.byref(t.t), _p(arr(claim2)), np_, nv, _p(arr(round_msgs)),
                                          _p(arr(final_eval))))


def basefold_table(nv, lr, nrounds, seed):
    n = L.or_basefold_table_len(nv, lr, nrounds)
    e = np.zeros(2 * n, dtype=np.uint64)
    L.or_basefold_table_generate(nv, lr, nrounds, seed, _p(e))
    return e


def basefold_commit(evals, nv, lr, hash_kind=MONOLITH, with_tree=True):
    cw = np.zeros(1 << (nv + lr), dtype=np.uint64)
    bh = np.zeros(1 << nv, dtype=np.uint64)
    tree = np.zeros(((1 << (nv + lr)) - 1) * 4, dtype=np.uint64) if with_tree else None
    L.or_basefold_commit(_p(arr(evals)), nv, lr, hash_kind, _p(cw), _p(bh), _p(tree) if with_tree else None)
    return cw, bh, tree


def merkle_tree(cw, num_leaves, hash_kind=MONOLITH, ext2=False):
    tree = np.zeros((2 * num_leaves - 1) * 4, dtype=np.uint64)
    f = L.or_merkle_tree_ext2 if ext2 else L.or_merkle_tree_gl
    f(_p(arr(cw)), num_leaves, hash_kind, _p(tree))
    return tree


def basefold_proof_len(nv, lr, nq):
    return int(L.or_basefold_proof_len(nv, lr, nq))


def basefold_open_ext2(cw, bh, root, nv, lr, hash_kind, point2, table, table_nv, t, nq, tree=None):
    """tree: the commitment's Merkle tree (as basefold_commit returns it); None rebuilds it."""
    proof = np.zeros(basefold_proof_len(nv, lr, nq), dtype=np.uint64)
    if tree is None:
        L.or_basefold_open_ext2(_p(arr(cw)), _p(arr(bh)), _p(arr(root)), nv, lr, hash_kind, _p(arr(point2)),
                                _p(arr(table)), table_nv, ctypes.byref(t.t), nq, _p(proof))
    else:
        L.or_basefold_open_ext2_tree(_p(arr(cw)), _p(arr(bh)), _p(arr(root)), _p(arr(tree)), nv, lr, hash_kind,
                                     _p(arr(point2)), _p(arr(table)), table_nv, ctypes.byref(t.t), nq, _p(proof))
    return proof


def set_threads(n):
    """Threads of the oracle's parallel loops (results do not depend on the count)."""
    L.or_set_threads(int(n))


def get_threads():
    return int(L.or_get_threads())


def basefold_verify_ext2(root, point2, nv, lr, hash_kind, proof, nq, table, table_nv, t):
    return bool(L.or_basefold_verify_ext2(_p(arr(root)), _p(arr(point2)), nv, lr, hash_kind, _p(arr(proof)), nq,
                                          _p(arr(table)), table_nv, ctypes.byref(t.t)))


def _evals_arrays(evals):
    poly = np.array([e[0] for e in evals], dtype=np.int32)
    point = np.array([e[1] for e in evals], dtype=np.int32)
    vals = np.array([[int(e[2][0]), int(e[2][1])] for e in evals], dtype=np.uint64).reshape(-1)
    return poly, point, vals


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ci))


def basefold_batch_proof_len(nv, lr, nq, num_evals):
    return int(L.or_basefold_batch_proof_len(nv, lr, nq, num_evals))


def basefold_batch_open_ext2(comms, nv, lr, hash_kind, points, evals, table, table_nv, t, nq):
    """comms: list of (codeword, bh_evals, root4); points: [num_points][nv][2];
    evals: list of (poly, point, value2)."""
    cws = [arr(c[0]) for c in comms]
    bhs = [arr(c[1]) for c in comms]
    roots = arr(np.concatenate([np.asarray(c[2], dtype=np.uint64).ravel() for c in comms]))
    pts = arr(np.asarray(points, dtype=np.uint64).ravel())
    poly, point, vals = _evals_arrays(evals)
    cwp = (u64p * len(comms))(*[_p(c) for c in cws])
    bhp = (u64p * len(comms))(*[_p(b) for b in bhs])
    proof = np.zeros(basefold_batch_proof_len(nv, lr, nq, len(evals)), dtype=np.uint64)
    L.or_basefold_batch_open_ext2(len(comms), cwp, bhp, _p(roots), nv, lr, hash_kind, len(points), _p(pts), len(evals),
                                  _ip(poly), _ip(point), _p(vals), _p(arr(table)), table_nv, ctypes.byref(t.t), nq,
                                  _p(proof))
    return proof


def basefold_batch_verify_ext2(roots, nv, lr, hash_kind, points, evals, proof, nq, table, table_nv, t):
    roots = arr(np.asarray(roots, dtype=np.uint64).ravel())
    pts = arr(np.asarray(points, dtype=np.uint64).ravel())
    poly, point, vals = _evals_arrays(evals)
    return bool(L.or_basefold_batch_verify_ext2(roots.size // 4, _p(roots), nv, lr, hash_kind, len(points), _p(pts),
                                                len(evals), _ip(poly), _ip(point), _p(vals), _p(arr(proof)), nq,
                                                _p(arr(table)), table_nv, ctypes.byref(t.t)))


def chacha_block(rounds, key_words8, counter, nonce3):
    k = np.ascontiguousarray(key_words8, dtype=np.uint32)
    nn = np.ascontiguousarray(nonce3, dtype=np.uint32)
    out = np.zeros(16, dtype=np.uint32)
    U32 = ctypes.POINTER(ctypes.c_uint32)
    L.or_chacha_block(rounds, k.ctypes.data_as(U32), counter, nn.ctypes.data_as(U32), out.ctypes.data_as(U32))
    return out


def ajtai_i8_row(key32, row, n):
    out = np.zeros(n, dtype=np.int8)
    L.or_ajtai_i8_row(bytes(key32), row, n, out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
    return out


def ajtai_i8_commit(key32, rows, n, z, batch):
    out = np.zeros(batch * rows, dtype=np.uint64)
    L.or_ajtai_i8_commit(bytes(key32), rows, n, _p(arr(z)), batch, _p(out))
    return out


# ---- almost-Goldilocks (oracle/q.c) ----
Q = 0xFFFFFFFEFFFFFFE1
KAPPA = 42


def _u32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))


def q_eq_table(r, nv, ext=True):
    r = arr(r)
    out = np.zeros((2 if ext else 1) << nv, dtype=np.uint64)
    L.or_q_eq_table(1 if ext else 0, _p(r), nv, _p(out))
    return out


def q_partial_eval(data, in_ext, nv, r, r_ext, m):
    data, r = arr(data), arr(r)
    out = np.zeros(2 << (nv - m), dtype=np.uint64)
    L.or_q_partial_eval(_p(data), in_ext, nv, _p(r), r_ext, m, _p(out))
    return out


def q_sumcheck_msg(ext, polys, d, original_size, half):
    polys = arr(polys)
    out = np.zeros(2 * (d + 1), dtype=np.uint64)
    L.or_q_sumcheck_msg(ext, _p(polys), d, original_size, half, _p(out))
    return out


def q_sumcheck_fold(ext, polys, ch2, d, original_size, half):
    polys = arr(polys)
    out = polys.copy()
    ch = arr(ch2)
    L.or_q_sumcheck_fold(ext, _p(polys), _p(out), _p(ch), d, original_size, half)
    return out


def ring_key_words(key32):
    return np.frombuffer(bytes(key32), dtype="<u4").astype(np.uint32).copy()


def ring_matrix_entry(key32, row, col):
    kw = ring_key_words(key32)
    out = np.zeros(64, dtype=np.uint64)
    L.or_ring_matrix_entry(_u32p(kw), row, col, _p(out))
    return out


def ring_commit_binary(key32, z, N, B, col_offset=0):
    kw, z = ring_key_words(key32), arr(z)
    out = np.zeros(B * KAPPA * 64, dtype=np.uint64)
    L.or_ring_commit_binary(_u32p(kw), _p(z), N, B, col_offset, _p(out))
    return out.reshape(B, KAPPA, 64)


def ring_commit_wide(key32, z_wide, N, col_offset=0):
    kw, z = ring_key_words(key32), arr(z_wide)
    out = np.zeros(KAPPA * 64, dtype=np.uint64)
    L.or_ring_commit_wide(_u32p(kw), _p(z), N, col_offset, _p(out))
    return out.reshape(KAPPA, 64)


def ring_commit_sparse(key32, positions):
    kw, pos = ring_key_words(key32), arr(positions)
    out = np.zeros(KAPPA * 64, dtype=np.uint64)
    L.or_ring_commit_sparse(_u32p(kw), _p(pos), len(pos), _p(out))
    return out.reshape(KAPPA, 64)


def split_luts(n, new_to_old, half):
    """LUTs of a variable permutation: bit p of the permuted index is bit new_to_old[p] of the
    stored index (goldilocks-cuda-rs/src/partial_eval.rs:266-321 builds the same tables)."""
    lo = np.zeros(1 << half, dtype=np.uint32)
    hi = np.zeros(1 << (n - half), dtype=np.uint32)
    for x in range(1 << half):
        lo[x] = sum(((x >> p) & 1) << new_to_old[p] for p in range(half))
    for y in range(1 << (n - half)):
        hi[y] = sum(((y >> p) & 1) << new_to_old[half + p] for p in range(n - half))
    return lo, hi


def fused_permute_peval(mod, evals, eq, lo, hi, n, m, half, out_size):
    evals, eq = arr(evals), arr(eq)
    out = np.zeros(2 * out_size, dtype=np.uint64)
    L.or_fused_permute_peval(mod, _p(evals), _p(eq), _u32p(lo), _u32p(hi), n, m, half, out_size, _p(out))
    return out


def lift_int(x):
    """Signed integers -> canonical Goldilocks words."""
    x = np.asarray(x, dtype=np.int64)
    return np.where(x < 0, np.uint64(P) - (-x).astype(np.uint64), x.astype(np.uint64)).astype(np.uint64)


def _ext(words):
    e = np.zeros((words.size, 2), dtype=np.uint64)
    e[:, 0] = words
    return e


def matmul_prove(a, b, label, table_entries, table_nv, lr, nq, key=None, rows=0, hash_kind=MONOLITH):
    """Matmul proof C = A B for integer A (m x k), B (k x n), powers of two, restated step by
    step (zk-torch-3/src/basicblock/einsum.rs:807-1098 is the statement: partial-evaluate A and
    B at the output point, then a sumcheck over the inner index with polys [A', B', ones]).
    Variables are LSB-first over the row-major index: C(j, i) with j the column bits first.
    Transcript: roots of A, B, C ("matmul-root"), optional int8 Ajtai of C ("matmul-ajtai"),
    point r = (r_j, r_i) ("matmul-point"), claim C(r) ("matmul-claim"), the degree-3 sumcheck
    over l, then Basefold openings of A at (r_l, r_i), B at (r_j, r_l), C at r."""
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    c = (a.astype(np.float64) @ b.astype(np.float64)).astype(np.int64)  # exact: |c| < 2^53
    m_, k_ = a.shape
    n_ = b.shape[1]
    lm, lk, ln = (m_ - 1).bit_length(), (k_ - 1).bit_length(), (n_ - 1).bit_length()
    A, B, C = lift_int(a).ravel(), lift_int(b).ravel(), lift_int(c).ravel()
    nvs = (lm + lk, lk + ln, lm + ln)
    t = Transcript(label)
    com = []
    for X, nv in zip((A, B, C), nvs):
        cw, bh, tree = basefold_commit(X, nv, lr, hash_kind)
        com.append((cw, bh, tree[-4:]))
        t.append_scalars(b"matmul-root", tree[-4:])
    out = {"roots": np.array([x[2] for x in com])}
    if key is not None:
        aj = ajtai_i8_commit(key, rows, 1 << nvs[2], C, 1)
        t.append_scalars(b"matmul-ajtai", aj)
        out["ajtai"] = aj
    point = np.array([t.challenge_ext2(b"matmul-point") for _ in range(ln + lm)], dtype=np.uint64)
    claim = partial_eval_ext2(_ext(C), nvs[2], point.ravel(), nvs[2]).ravel()[:2]
    t.append_ext2(b"matmul-claim", claim)
    r_j, r_i = point[:ln], point[ln:]
    a_perm = A.reshape(m_, k_).T.ravel()  # row variables first
    a1 = partial_eval_ext2(_ext(a_perm), nvs[0], r_i.ravel(), lm).reshape(-1, 2)
    b1 = partial_eval_ext2(_ext(B), nvs[1], r_j.ravel(), ln).reshape(-1, 2)
    ones = np.zeros((k_, 2), dtype=np.uint64)
    ones[:, 0] = 1
    sc = sumcheck_prove_ext2(t, np.stack([a1, b1, ones]), 3, lk)
    r_l = sc["challenges"]
    pts = (np.concatenate([r_l, r_i]), np.concatenate([r_j, r_l]), point)
    for name, (cw, bh, root), nv, pt in zip(("proof_a", "proof_b", "proof_c"), com, nvs, pts):
        out[name] = basefold_open_ext2(cw, bh, root, nv, lr, hash_kind, pt.ravel(), table_entries, table_nv, t, nq)
    out.update(point=point, claim=claim, round_msgs=sc["round_msgs"], challenges=r_l, final_evals=sc["final_evals"])
    return out


def sumcheck_proof_bincode(final_eval, round_msgs):
    """bincode 1.x (fixed-width little-endian integers, u64 Vec lengths) of zk-torch-3's
    SumcheckProof { final_eval: Ext2, round_messages: Vec<Vec<Ext2>> }
    (zk-torch-3/src/sumcheck/mod.rs:16-23), field words as canonical residues."""
    import struct
    msgs = np.asarray(round_msgs, dtype=np.uint64)
    out = struct.pack("<QQ", int(final_eval[0]) % P, int(final_eval[1]) % P)
    out += struct.pack("<Q", msgs.shape[0])
    for r in range(msgs.shape[0]):
        out += struct.pack("<Q", msgs.shape[1])
        for e in range(msgs.shape[1]):
            out += struct.pack("<QQ", int(msgs[r, e, 0]) % P, int(msgs[r, e, 1]) % P)
    return out


# ---- folding, splitting and ternary commits of ring witnesses (q crate) ----
L.or_ring_fold_witness.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
L.or_ring_fold_commitment.argtypes = [ctypes.c_void_p] * 4
L.or_ring_multifold_mixed_witness.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int, ctypes.c_int, ctypes.c_uint64,
                                                                       ctypes.c_void_p]
L.or_ring_multifold_commitment.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
L.or_ring_split_witness.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
L.or_ring_commit_ternary.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_uint64, ctypes.c_void_p]
L.or_ring_materialize.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
L.or_q_selective_add_planes.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int,
                                        ctypes.c_uint64, ctypes.c_void_p]


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def ring_fold_witness(z1, z2, r64):
    z1, z2, r = _c(z1, np.uint64), _c(z2, np.uint64), _c(r64, np.int8)
    out = np.zeros(z1.size * 64, dtype=np.uint64)
    L.or_ring_fold_witness(z1.ctypes.data, z2.ctypes.data, r.ctypes.data, z1.size, out.ctypes.data)
    return out


def ring_fold_commitment(c1, c2, r64):
    c1, c2, r = _c(c1, np.uint64), _c(c2, np.uint64), _c(r64, np.int8)
    out = np.zeros(KAPPA * 64, dtype=np.uint64)
    L.or_ring_fold_commitment(c1.ctypes.data, c2.ctypes.data, r.ctypes.data, out.ctypes.data)
    return out


def ring_multifold_mixed_witness(bin_packed, pos, neg, r_all, K, T, N):
    b = _c(bin_packed if K else np.zeros(1), np.uint64)
    p = _c(pos if T else np.zeros(1), np.uint64)
    n = _c(neg if T else np.zeros(1), np.uint64)
    r = _c(r_all, np.int8)
    out = np.zeros(N * 64, dtype=np.int16)
    L.or_ring_multifold_mixed_witness(b.ctypes.data, p.ctypes.data, n.ctypes.data, r.ctypes.data, K, T, N,
                                      out.ctypes.data)
    return out


def ring_multifold_commitment(c_all, r_all, M):
    c, r = _c(c_all, np.uint64), _c(r_all, np.int8)
    out = np.zeros(KAPPA * 64, dtype=np.uint64)
    L.or_ring_multifold_commitment(c.ctypes.data, r.ctypes.data, M, out.ctypes.data)
    return out


def ring_split_witness(z_wide_i16, N):
    z = _c(z_wide_i16, np.int16)
    pos = np.zeros(13 * N, dtype=np.uint64)
    neg = np.zeros(13 * N, dtype=np.uint64)
    L.or_ring_split_witness(z.ctypes.data, N, pos.ctypes.data, neg.ctypes.data)
    return pos, neg


def ring_commit_ternary(key32, pos, neg, N):
    kw = ring_key_words(key32)
    p, n = _c(pos, np.uint64), _c(neg, np.uint64)
    out = np.zeros(13 * KAPPA * 64, dtype=np.uint64)
    L.or_ring_commit_ternary(kw.ctypes.data, p.ctypes.data, n.ctypes.data, N, out.ctypes.data)
    return out.reshape(13, KAPPA, 64)


def ring_materialize(key32, N):
    kw = ring_key_words(key32)
    out = np.zeros(N * KAPPA * 64, dtype=np.uint64)
    L.or_ring_materialize(kw.ctypes.data, N, out.ctypes.data)
    return out


def q_selective_add_planes(eq2, planes, total, n_planes, packed):
    e, pl = _c(eq2, np.uint64), _c(planes, np.uint64)
    out = np.zeros(2 * n_planes, dtype=np.uint64)
    L.or_q_selective_add_planes(e.ctypes.data, pl.ctypes.data, total, n_planes, packed, out.ctypes.data)
    return out.reshape(n_planes, 2)


# ---- zk-torch-4 transcript (Monolith-w12 over q) and the q sumcheck prover ----
import hashlib as _hashlib
import struct as _struct


def monoq_round_constants():
    """[5][12] round constants of rounds 1..5 (monolith.rs:67-110): the first 8 bytes
    (little-endian) of SHA-256("AGL-Monolith-RC-v1" || round || pos || salt), u32 LE each,
    for the first salt giving a value below q."""
    out = np.zeros((5, 12), dtype=np.uint64)
    for r in range(1, 6):
        for pos in range(12):
            salt = 0
            while True:
                d = _hashlib.sha256(b"AGL-Monolith-RC-v1" + _struct.pack("<III", r, pos, salt)).digest()
                v = int.from_bytes(d[:8], "little")
                if v < Q:
                    break
                salt += 1
            out[r - 1, pos] = v
    return out


MONOQ_RC = monoq_round_constants().ravel()


class QTr_t(ctypes.Structure):
    _fields_ = [("st", u64 * 12), ("inb", u64 * 8), ("outb", u64 * 8), ("nin", ctypes.c_int64),
                ("nout", ctypes.c_int64), ("rc", u64 * 60)]


_qtrp = ctypes.POINTER(QTr_t)
for _n, (_r, _a) in {
    "or_monoq_permute": (None, [u64p, u64p]),
    "or_qtr_init": (None, [_qtrp, ctypes.c_char_p, cl, u64p]),
    "or_qtr_append": (None, [_qtrp, ctypes.c_char_p, cl, u64p, cl]),
    "or_qtr_challenge": (None, [_qtrp, ctypes.c_char_p, cl, u64p, cl]),
    "or_q_sumcheck_prove": (None, [_qtrp, ci, u64p, ci, ci, u64p, u64p, u64p, u64p]),
}.items():
    _f = getattr(L, _n)
    _f.restype = _r
    _f.argtypes = _a


def monoq_permute(state12):
    s = arr(state12).copy()
    L.or_monoq_permute(_p(s), _p(MONOQ_RC))
    return s


class QTranscript:
    """zk-torch-4 Transcript (transcript.rs:21-240)."""

    def __init__(self, label=b""):
        self.t = QTr_t()
        L.or_qtr_init(ctypes.byref(self.t), label, len(label), _p(MONOQ_RC))

    def clone(self):
        c = QTranscript.__new__(QTranscript)
        c.t = QTr_t()
        ctypes.pointer(c.t)[0] = self.t
        return c

    def append_u64s(self, label, v):
        a = arr(np.atleast_1d(np.asarray(v, dtype=np.uint64)))
        L.or_qtr_append(ctypes.byref(self.t), label, len(label), _p(a), a.size)

    def append_ext2(self, label, v2):
        self.append_u64s(label, v2)

    def challenge(self, label, n):
        o = np.zeros(n, dtype=np.uint64)
        L.or_qtr_challenge(ctypes.byref(self.t), label, len(label), _p(o), n)
        return o

    def fork(self, family, branch_id):
        c = self.clone()
        c.append_u64s(family, [branch_id])
        return c

    def state(self):
        return np.array(self.t.st, dtype=np.uint64)


def q_sumcheck_prove(t, polys, np_, nv, ext):
    """CpuLinearSumcheckProverExt2 (cpu_ext2_prover.rs:22-100) on a QTranscript: returns
    (round messages [nv][np+1][2], challenges [nv][2], finals [np][2], final_eval [2])."""
    p = arr(polys)
    msgs = np.zeros(nv * (np_ + 1) * 2, dtype=np.uint64)
    chs = np.zeros(2 * nv, dtype=np.uint64)
    fin = np.zeros(2 * np_, dtype=np.uint64)
    fe = np.zeros(2, dtype=np.uint64)
    L.or_q_sumcheck_prove(ctypes.byref(t.t), 1 if ext else 0, _p(p), np_, nv, _p(msgs), _p(chs), _p(fin), _p(fe))
    return msgs.reshape(nv, np_ + 1, 2), chs.reshape(nv, 2), fin.reshape(np_, 2), fe
