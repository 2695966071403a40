"""Prover-level API over the C-ABI, shaped like the reference's Rust types.

  Transcript              zk-torch-3/src/transcript.rs:167-330
  sumcheck_prove          zk-torch-3/src/sumcheck/gpu_prover.rs:21-89 (GpuLinearSumcheckProver)
  GpuSumcheckStateExt2    goldilocks-cuda-rs/src/sumcheck_prover.rs:205-423
  BasefoldTable           goldilocks-cuda-rs/src/basefold.rs:281-360
  BasefoldCommitment      goldilocks-cuda-rs/src/basefold.rs:366-863
  sumcheck_verify         zk-torch-3/src/sumcheck/verifier.rs:17-67 (SumcheckVerifier::verify)
  verify_ext2             goldilocks-cuda-rs/src/basefold.rs:1759-1925
  batch_open_ext2         goldilocks-cuda-rs/src/basefold.rs:1202-1525
  batch_verify_ext2       goldilocks-cuda-rs/src/basefold.rs:1931-2221
"""
import ctypes

import numpy as np

from ._lib import Transcript_t, lib
from .device import DeviceBuffer, check

P = 0xFFFFFFFF00000001
HASH_MONOLITH = 0
HASH_POSEIDON2 = 1

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _ptr(a):
    return a.ctypes.data_as(_u64p)


def _b(label):
    return label if isinstance(label, (bytes, bytearray)) else label.encode()


class Transcript:
    """Poseidon2-w8 duplex Fiat-Shamir transcript (value type; ``clone`` copies it)."""

    def __init__(self, label=b""):
        self.t = Transcript_t()
        lb = _b(label)
        lib.mam_transcript_init(ctypes.byref(self.t), lb, len(lb))

    def clone(self):
        c = Transcript.__new__(Transcript)
        c.t = Transcript_t()
        ctypes.pointer(c.t)[0] = self.t
        return c

    def append_u64(self, label, v):
        lb = _b(label)
        lib.mam_transcript_append_u64(ctypes.byref(self.t), lb, len(lb), int(v))

    def append_scalars(self, label, vals):
        lb = _b(label)
        a = np.ascontiguousarray(vals, dtype=np.uint64)
        lib.mam_transcript_append_scalars(ctypes.byref(self.t), lb, len(lb), _ptr(a), a.size)

    def append_ext2(self, label, v2):
        lb = _b(label)
        a = np.ascontiguousarray(v2, dtype=np.uint64)
        lib.mam_transcript_append_ext2(ctypes.byref(self.t), lb, len(lb), _ptr(a))

    def challenge_scalar(self, label=b""):
        lb = _b(label)
        return int(lib.mam_transcript_challenge_scalar(ctypes.byref(self.t), lb, len(lb)))

    def challenge_ext2(self, label=b""):
        lb = _b(label)
        out = np.zeros(2, dtype=np.uint64)
        lib.mam_transcript_challenge_ext2(ctypes.byref(self.t), lb, len(lb), _ptr(out))
        return out

    def state(self):
        return np.array(self.t.state, dtype=np.uint64)


def transcript_program(ops):
    """Encode [("append", label, values) | ("challenge", label, n), ...] for the device
    transcript: per operation [kind, label_len, label words, n]."""
    words, src = [], []
    for op in ops:
        kind, label = op[0], _b(op[1])
        lw = (len(label) + 7) // 8
        packed = np.frombuffer(label + bytes(8 * lw - len(label)), dtype="<u8") if lw else np.zeros(0, np.uint64)
        if kind == "append":
            vals = [int(v) for v in np.atleast_1d(np.asarray(op[2], dtype=np.uint64))]
            words += [0, len(label)] + [int(v) for v in packed] + [len(vals)]
            src += vals
        elif kind == "challenge":
            words += [1, len(label)] + [int(v) for v in packed] + [int(op[2])]
        else:
            raise ValueError(f"unknown transcript operation {kind!r}")
    return np.array(words, dtype=np.uint64), np.array(src, dtype=np.uint64)


class DeviceTranscript:
    """The zk-torch-3 transcript held in device memory; ``run`` queues a program of labelled
    appends and challenges for one wavefront (no host round trip per operation)."""

    def __init__(self, transcript=None):
        t = transcript if transcript is not None else Transcript()
        self.state = DeviceBuffer.from_numpy(self._words(t))

    @staticmethod
    def _words(t):
        return np.frombuffer(bytes(t.t), dtype=np.uint64).copy()

    def run(self, ops):
        prog, src = transcript_program(ops)
        n_out = sum(int(op[2]) for op in ops if op[0] == "challenge")
        dp = DeviceBuffer.from_numpy(prog) if prog.size else None
        ds = DeviceBuffer.from_numpy(src) if src.size else None
        dd = DeviceBuffer(max(n_out, 1))
        err = DeviceBuffer(1)
        check(lib.mam_dtranscript_run(self.state.ptr, dp.ptr if dp else None, prog.size, ds.ptr if ds else None,
                                      src.size, dd.ptr, n_out, err.ptr), "mam_dtranscript_run")
        e = int(err.to_numpy()[0]) & 0xFFFFFFFF
        if e:
            raise ValueError("device transcript program was rejected")
        return dd.to_numpy()[:n_out]

    def to_host(self):
        t = Transcript()
        ctypes.memmove(ctypes.addressof(t.t), self.state.to_numpy().ctypes.data, ctypes.sizeof(t.t))
        return t


def sumcheck_prove(transcript, d_polys, num_polys, num_vars, base=True):
    """Whole linear sumcheck proof over Ext2 on the GPU.

    ``d_polys``: DeviceBuffer with ``num_polys`` polynomials of ``2**num_vars`` elements
    back to back (base words when ``base``, else interleaved Ext2).
    Returns dict(round_msgs [nv, np+1, 2], challenges [nv, 2], final_evals [np, 2], final_eval [2]).
    """
    msgs = np.zeros((num_vars, num_polys + 1, 2), dtype=np.uint64)
    chs = np.zeros((num_vars, 2), dtype=np.uint64)
    fin = np.zeros((num_polys, 2), dtype=np.uint64)
    fe = np.zeros(2, dtype=np.uint64)
    check(lib.mam_sumcheck_prove(ctypes.byref(transcript.t), d_polys.ptr, num_polys, num_vars, 1 if base else 0,
                                 _ptr(msgs), _ptr(chs), _ptr(fin), _ptr(fe)), "mam_sumcheck_prove")
    return {"round_msgs": msgs, "challenges": chs, "final_evals": fin, "final_eval": fe}


def sumcheck_verify(transcript, claimed_sum, num_polys, num_vars, round_msgs, final_eval):
    """SumcheckVerifier::verify: absorbs num_var and num_poly, then per round checks
    s(0) + s(1) against the running claim, absorbs the num_polys + 1 evaluations and draws the
    challenge, and moves the claim to s(challenge); finally compares with final_eval.
    round_msgs: [num_vars][num_polys + 1][2] words. Returns (accepted, challenges)."""
    from .hostfield import e2add, interp_eval
    transcript.append_u64(b"num_var", num_vars)
    transcript.append_u64(b"num_poly", num_polys)
    msgs = np.asarray(round_msgs, dtype=np.uint64).reshape(-1, num_polys + 1, 2)
    if msgs.shape[0] != num_vars:
        return False, []
    cur = (int(claimed_sum[0]) % P, int(claimed_sum[1]) % P)
    chs = []
    for r in range(num_vars):
        ev = [(int(x) % P, int(y) % P) for x, y in msgs[r]]
        if e2add(ev[0], ev[1]) != cur:
            return False, chs
        for e in msgs[r]:
            transcript.append_ext2(b"round_message", e)
        ch = transcript.challenge_ext2(b"challenge")
        chs.append(ch)
        cur = interp_eval(ev, (int(ch[0]), int(ch[1])))
    return cur == (int(final_eval[0]) % P, int(final_eval[1]) % P), chs


class GpuSumcheckStateExt2:
    """Round-by-round sumcheck state (compute_round_message / fold / final_evaluations)."""

    def __init__(self, d_polys, num_polys, num_vars, base=True):
        self._src = d_polys  # keep the input alive: the state reads it until the first fold
        self.num_polys = num_polys
        h = ctypes.c_void_p()
        check(lib.mam_sc_state_new(d_polys.ptr, num_polys, num_vars, 1 if base else 0, ctypes.byref(h)),
              "mam_sc_state_new")
        self.h = h.value

    def compute_round_message(self):
        out = np.zeros((self.num_polys + 1) * 2, dtype=np.uint64)
        check(lib.mam_sc_state_round_message(self.h, _ptr(out)), "mam_sc_state_round_message")
        return out.reshape(-1, 2)

    def fold(self, challenge):
        check(lib.mam_sc_state_fold(self.h, int(challenge[0]), int(challenge[1])), "mam_sc_state_fold")

    def final_evaluations(self):
        out = np.zeros(self.num_polys * 2, dtype=np.uint64)
        check(lib.mam_sc_state_final_evaluations(self.h, _ptr(out)), "mam_sc_state_final_evaluations")
        return out.reshape(-1, 2)

    def free(self):
        if self.h:
            lib.mam_sc_state_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class BasefoldTable:
    """Folding table: host entries (point, weight) and their device copy."""

    def __init__(self, num_vars, log_rate, num_rounds, seed):
        self.num_vars, self.log_rate, self.num_rounds = num_vars, log_rate, num_rounds
        n = lib.mam_basefold_table_len(num_vars, log_rate, num_rounds)
        self.entries = np.zeros(2 * n, dtype=np.uint64)
        check(lib.mam_basefold_table_generate(num_vars, log_rate, num_rounds, seed, _ptr(self.entries)),
              "mam_basefold_table_generate")
        self.d_entries = None

    @classmethod
    def generate(cls, num_vars, log_rate, num_rounds, seed):
        return cls(num_vars, log_rate, num_rounds, seed)

    def upload(self):
        if self.d_entries is None:
            self.d_entries = DeviceBuffer.from_numpy(self.entries)
        return self.d_entries


def proof_len(num_vars, log_rate, num_queries):
    return int(lib.mam_basefold_proof_len(num_vars, log_rate, num_queries))


def split_proof(proof, num_vars, log_rate, num_queries):
    """Split the flat proof into its fields (same layout as the C-ABI documents)."""
    nv, lr = num_vars, log_rate
    p = 0
    out = {"eval": proof[p:p + 2]}
    p += 2
    out["oracles"] = proof[p:p + 6 * nv].reshape(nv, 3, 2)
    p += 6 * nv
    out["folded_roots"] = proof[p:p + 4 * (nv - 1)].reshape(nv - 1, 4)
    p += 4 * (nv - 1)
    out["final_codeword"] = proof[p:p + 2 * (1 << lr)].reshape(-1, 2)
    p += 2 * (1 << lr)
    qs = []
    for _ in range(num_queries):
        q = {"index": int(proof[p])}
        p += 1
        q["values"] = proof[p:p + 4 * nv].reshape(nv, 2, 2)
        p += 4 * nv
        paths = []
        for k in range(nv):
            d = nv + lr - 1 - k
            paths.append(proof[p:p + 4 * d].reshape(d, 4))
            p += 4 * d
        q["paths"] = paths
        qs.append(q)
    out["queries"] = qs
    assert p == len(proof)
    return out


class BasefoldCommitment:
    """Commitment whose codeword, bh_evals and Merkle tree live on the GPU."""

    def __init__(self, handle, root, num_vars, log_rate, hash_kind):
        self.h, self.root = handle, root
        self.num_vars, self.log_rate, self.hash = num_vars, log_rate, hash_kind

    @classmethod
    def commit_device(cls, d_evals, num_vars, log_rate, hash_kind=HASH_MONOLITH):
        h = ctypes.c_void_p()
        root = np.zeros(4, dtype=np.uint64)
        check(lib.mam_basefold_commit(d_evals.ptr, num_vars, log_rate, hash_kind, ctypes.byref(h), _ptr(root)),
              "mam_basefold_commit")
        return cls(h.value, root, num_vars, log_rate, hash_kind)

    @classmethod
    def commit(cls, evals, num_vars, log_rate, hash_kind=HASH_MONOLITH):
        d = DeviceBuffer.from_numpy(evals)
        c = cls.commit_device(d, num_vars, log_rate, hash_kind)
        d.free()
        return c

    def buffers(self):
        cw, bh, tr = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        check(lib.mam_basefold_commitment_buffers(self.h, ctypes.byref(cw), ctypes.byref(bh), ctypes.byref(tr),
                                                  None), "mam_basefold_commitment_buffers")
        return cw.value, bh.value, tr.value

    def to_host_cache(self):
        """Codeword and bh_evals on the host (HostCommitmentCache, basefold.rs:873-900)."""
        from ._lib import lib as L
        cw, bh, _ = self.buffers()
        n_cw = 1 << (self.num_vars + self.log_rate)
        a = np.empty(n_cw, dtype=np.uint64)
        b = np.empty(1 << self.num_vars, dtype=np.uint64)
        check(L.cuda_memcpy_dtoh(a.ctypes.data, cw, a.size * 8), "cuda_memcpy_dtoh")
        check(L.cuda_memcpy_dtoh(b.ctypes.data, bh, b.size * 8), "cuda_memcpy_dtoh")
        return a, b

    def open_ext2(self, point2, table, transcript, num_queries):
        point2 = np.ascontiguousarray(point2, dtype=np.uint64).reshape(-1)
        assert point2.size == 2 * self.num_vars
        d_table = table.upload()
        proof = np.zeros(proof_len(self.num_vars, self.log_rate, num_queries), dtype=np.uint64)
        check(lib.mam_basefold_open_ext2(self.h, _ptr(point2), d_table.ptr, table.num_vars,
                                         ctypes.byref(transcript.t), num_queries, _ptr(proof)),
              "mam_basefold_open_ext2")
        return proof

    def free(self):
        if self.h:
            lib.mam_basefold_commitment_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def verify_ext2(root, point2, num_vars, log_rate, hash_kind, proof, num_queries, table, transcript):
    root = np.ascontiguousarray(root, dtype=np.uint64)
    point2 = np.ascontiguousarray(point2, dtype=np.uint64).reshape(-1)
    proof = np.ascontiguousarray(proof, dtype=np.uint64)
    return bool(lib.mam_basefold_verify_ext2(_ptr(root), _ptr(point2), num_vars, log_rate, hash_kind, _ptr(proof),
                                             num_queries, _ptr(table.entries), table.num_vars,
                                             ctypes.byref(transcript.t)))


class EvaluationExt2:
    """One claim: poly[poly] at points[point] = value (basefold.rs:148-160)."""

    def __init__(self, poly, point, value):
        self.poly, self.point = int(poly), int(point)
        self.value = [int(value[0]), int(value[1])]


def _claims(evals):
    from ._lib import EvalClaim_t
    arr = (EvalClaim_t * len(evals))()
    for i, e in enumerate(evals):
        arr[i].poly, arr[i].point = e.poly, e.point
        arr[i].value[0], arr[i].value[1] = e.value
    return arr


def batch_proof_len(num_vars, log_rate, num_queries, num_evals):
    return int(lib.mam_basefold_batch_proof_len(num_vars, log_rate, num_queries, num_evals))


def batch_open_ext2(comms, points, evals, table, transcript, num_queries):
    """One proof for every claim; comms share num_vars, log_rate and hash (basefold.rs:1202)."""
    nv, lr = comms[0].num_vars, comms[0].log_rate
    pts = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1)
    assert pts.size % (2 * nv) == 0
    handles = (ctypes.c_void_p * len(comms))(*[c.h for c in comms])
    proof = np.zeros(batch_proof_len(nv, lr, num_queries, len(evals)), dtype=np.uint64)
    d_table = table.upload()
    check(lib.mam_basefold_batch_open_ext2(handles, len(comms), _ptr(pts), pts.size // (2 * nv), _claims(evals),
                                           len(evals), d_table.ptr, table.num_vars, ctypes.byref(transcript.t),
                                           num_queries, _ptr(proof)), "mam_basefold_batch_open_ext2")
    return proof


def batch_verify_ext2(roots, points, evals, proof, table, transcript, num_vars, log_rate, num_queries,
                      hash_kind=HASH_MONOLITH):
    roots = np.ascontiguousarray(roots, dtype=np.uint64).reshape(-1)
    pts = np.ascontiguousarray(points, dtype=np.uint64).reshape(-1)
    proof = np.ascontiguousarray(proof, dtype=np.uint64)
    return bool(lib.mam_basefold_batch_verify_ext2(_ptr(roots), roots.size // 4, num_vars, log_rate, hash_kind,
                                                   _ptr(pts), pts.size // (2 * num_vars), _claims(evals), len(evals),
                                                   _ptr(proof), num_queries, _ptr(table.entries), table.num_vars,
                                                   ctypes.byref(transcript.t)))


def sumcheck_proof_bincode(proof):
    """bincode 1.x bytes of zk-torch-3's SumcheckProof for a sumcheck_prove result."""
    msgs = np.ascontiguousarray(proof["round_msgs"], dtype=np.uint64)
    fe = np.ascontiguousarray(proof["final_eval"], dtype=np.uint64)
    rounds, entries = msgs.shape[0], msgs.shape[1]
    n = int(lib.mam_sumcheck_proof_bincode_len(rounds, entries))
    out = (ctypes.c_uint8 * max(n, 1))()
    got = lib.mam_sumcheck_proof_bincode(_ptr(msgs), rounds, entries, _ptr(fe), out, n)
    check(0 if got == n else -1, "mam_sumcheck_proof_bincode")
    return bytes(out)[:n]


def fingerprint(data):
    """ser_fingerprint (zk-torch-4/src/lib.rs:151-161): SHA-256 of the bytes, first 8 bytes as hex."""
    import hashlib
    return hashlib.sha256(data).digest()[:8].hex()
