"""zk-torch-4 (almost-Goldilocks field q = 2^64 - 2^32 - 31) prover API over the C-ABI.

  QTranscript       zk-torch-4/src/transcript.rs:21-240 (Monolith-w12 over q, rate 8)
  q_sumcheck_prove  zk-torch-4/src/sumcheck/gpu_prover.rs:26-112 (GpuLinearSumcheckProver)
  q_sumcheck_verify zk-torch-4/src/sumcheck/verifier.rs:15-86 (SumcheckVerifier::verify)
"""
import ctypes

import numpy as np

from ._lib import QTranscript_t, lib
from .device import check

Q = 0xFFFFFFFEFFFFFFE1
W = 3  # Ext2 = F_q[X] / (X^2 - 3)

_u64p = ctypes.POINTER(ctypes.c_uint64)


def _ptr(a):
    return a.ctypes.data_as(_u64p)


def _b(label):
    return label if isinstance(label, (bytes, bytearray)) else label.encode()


class QTranscript:
    """The zk-torch-4 Fiat-Shamir transcript (value type; ``clone`` copies it)."""

    def __init__(self, label=b""):
        self.t = QTranscript_t()
        lb = _b(label)
        lib.mam_qtranscript_init(ctypes.byref(self.t), lb, len(lb))

    def clone(self):
        c = QTranscript.__new__(QTranscript)
        c.t = QTranscript_t()
        ctypes.pointer(c.t)[0] = self.t
        return c

    def _append(self, label, vals):
        lb = _b(label)
        a = np.ascontiguousarray(np.atleast_1d(np.asarray(vals, dtype=np.uint64)))
        lib.mam_qtranscript_append_u64s(ctypes.byref(self.t), lb, len(lb), _ptr(a), a.size)

    # append_scalar / append_scalars / append_u64 / append_u64_slice / append_digest all absorb
    # the label and then the values reduced mod q (transcript.rs:65-113)
    def append_u64(self, label, v):
        self._append(label, [int(v)])

    append_scalar = append_u64

    def append_scalars(self, label, vals):
        self._append(label, vals)

    append_u64_slice = append_scalars
    append_digest = append_scalars

    def append_ext2(self, label, v2):
        self._append(label, v2)

    def _challenge(self, label, n):
        lb = _b(label)
        out = np.zeros(max(n, 1), dtype=np.uint64)
        lib.mam_qtranscript_challenge(ctypes.byref(self.t), lb, len(lb), _ptr(out), n)
        return out[:n]

    def challenge_scalar(self, label=b""):
        return int(self._challenge(label, 1)[0])

    def challenge_ext2(self, label=b""):
        return self._challenge(label, 2)

    def challenge_vector(self, label, n):
        return self._challenge(label, n)

    def fork(self, family, branch_id):
        c = QTranscript.__new__(QTranscript)
        c.t = QTranscript_t()
        fb = _b(family)
        lib.mam_qtranscript_fork(ctypes.byref(self.t), fb, len(fb), int(branch_id), ctypes.byref(c.t))
        return c

    def fingerprint(self):
        return int(lib.mam_qtranscript_fingerprint(ctypes.byref(self.t)))

    def get_state(self):
        return np.array(self.t.state, dtype=np.uint64)


def monolith_q_permute(state12):
    s = np.ascontiguousarray(state12, dtype=np.uint64).copy()
    assert s.size == 12
    lib.mam_monolith_q_permute(_ptr(s))
    return s


def q_sumcheck_prove(transcript, d_polys, num_polys, num_vars, base=True):
    """Whole linear sumcheck proof over the q Ext2 on the GPU, zk-torch-4 transcript.

    ``d_polys``: DeviceBuffer with ``num_polys`` polynomials of ``2**num_vars`` elements back
    to back (base words when ``base`` (prove), else interleaved Ext2 (prove_ext2)).
    Returns dict(round_msgs [nv, np+1, 2], challenges [nv, 2], final_evals [np, 2], final_eval [2]).
    """
    msgs = np.zeros((num_vars, num_polys + 1, 2), dtype=np.uint64)
    chs = np.zeros((num_vars, 2), dtype=np.uint64)
    fin = np.zeros((num_polys, 2), dtype=np.uint64)
    fe = np.zeros(2, dtype=np.uint64)
    check(lib.mam_q_sumcheck_prove(ctypes.byref(transcript.t), d_polys.ptr, num_polys, num_vars, 1 if base else 0,
                                   _ptr(msgs), _ptr(chs), _ptr(fin), _ptr(fe)), "mam_q_sumcheck_prove")
    return {"round_msgs": msgs, "challenges": chs, "final_evals": fin, "final_eval": fe}


# ---- Ext2 over q on Python integers (c0 + c1 X, X^2 = 3) for the host verifier ----
def q2add(a, b):
    return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)


def q2sub(a, b):
    return ((a[0] - b[0]) % Q, (a[1] - b[1]) % Q)


def q2mul(a, b):
    return ((a[0] * b[0] + W * a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def q_interp_eval(evals, x):
    """The polynomial through (0, evals[0]), ..., (d, evals[d]) at x (verifier.rs:92-118)."""
    d = len(evals)
    res = (0, 0)
    for i in range(d):
        num, den = (1, 0), 1
        for j in range(d):
            if j != i:
                num = q2mul(num, q2sub(x, (j, 0)))
                den = den * (i - j) % Q
        res = q2add(res, q2mul(evals[i], q2mul(num, (pow(den, Q - 2, Q), 0))))
    return res


def q_sumcheck_verify(transcript, claimed_sum, num_polys, num_vars, round_msgs, final_eval):
    """SumcheckVerifier::verify over q: absorbs num_var and num_poly, per round checks
    s(0) + s(1) against the running claim, absorbs the evaluations, draws the challenge and
    moves the claim to s(challenge); finally compares with final_eval. Returns (accepted,
    challenges)."""
    transcript.append_u64(b"num_var", num_vars)
    transcript.append_u64(b"num_poly", num_polys)
    msgs = np.asarray(round_msgs, dtype=np.uint64).reshape(-1, num_polys + 1, 2)
    if msgs.shape[0] != num_vars:
        return False, []
    cur = (int(claimed_sum[0]) % Q, int(claimed_sum[1]) % Q)
    chs = []
    for r in range(num_vars):
        ev = [(int(x) % Q, int(y) % Q) for x, y in msgs[r]]
        if q2add(ev[0], ev[1]) != cur:
            return False, []
        for e in msgs[r]:
            transcript.append_ext2(b"round_message", e)
        ch = transcript.challenge_ext2(b"challenge")
        chs.append(ch)
        cur = q_interp_eval(ev, (int(ch[0]), int(ch[1])))
    return cur == (int(final_eval[0]) % Q, int(final_eval[1]) % Q), chs
