"""Matmul proof C = A B (include/morpheus_prover.h, mam_matmul_prove): Basefold commitments
of A, B, C, an optional int8 Ajtai commitment of C, the claim C(r), one degree-3 sumcheck over
the inner index, and Basefold openings of A, B and C on one transcript."""
import ctypes

import numpy as np

from ._lib import MatmulInstance_t, ServeRequest_t, lib
from .device import DeviceBuffer, check
from .prover import proof_len

_u64p = ctypes.POINTER(ctypes.c_uint64)
P = 0xFFFFFFFF00000001


def lift_int(x):
    """Signed integers -> canonical Goldilocks words."""
    x = np.asarray(x, dtype=np.int64)
    return np.where(x < 0, np.uint64(P) - (-x).astype(np.uint64), x.astype(np.uint64)).astype(np.uint64)


class MatmulJob:
    """One statement with A, B, C resident on the device (canonical words, row-major)."""

    def __init__(self, lm, lk, ln, d_a, d_b, d_c, label):
        self.lm, self.lk, self.ln = lm, lk, ln
        self.d_a, self.d_b, self.d_c = d_a, d_b, d_c
        self.label = bytes(label)

    @classmethod
    def from_int8(cls, a_i8, b_i8, label):
        """A (m x k) and B (k x n) int8 arrays; C = A B computed exactly on the host."""
        a = np.asarray(a_i8, dtype=np.int64)
        b = np.asarray(b_i8, dtype=np.int64)
        if a.min(initial=0) < -128 or a.max(initial=0) > 127 or b.min(initial=0) < -128 or b.max(initial=0) > 127:
            raise ValueError("from_int8 takes entries in [-128, 127]")
        # float64 BLAS is exact here: |c| <= 2^14 k < 2^53 for any k below 2^39
        c = (a.astype(np.float64) @ b.astype(np.float64)).astype(np.int64)
        lm, lk = (a.shape[0] - 1).bit_length(), (a.shape[1] - 1).bit_length()
        ln = (b.shape[1] - 1).bit_length()
        return cls(lm, lk, ln, DeviceBuffer.from_numpy(lift_int(a).ravel()),
                   DeviceBuffer.from_numpy(lift_int(b).ravel()), DeviceBuffer.from_numpy(lift_int(c).ravel()), label)

    def outputs(self, log_rate, num_queries, ajtai_rows):
        lm, lk, ln = self.lm, self.lk, self.ln
        o = {"roots": np.zeros((3, 4), dtype=np.uint64), "point": np.zeros((ln + lm, 2), dtype=np.uint64),
             "claim": np.zeros(2, dtype=np.uint64), "round_msgs": np.zeros((lk, 4, 2), dtype=np.uint64),
             "challenges": np.zeros((lk, 2), dtype=np.uint64), "final_evals": np.zeros((3, 2), dtype=np.uint64),
             "proof_a": np.zeros(proof_len(lm + lk, log_rate, num_queries), dtype=np.uint64),
             "proof_b": np.zeros(proof_len(lk + ln, log_rate, num_queries), dtype=np.uint64),
             "proof_c": np.zeros(proof_len(lm + ln, log_rate, num_queries), dtype=np.uint64)}
        if ajtai_rows:
            o["ajtai"] = np.zeros(ajtai_rows, dtype=np.uint64)
        return o

    def fill(self, s, o):
        s.d_a, s.d_b, s.d_c = self.d_a.ptr, self.d_b.ptr, self.d_c.ptr
        s.lm, s.lk, s.ln = self.lm, self.lk, self.ln
        s.label, s.label_len = self.label, len(self.label)
        for k in ("roots", "point", "claim", "round_msgs", "challenges", "final_evals", "proof_a", "proof_b",
                  "proof_c"):
            setattr(s, k, o[k].ctypes.data_as(_u64p))
        s.ajtai = o["ajtai"].ctypes.data_as(_u64p) if "ajtai" in o else None


def prove_matmul(job, table, log_rate=1, hash_kind=0, num_queries=32, ajtai=None):
    """Returns the proof fields as numpy arrays. ajtai: an AjtaiI8 set-up with at least |C|
    columns (or None to skip the Ajtai commitment)."""
    rows = ajtai.rows if ajtai is not None else 0
    o = job.outputs(log_rate, num_queries, rows)
    s = MatmulInstance_t()
    job.fill(s, o)
    check(lib.mam_matmul_prove(ctypes.byref(s), table.upload().ptr, table.num_vars, log_rate, hash_kind, num_queries,
                               ajtai.h if ajtai is not None else None, rows), "mam_matmul_prove")
    return o


def serve(jobs, arrivals_ms, table, log_rate=1, hash_kind=0, num_queries=32, ajtai=None, workers=4,
          hbm_budget=64 << 30):
    """Run the proving service over jobs released at arrivals_ms (offsets from the call).
    Returns (outputs per job, admit_ms, done_ms arrays)."""
    rows = ajtai.rows if ajtai is not None else 0
    n = len(jobs)
    insts = (MatmulInstance_t * max(n, 1))()
    reqs = (ServeRequest_t * max(n, 1))()
    outs = []
    for i, job in enumerate(jobs):
        o = job.outputs(log_rate, num_queries, rows)
        outs.append(o)
        job.fill(insts[i], o)
        reqs[i].inst = ctypes.pointer(insts[i])
        reqs[i].arrival_ms = float(arrivals_ms[i])
    check(lib.mam_serve_matmul(reqs, n, table.upload().ptr, table.num_vars, log_rate, hash_kind, num_queries,
                               ajtai.h if ajtai is not None else None, rows, workers, int(hbm_budget)),
          "mam_serve_matmul")
    admit = np.array([reqs[i].admit_ms for i in range(n)])
    done = np.array([reqs[i].done_ms for i in range(n)])
    return outs, admit, done


def verify_matmul(o, lm, lk, ln, label, table, log_rate=1, hash_kind=0, num_queries=32):
    """Host verifier of a matmul proof (the fields prove_matmul returns): replays the transcript
    (roots of A, B, C, the Ajtai commitment when present, the point, the claim), checks the
    degree-3 sumcheck over the inner index against the claim C(r) ending at A(r_l, r_i) B(r_j, r_l),
    and checks the Basefold openings of A at (r_l, r_i), B at (r_j, r_l) and C at r, each
    opening's evaluation being the value the sumcheck relies on."""
    from .hostfield import e2mul
    from .prover import Transcript, sumcheck_verify, verify_ext2

    def pair(v):
        return (int(v[0]), int(v[1]))
    t = Transcript(label)
    for root in o["roots"]:
        t.append_scalars(b"matmul-root", root)
    if "ajtai" in o:
        t.append_scalars(b"matmul-ajtai", o["ajtai"])
    point = np.array([t.challenge_ext2(b"matmul-point") for _ in range(ln + lm)], dtype=np.uint64)
    if not np.array_equal(point, o["point"]):
        return False
    claim = np.asarray(o["claim"], dtype=np.uint64)
    t.append_ext2(b"matmul-claim", claim)
    fe = np.asarray(o["final_evals"], dtype=np.uint64)
    if pair(fe[2]) != (1, 0):
        return False
    ok, chs = sumcheck_verify(t, claim, 3, lk, o["round_msgs"], e2mul(pair(fe[0]), pair(fe[1])))
    if not ok or not np.array_equal(np.array(chs, dtype=np.uint64).reshape(-1, 2), o["challenges"]):
        return False
    r_j, r_i, r_l = point[:ln], point[ln:], np.asarray(o["challenges"], dtype=np.uint64)
    opens = (("proof_a", lm + lk, np.concatenate([r_l, r_i]), fe[0]),
             ("proof_b", lk + ln, np.concatenate([r_j, r_l]), fe[1]),
             ("proof_c", lm + ln, point, claim))
    for (name, nv, pt, ev), root in zip(opens, o["roots"]):
        if pair(o[name][:2]) != pair(ev):
            return False
        if not verify_ext2(root, pt.ravel(), nv, log_rate, hash_kind, o[name], num_queries, table, t):
            return False
    return True
