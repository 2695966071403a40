"""Pins of the oracle's batch opening (or_basefold_batch_open_ext2 / _verify_ext2): the
verifier accepts honest proofs, the evaluation claims and the bridged inner evaluation match
direct MLE sums computed in Python integers from the committed evaluations, and tampering with
a claim, an outer oracle, a combined query value or an individual opening is rejected."""
import numpy as np
import pytest

import pyoracle as O

P = O.P
W = 7


def emul(a, b):
    return [(a[0] * b[0] + W * a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P]


def eadd(a, b):
    return [(a[0] + b[0]) % P, (a[1] + b[1]) % P]


def mle(evals, pt):
    """sum_x eq(pt, x) f(x), x_v = bit v of x, Python ints."""
    nv = len(pt)
    acc = [0, 0]
    for x in range(1 << nv):
        w = [1, 0]
        for v in range(nv):
            f = pt[v] if (x >> v) & 1 else [(1 - pt[v][0]) % P, (-pt[v][1]) % P]
            w = emul(w, f)
        acc = eadd(acc, [w[0] * int(evals[x]) % P, w[1] * int(evals[x]) % P])
    return acc


def eq_pt(r, pt):
    res = [1, 0]
    for a, b in zip(r, pt):
        res = emul(res, eadd(emul(a, b), emul([(1 - a[0]) % P, (-a[1]) % P], [(1 - b[0]) % P, (-b[1]) % P])))
    return res


def setup(nv, lr, npoly, npts, evals_spec, seed):
    rng = np.random.default_rng(seed)
    polys = [rng.integers(0, P, 1 << nv, dtype=np.uint64) for _ in range(npoly)]
    comms = []
    for f in polys:
        cw, bh, tree = O.basefold_commit(f, nv, lr, O.MONOLITH)
        comms.append((cw, bh, tree[-4:]))
    pts = [[[int(x) for x in rng.integers(0, P, 2, dtype=np.uint64)] for _ in range(nv)] for _ in range(npts)]
    evals = [(pi, qi, mle(polys[pi], pts[qi])) for pi, qi in evals_spec]
    return polys, comms, pts, evals


def replay_outer(comms, pts, evals, proof, nv):
    """Outer challenges r and batching weights from a replayed transcript."""
    t = O.Transcript(b"batch")
    for c in comms:
        t.append_scalars(b"", c[2])
    for pt in pts:
        for v in pt:
            t.append_ext2(b"", v)
    for e in evals:
        t.append_ext2(b"", e[2])
    ell = 1 if len(evals) <= 1 else (len(evals) - 1).bit_length()
    tch = [t.challenge_scalar(b"") for _ in range(ell)]
    w = [1]
    for ti in tch:  # first challenge on the most significant index bit
        w = [x for v in w for x in (v * (1 - ti) % P, v * ti % P)]
    r = []
    for rnd in range(nv):
        for k in range(3):
            t.append_ext2(b"", proof[6 * rnd + 2 * k:6 * rnd + 2 * k + 2])
        r.append([int(x) for x in t.challenge_ext2(b"")])
    return w, r


@pytest.mark.parametrize("nv,lr,npoly,npts,spec", [
    (4, 1, 3, 2, [(0, 0), (1, 0), (2, 1), (0, 1)]),
    (5, 1, 1, 1, [(0, 0)]),
    (3, 2, 2, 3, [(1, 2), (0, 0), (1, 1)]),
])
def test_batch_open_pinned(nv, lr, npoly, npts, spec):
    nq = 6
    polys, comms, pts, evals = setup(nv, lr, npoly, npts, spec, nv * 10 + npoly)
    table = O.basefold_table(nv, lr, nv, 7)
    proof = O.basefold_batch_open_ext2(comms, nv, lr, O.MONOLITH, pts, evals, table, nv, O.Transcript(b"batch"), nq)
    assert proof.size == O.basefold_batch_proof_len(nv, lr, nq, len(evals))
    roots = [c[2] for c in comms]

    def verify(pf, ev=evals):
        return O.basefold_batch_verify_ext2(roots, nv, lr, O.MONOLITH, pts, ev, pf, nq, table, nv,
                                            O.Transcript(b"batch"))

    assert verify(proof)
    # inner_eval = sum_e w_e eq(r, point_e) f_e(r), from Python integers
    w, r = replay_outer(comms, pts, evals, proof, nv)
    want = [0, 0]
    for e, (pi, qi, _) in enumerate(evals):
        want = eadd(want, emul([w[e], 0], emul(eq_pt(r, pts[qi]), mle(polys[pi], r))))
    off = 6 * nv + 4
    assert [int(proof[off]), int(proof[off + 1])] == want
    # tampering
    bad_evals = [(evals[0][0], evals[0][1], [(evals[0][2][0] + 1) % P, evals[0][2][1]])] + evals[1:]
    assert not verify(proof, bad_evals)
    for pos in (0, off + 2 + 6 * nv + 4 * (nv - 1) + 2 * (1 << lr) + 2, proof.size - 1):
        bad = proof.copy()
        bad[pos] ^= np.uint64(1 << 20)
        assert not verify(bad), pos
