"""Pins of the oracle's matmul proof composition (pyoracle.matmul_prove): the claim and the
folded A', B' against closed forms computed with Python integers from the integer matrices,
the identity C(r) = sum_l A'(l) B'(l), and the oracle's own verifiers (sumcheck and the three
Basefold openings) accepting the proof on a replayed transcript and rejecting a changed one."""
import numpy as np
import pytest

import pyoracle as O

P = O.P
W = 7  # Ext2 = F_p[X] / (X^2 - 7)


def emul(a, b):
    return [(a[0] * b[0] + W * a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P]


def eq_weights(r2):
    """eq(r, x) for x in [0, 2^len(r)), x_v = bit v of x (LSB first), Python ints."""
    nv = len(r2) // 2
    out = []
    for x in range(1 << nv):
        acc = [1, 0]
        for v in range(nv):
            rv = [int(r2[2 * v]), int(r2[2 * v + 1])]
            f = rv if (x >> v) & 1 else [(1 - rv[0]) % P, (-rv[1]) % P]
            acc = emul(acc, f)
        out.append(acc)
    return out


def mle(vals, r2):
    s = [0, 0]
    for w, v in zip(eq_weights(r2), vals):
        s = [(s[0] + w[0] * v) % P, (s[1] + w[1] * v) % P]
    return s


@pytest.mark.parametrize("dims,ajtai", [((2, 3, 2), False), ((3, 2, 3), True)])
def test_matmul_proof_composition(dims, ajtai):
    lm, lk, ln = dims
    m_, k_, n_ = 1 << lm, 1 << lk, 1 << ln
    rng = np.random.default_rng(sum(dims) + 40)
    a = rng.integers(-128, 128, (m_, k_))
    b = rng.integers(-128, 128, (k_, n_))
    c = a @ b
    lr, nq = 1, 5
    tnv = max(lm + lk, lk + ln, lm + ln)
    table = O.basefold_table(tnv, lr, tnv, 3)
    key = bytes(range(32)) if ajtai else None
    pr = O.matmul_prove(a, b, b"mm-pin", table, tnv, lr, nq, key, 16)
    point = pr["point"].ravel()
    r_j, r_i = point[:2 * ln], point[2 * ln:]
    # claim = C(r), r = (r_j, r_i) over the row-major index i * n + j
    cw = [int(x) % P for x in c.ravel()]
    assert [int(x) for x in pr["claim"]] == mle(cw, point)
    # A'(l) = sum_i eq(r_i, i) A[i, l], B'(l) = sum_j eq(r_j, j) B[l, j]; C(r) = sum_l A'(l) B'(l)
    wi, wj = eq_weights(r_i), eq_weights(r_j)
    a1 = [[sum(wi[i][e] * (int(a[i, l]) % P) for i in range(m_)) % P for e in range(2)] for l in range(k_)]
    b1 = [[sum(wj[j][e] * (int(b[l, j]) % P) for j in range(n_)) % P for e in range(2)] for l in range(k_)]
    tot = [0, 0]
    for l in range(k_):
        pl = emul(a1[l], b1[l])
        tot = [(tot[0] + pl[0]) % P, (tot[1] + pl[1]) % P]
    assert tot == [int(x) for x in pr["claim"]]
    # the sumcheck's final evaluations are A'(r_l), B'(r_l), 1
    r_l = pr["challenges"].ravel()
    wl = eq_weights(r_l)
    for f, col in ((a1, 0), (b1, 1)):
        s = [0, 0]
        for l in range(k_):
            pl = emul(wl[l], f[l])
            s = [(s[0] + pl[0]) % P, (s[1] + pl[1]) % P]
        assert [int(x) for x in pr["final_evals"][col]] == s
    assert [int(x) for x in pr["final_evals"][2]] == [1, 0]

    def verify(proof_b):
        t = O.Transcript(b"mm-pin")
        for x in range(3):
            t.append_scalars(b"matmul-root", pr["roots"][x])
        if ajtai:
            assert np.array_equal(pr["ajtai"], O.ajtai_i8_commit(key, 16, 1 << (lm + ln), O.lift_int(c).ravel(), 1))
            t.append_scalars(b"matmul-ajtai", pr["ajtai"])
        pt = np.array([t.challenge_ext2(b"matmul-point") for _ in range(ln + lm)], dtype=np.uint64)
        assert np.array_equal(pt, pr["point"])
        t.append_ext2(b"matmul-claim", pr["claim"])
        fe = emul(emul([int(x) for x in pr["final_evals"][0]], [int(x) for x in pr["final_evals"][1]]),
                  [int(x) for x in pr["final_evals"][2]])
        ok = O.sumcheck_verify_ext2(t, pr["claim"], 3, lk, pr["round_msgs"], np.array(fe, dtype=np.uint64))
        pts = (np.concatenate([r_l, r_i]), np.concatenate([r_j, r_l]), point)
        nvs = (lm + lk, lk + ln, lm + ln)
        for x, (proof, nv, pt2) in enumerate(zip((pr["proof_a"], proof_b, pr["proof_c"]), nvs, pts)):
            ok = O.basefold_verify_ext2(pr["roots"][x], pt2, nv, lr, O.MONOLITH, proof, nq, table, tnv, t) and ok
        return ok

    assert verify(pr["proof_b"])
    # a changed word in B's opening (a query path word near the end) is rejected
    bad = pr["proof_b"].copy()
    bad[-3] ^= np.uint64(1)
    assert not verify(bad)


def test_host_matmul_verifier_on_oracle_proofs():
    """morpheus_amd.matmul.verify_matmul (host code of the library) accepts the oracle's matmul
    proofs and rejects changed claim, round-message and opening words."""
    import morpheus_amd as M
    from morpheus_amd.matmul import verify_matmul
    rng = np.random.default_rng(44)
    for (lm, lk, ln), key in (((3, 4, 3), bytes(range(5, 37))), ((2, 5, 4), None)):
        a = rng.integers(-128, 128, (1 << lm, 1 << lk))
        b = rng.integers(-128, 128, (1 << lk, 1 << ln))
        tnv = max(lm + lk, lk + ln, lm + ln)
        table = M.BasefoldTable.generate(tnv, 1, tnv, 42)
        o = O.matmul_prove(a, b, b"hv-mm", table.entries, tnv, 1, 9, key, 64)
        assert verify_matmul(o, lm, lk, ln, b"hv-mm", table, 1, 0, 9)
        for field, idx in (("claim", (0,)), ("round_msgs", (lk - 1, 3, 1)), ("proof_b", (40,)), ("proof_c", (0,))):
            bad = {k: np.array(v, copy=True) for k, v in o.items()}
            bad[field][idx] ^= np.uint64(1)
            assert not verify_matmul(bad, lm, lk, ln, b"hv-mm", table, 1, 0, 9), field
