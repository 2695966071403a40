"""Independent big-integer formulations used to pin the oracle (plain Python ints).

Nothing here is a copy of the oracle's C code: each function computes its quantity from
the mathematical definition (MLE sums, explicit matrices, pow() inverses).
"""
import json
import os

import numpy as np

P = 0xFFFFFFFF00000001
W = 7
GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_vectors.json")))


def e2mul(a, b):
    # (a0 + a1 X)(b0 + b1 X) mod (X^2 - 7)
    return ((a[0] * b[0] + W * a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def e2add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def e2sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def e2inv(a):
    # a^(p^2 - 2) in the field of p^2 elements
    r, x, e = (1, 0), (a[0] % P, a[1] % P), P * P - 2
    while e:
        if e & 1:
            r = e2mul(r, x)
        x = e2mul(x, x)
        e >>= 1
    return r


def mle_eval(vals, point):
    """Multilinear extension of vals (list of Ext2 tuples, 2^nv) at point (nv Ext2), LSB first:
    sum_x vals[x] * prod_i (x_i ? r_i : 1 - r_i)."""
    nv = len(point)
    acc = (0, 0)
    for x in range(1 << nv):
        w = (1, 0)
        for i in range(nv):
            w = e2mul(w, point[i] if (x >> i) & 1 else e2sub((1, 0), point[i]))
        acc = e2add(acc, e2mul(vals[x], w))
    return acc


def eq_direct(point, x):
    w = (1, 0)
    for i, r in enumerate(point):
        w = e2mul(w, r if (x >> i) & 1 else e2sub((1, 0), r))
    return w


def pairs(a):
    a = np.asarray(a, dtype=np.uint64).reshape(-1, 2)
    return [(int(x), int(y)) for x, y in a]


def rand_field(rng, n):
    return rng.integers(0, P, size=n, dtype=np.uint64)


def poseidon2_matrix_form(state):
    """Poseidon2-w8 from its definition with explicit matrices: M_E = [[2M4, M4],[M4, 2M4]],
    M4 = circ(2,3,1,1) rows [2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]; M_I = 1 + diag(D)."""
    c = GOLDEN["poseidon2_w8_constants"]
    m4 = [[2, 3, 1, 1], [1, 2, 3, 1], [1, 1, 2, 3], [3, 1, 1, 2]]
    me = [[(2 if (i // 4) == (j // 4) else 1) * m4[i % 4][j % 4] for j in range(8)] for i in range(8)]
    mi = [[(1 + (c["diag"][i] if i == j else 0)) for j in range(8)] for i in range(8)]

    def mat(m, s):
        return [sum(m[i][j] * s[j] for j in range(8)) % P for i in range(8)]

    s = [int(v) % P for v in state]
    for r in range(4):
        s = mat(me, [pow((s[i] + c["rc_ext"][r][i]) % P, 7, P) for i in range(8)])
    for r in range(22):
        s[0] = pow((s[0] + c["rc_int"][r]) % P, 7, P)
        s = mat(mi, s)
    for r in range(4):
        s = mat(me, [pow((s[i] + c["rc_ext"][4 + r][i]) % P, 7, P) for i in range(8)])
    return s


# vectorised Goldilocks arithmetic on uint64 arrays (the library's host-side helpers; pinned
# against Python integers in test_oracle.py)
from morpheus_amd.hostfield import gl_dot_np, gl_mul_np, gl_sum_np  # noqa: E402,F401
