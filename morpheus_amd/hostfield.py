"""Host-side Goldilocks and Ext2 (X^2 - 7) arithmetic for checking results: vectorised field
operations on uint64 arrays, Python-integer Ext2 operations, multilinear evaluation and
Lagrange interpolation (the pieces of zk-torch-3/src/sumcheck/verifier.rs:116-140 and
zk-torch-3/src/poly/mod.rs:106-163 a verifier needs)."""
import numpy as np

P = 0xFFFFFFFF00000001
W = 7
_M32 = np.uint64(0xFFFFFFFF)
_EPS = np.uint64(0xFFFFFFFF)  # 2^64 mod p
_P64 = np.uint64(P)
_S32 = np.uint64(32)


def gl_mul_np(a, b):
    """Elementwise a * b mod p of uint64 arrays (any words in [0, 2^64)): the 128-bit product
    from four 32 x 32 partial products, then 2^64 = 2^32 - 1 and 2^96 = -1 mod p."""
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    with np.errstate(over="ignore"):
        a0, a1, b0, b1 = a & _M32, a >> _S32, b & _M32, b >> _S32
        ll, m1, m2, hh = a0 * b0, a0 * b1, a1 * b0, a1 * b1
        mid = m1 + m2
        cmid = (mid < m1).astype(np.uint64)             # a carry out of mid is worth 2^96
        lo = ll + (mid << _S32)
        clo = (lo < ll).astype(np.uint64)
        hi = hh + (mid >> _S32) + clo + (cmid << _S32)
        hi0, hi1 = hi & _M32, hi >> _S32
        t = lo - hi1
        t = np.where(lo < hi1, t - _EPS, t)             # borrow: + p = - (2^32 - 1) mod 2^64
        u = (hi0 << _S32) - hi0                         # hi0 (2^32 - 1)
        r = t + u
        r = np.where(r < u, r + _EPS, r)                # carry: + 2^64 = + (2^32 - 1)
        return np.where(r >= _P64, r - _P64, r)


def gl_add_np(a, b):
    """a + b mod p for canonical uint64 arrays."""
    with np.errstate(over="ignore"):
        s = np.asarray(a, dtype=np.uint64) + np.asarray(b, dtype=np.uint64)
        s = np.where(s < a, s + _EPS, s)
        return np.where(s >= _P64, s - _P64, s)


def gl_sub_np(a, b):
    """a - b mod p for canonical uint64 arrays."""
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    with np.errstate(over="ignore"):
        d = a - b
        return np.where(a < b, d - _EPS, d)


def gl_sum_np(a):
    """(sum of a uint64 array) mod p, exact."""
    a = np.asarray(a, dtype=np.uint64).ravel()
    total = 0
    for s in range(0, a.size, 1 << 24):
        c = a[s:s + (1 << 24)]
        total += int((c & _M32).sum(dtype=np.uint64)) + (int((c >> _S32).sum(dtype=np.uint64)) << 32)
    return total % P


def gl_dot_np(a, b):
    """sum_i a_i b_i mod p."""
    return gl_sum_np(gl_mul_np(a, b))


def canon_np(a):
    a = np.asarray(a, dtype=np.uint64)
    return np.where(a >= _P64, a - _P64, a)


# ---- Ext2 on Python integers: (c0, c1) = c0 + c1 X, X^2 = 7 ----
def e2add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def e2sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def e2mul(a, b):
    return ((a[0] * b[0] + W * a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def e2inv(a):
    """a^-1 = conj(a) / norm(a), conj(c0 + c1 X) = c0 - c1 X, norm = c0^2 - 7 c1^2."""
    n = (a[0] * a[0] - W * a[1] * a[1]) % P
    ni = pow(n, P - 2, P)
    return (a[0] * ni % P, (-a[1]) * ni % P)


def interp_eval(evals, x):
    """The polynomial through (0, evals[0]), ..., (d, evals[d]) at x (Ext2 tuples)."""
    d = len(evals)
    res = (0, 0)
    for i in range(d):
        num, den = (1, 0), 1
        for j in range(d):
            if j != i:
                num = e2mul(num, e2sub(x, (j, 0)))
                den = den * (i - j) % P
        res = e2add(res, e2mul(evals[i], e2mul(num, (pow(den, P - 2, P), 0))))
    return res


def mle_eval(evals, point2):
    """Multilinear extension of base-field evaluations (2^nv uint64) at an Ext2 point
    (nv pairs, variable i = bit i of the index, LSB first), as an Ext2 tuple."""
    pt = np.asarray(point2, dtype=np.uint64).reshape(-1, 2)
    v = canon_np(evals)
    if v.size != 1 << pt.shape[0]:
        raise ValueError("evaluations and point disagree in size")
    if v.size == 1:
        return (int(v[0]), 0)
    # first variable: base -> Ext2
    r0, r1 = np.uint64(int(pt[0, 0]) % P), np.uint64(int(pt[0, 1]) % P)
    a, b = v[0::2], v[1::2]
    d = gl_sub_np(b, a)
    c0, c1 = gl_add_np(a, gl_mul_np(d, r0)), gl_mul_np(d, r1)
    seven = np.uint64(W)
    for i in range(1, pt.shape[0]):
        r0, r1 = np.uint64(int(pt[i, 0]) % P), np.uint64(int(pt[i, 1]) % P)
        a0, a1, b0, b1 = c0[0::2], c1[0::2], c0[1::2], c1[1::2]
        d0, d1 = gl_sub_np(b0, a0), gl_sub_np(b1, a1)
        # r d = (r0 d0 + 7 r1 d1, r0 d1 + r1 d0)
        t0 = gl_add_np(gl_mul_np(d0, r0), gl_mul_np(gl_mul_np(d1, r1), seven))
        t1 = gl_add_np(gl_mul_np(d1, r0), gl_mul_np(d0, r1))
        c0, c1 = gl_add_np(a0, t0), gl_add_np(a1, t1)
    return (int(c0[0]), int(c1[0]))
