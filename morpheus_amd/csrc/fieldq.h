// Almost-Goldilocks field q = 2^64 - 2^32 - 31 and its quadratic extension
// F_q[X]/(X^2 - 3), host and device.
//
// Reference semantics: cuda_almost_goldilocks/almost_goldilocks.cuh:23-300 (constants,
// 2^64 = c = 2^32 + 31 mod q, multi-pass Solinas reduction) and almost_extension.cuh:31-213
// (W = 3, Karatsuba product, inverse by the norm, Frobenius = conjugation); host twins in
// almost-goldilocks-cuda-rs/src/field.rs:11-133 and extension.rs:55-86. As with the
// Goldilocks code, this library stores canonical residues; every function here accepts
// any 64-bit representative and returns the canonical one.
#pragma once

#include "field.h"

namespace fq {

static constexpr u64 Q = 0xFFFFFFFEFFFFFFE1ULL;
static constexpr u64 C = 0x10000001FULL;           // 2^64 mod q
static constexpr u64 HALF = 0x7FFFFFFF7FFFFFF1ULL;  // (q + 1) / 2
static constexpr u64 W = 3;                         // Ext2 non-residue

MAM_HD u64 canon(u64 a) { return a >= Q ? a - Q : a; }

// canonical inputs
MAM_HD u64 add(u64 a, u64 b) {
    const u64 nb = Q - b;
    const u64 r = a - nb;
    return a < nb ? r + Q : r;
}
MAM_HD u64 sub(u64 a, u64 b) {
    const u64 r = a - b;
    return a < b ? r + Q : r;
}
MAM_HD u64 neg(u64 a) { return a ? Q - a : 0; }

// lo + 2^64 hi (any 128-bit value) -> canonical residue. hi * C < 2^97; two folds bring the
// value below 2^64 + 2^38, a third (tiny) fold below 2^64, one subtraction below q.
MAM_HD u64 reduce128(u64 lo, u64 hi) {
    u64 l1, h1;
    mul_wide(hi, C, l1, h1);  // hi * C
    u64 s = lo + l1;
    h1 += (s < l1);           // h1 <= 2^33 + 1
    u64 l2, h2;
    mul_wide(h1, C, l2, h2);  // < 2^67
    u64 t = s + l2;
    h2 += (t < l2);           // h2 tiny
    const u64 corr = h2 * C;
    u64 v = t + corr;
    if (v < corr) v += C;     // 2^64 wrapped: add c (cannot wrap again: v < corr <= 2^38)
    return canon(v);
}

MAM_HD u64 mul(u64 a, u64 b) {
    u64 lo, hi;
    mul_wide(a, b, lo, hi);
    return reduce128(lo, hi);
}
MAM_HD u64 sqr(u64 a) { return mul(a, a); }
MAM_HD u64 dbl(u64 a) { return add(a, a); }
MAM_HD u64 mul3(u64 a) { return add(dbl(a), a); }
MAM_HD u64 halve(u64 a) { return (a & 1) ? (a >> 1) + HALF : (a >> 1); }

MAM_HD u64 pow(u64 a, u64 e) {
    u64 r = 1;
    while (e) {
        if (e & 1) r = mul(r, a);
        a = mul(a, a);
        e >>= 1;
    }
    return r;
}
// Fermat inverse; maps 0 to 0 (the reference's agl_inverse computes 0^(q-2) = 0)
MAM_HD u64 inv(u64 a) { return pow(a, Q - 2); }

// Sum of two canonical values and a 128-bit accumulator helper: acc (lo, hi) += v.
MAM_HD void acc_add(u64 &lo, u64 &hi, u64 v) {
    lo += v;
    hi += (lo < v);
}

}  // namespace fq

// Ext2 over q: c0 + c1 X with X^2 = 3. Same storage as ext2 (two canonical words).
struct qe2 {
    u64 c0, c1;
};

MAM_HD qe2 q2(u64 a, u64 b) { return qe2{a, b}; }
MAM_HD qe2 q2_zero() { return qe2{0, 0}; }
MAM_HD qe2 q2_one() { return qe2{1, 0}; }
MAM_HD qe2 q2_base(u64 a) { return qe2{a, 0}; }
MAM_HD qe2 q2_canon(qe2 a) { return qe2{fq::canon(a.c0), fq::canon(a.c1)}; }
MAM_HD qe2 q2_add(qe2 a, qe2 b) { return qe2{fq::add(a.c0, b.c0), fq::add(a.c1, b.c1)}; }
MAM_HD qe2 q2_sub(qe2 a, qe2 b) { return qe2{fq::sub(a.c0, b.c0), fq::sub(a.c1, b.c1)}; }
MAM_HD qe2 q2_neg(qe2 a) { return qe2{fq::neg(a.c0), fq::neg(a.c1)}; }
MAM_HD qe2 q2_scale(u64 s, qe2 a) { return qe2{fq::mul(s, a.c0), fq::mul(s, a.c1)}; }
// (a0 + a1 X)(b0 + b1 X) = a0 b0 + 3 a1 b1 + ((a0 + a1)(b0 + b1) - a0 b0 - a1 b1) X
MAM_HD qe2 q2_mul(qe2 a, qe2 b) {
    const u64 t0 = fq::mul(a.c0, b.c0);
    const u64 t1 = fq::mul(a.c1, b.c1);
    const u64 t2 = fq::mul(fq::add(a.c0, a.c1), fq::add(b.c0, b.c1));
    return qe2{fq::add(t0, fq::mul3(t1)), fq::sub(fq::sub(t2, t0), t1)};
}
MAM_HD qe2 q2_sqr(qe2 a) { return q2_mul(a, a); }
MAM_HD qe2 q2_conj(qe2 a) { return qe2{a.c0, fq::neg(a.c1)}; }  // = Frobenius (3 is a non-residue)
MAM_HD u64 q2_norm(qe2 a) { return fq::sub(fq::sqr(a.c0), fq::mul3(fq::sqr(a.c1))); }
MAM_HD qe2 q2_inv(qe2 a) {
    const u64 ni = fq::inv(q2_norm(a));
    return qe2{fq::mul(a.c0, ni), fq::mul(fq::neg(a.c1), ni)};
}
MAM_HD qe2 q2_pow(qe2 a, u64 e) {
    qe2 r = q2_one();
    while (e) {
        if (e & 1) r = q2_mul(r, a);
        a = q2_mul(a, a);
        e >>= 1;
    }
    return r;
}
MAM_HD qe2 q2_load(const u64 *p, size_t i) { return qe2{fq::canon(p[2 * i]), fq::canon(p[2 * i + 1])}; }
MAM_HD void q2_store(u64 *p, size_t i, qe2 v) {
    p[2 * i] = v.c0;
    p[2 * i + 1] = v.c1;
}
