/*
 * oracle/gl.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Scalar C restatement of Goldilocks (p = 2^64 - 2^32 + 1) and its quadratic
 * extension F_p[X]/(X^2 - 7).  Every value produced here is canonical
 * (in [0, p)); the reference's own CPU/GPU comparisons are made on residues
 * (zk-torch-3/src/commit/cpu_basefold.rs:96-99, goldilocks-cuda-rs/src/basefold.rs:26-30).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this code, and only as the checker.
 *
 * Follows:
 *   field ops      goldilocks-cuda-rs/src/field.rs:55-106 (Add/Sub/Mul),
 *                  zk-torch-3/src/util/arith.rs:94-131, 206-237 (reduce128)
 *   Ext2 (W = 7)   cuda/extension.cuh:22-24, 112-125 (mul), 167-183 (inverse)
 */
#ifndef ORACLE_GL_H
#define ORACLE_GL_H

#include <stdint.h>

typedef uint64_t u64;
typedef unsigned __int128 u128;

#define GL_P 0xFFFFFFFF00000001ULL
#define GL_EPS 0x00000000FFFFFFFFULL /* 2^64 mod p */

/* Reduce any 128-bit value to its canonical residue.  2^64 = 2^32 - 1 (mod p),
 * applied twice, then a final conditional subtraction (arith.rs:206-237). */
static inline u64 gl_reduce128(u128 x) {
    u64 lo = (u64)x, hi = (u64)(x >> 64);
    while (hi) { /* hi shrinks 2^64 -> 2^33 -> <=2 -> <=1 -> 0 */
        u128 s = (u128)lo + (u128)hi * GL_EPS;
        lo = (u64)s;
        hi = (u64)(s >> 64);
    }
    if (lo >= GL_P) lo -= GL_P;
    return lo;
}

static inline u64 gl_canon(u64 a) { return a >= GL_P ? a - GL_P : a; }
static inline u64 gl_add(u64 a, u64 b) { return gl_reduce128((u128)a + b); }
static inline u64 gl_sub(u64 a, u64 b) {
    a = gl_canon(a);
    b = gl_canon(b);
    return a >= b ? a - b : GL_P - (b - a);
}
static inline u64 gl_neg(u64 a) { a = gl_canon(a); return a ? GL_P - a : 0; }
static inline u64 gl_mul(u64 a, u64 b) { return gl_reduce128((u128)a * b); }

static inline u64 gl_pow(u64 a, u64 e) {
    u64 r = 1;
    a = gl_canon(a);
    while (e) {
        if (e & 1) r = gl_mul(r, a);
        a = gl_mul(a, a);
        e >>= 1;
    }
    return r;
}
/* Fermat inverse a^(p-2); 0 maps to 0 (arith.rs:126-132). */
static inline u64 gl_inv(u64 a) { return gl_pow(a, GL_P - 2); }

/* ---- Ext2: c0 + c1*X, X^2 = 7 ---- */
typedef struct { u64 c0, c1; } ext2;

static inline ext2 e2(u64 c0, u64 c1) { ext2 r = {gl_canon(c0), gl_canon(c1)}; return r; }
static inline ext2 e2_zero(void) { ext2 r = {0, 0}; return r; }
static inline ext2 e2_one(void) { ext2 r = {1, 0}; return r; }
static inline ext2 e2_base(u64 a) { ext2 r = {gl_canon(a), 0}; return r; }
static inline ext2 e2_add(ext2 a, ext2 b) { ext2 r = {gl_add(a.c0, b.c0), gl_add(a.c1, b.c1)}; return r; }
static inline ext2 e2_sub(ext2 a, ext2 b) { ext2 r = {gl_sub(a.c0, b.c0), gl_sub(a.c1, b.c1)}; return r; }
static inline ext2 e2_mul(ext2 a, ext2 b) {
    /* (a0 + a1 X)(b0 + b1 X) = a0 b0 + 7 a1 b1 + (a0 b1 + a1 b0) X */
    u128 m0 = (u128)a.c0 * b.c0;
    u64 m1 = gl_mul(a.c1, b.c1);
    ext2 r;
    r.c0 = gl_add(gl_reduce128(m0), gl_mul(7, m1));
    r.c1 = gl_add(gl_mul(a.c0, b.c1), gl_mul(a.c1, b.c0));
    return r;
}
static inline ext2 e2_scale(u64 s, ext2 a) { ext2 r = {gl_mul(s, a.c0), gl_mul(s, a.c1)}; return r; }
static inline int e2_eq(ext2 a, ext2 b) { return gl_canon(a.c0) == gl_canon(b.c0) && gl_canon(a.c1) == gl_canon(b.c1); }
/* (a0 - a1 X) / (a0^2 - 7 a1^2)  (extension.cuh:167-183) */
static inline ext2 e2_inv(ext2 a) {
    u64 norm = gl_sub(gl_mul(a.c0, a.c0), gl_mul(7, gl_mul(a.c1, a.c1)));
    u64 ni = gl_inv(norm);
    ext2 r = {gl_mul(a.c0, ni), gl_neg(gl_mul(a.c1, ni))};
    return r;
}

#endif
