// Goldilocks field p = 2^64 - 2^32 + 1 and its quadratic extension F_p[X]/(X^2 - 7),
// usable from host and device code.
//
// Every function returns a canonical residue in [0, p) when its inputs are canonical;
// `gl_canon` maps any 64-bit word to its residue, and the kernels call it on every value
// they read from caller memory, so non-canonical inputs are accepted everywhere.
//
// Reference semantics: cuda/goldilocks.cuh:181-311 (add/sub/mul, reduce128 via
// 2^64 = 2^32 - 1 mod p), cuda/extension.cuh:112-183 (Ext2 with W = 7). The reference keeps
// GPU values non-canonical; canonical residues are what its CPU prover compares
// (zk-torch-3/src/commit/cpu_basefold.rs:96-99), so this library stores residues only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define MAM_HD __host__ __device__ __forceinline__

typedef uint64_t u64;
typedef uint32_t u32;

static constexpr u64 GL_P = 0xFFFFFFFF00000001ULL;
static constexpr u64 GL_EPS = 0x00000000FFFFFFFFULL;  // 2^64 mod p = 2^32 - 1
static constexpr u64 GL_INV2 = 0x7FFFFFFF80000001ULL;  // (p + 1) / 2
static constexpr u64 EXT2_W = 7;

MAM_HD u64 gl_canon(u64 a) { return a >= GL_P ? a - GL_P : a; }

MAM_HD u64 gl_add(u64 a, u64 b) {
    // a + b - p computed as a - (p - b); valid for canonical inputs.
    u64 nb = GL_P - b;
    u64 r = a - nb;
    return a < nb ? r + GL_P : r;
}

MAM_HD u64 gl_sub(u64 a, u64 b) {
    u64 r = a - b;
    return a < b ? r + GL_P : r;
}

MAM_HD u64 gl_neg(u64 a) { return a ? GL_P - a : 0; }
MAM_HD u64 gl_double(u64 a) { return gl_add(a, a); }

// The full 128-bit product. On the device: four 32 x 32 -> 64 multiply-adds
// (v_mad_u64_u32) chained through the middle words, which yield both halves at once (a * b
// and __umul64hi separately take two more 32-bit multiplies for the low half).
MAM_HD void mul_wide(u64 a, u64 b, u64 &lo, u64 &hi) {
#ifdef __HIP_DEVICE_COMPILE__
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    const u64 p00 = (u64)a0 * b0;
    const u64 p01 = (u64)a0 * b1 + (p00 >> 32);  // < 2^64
    const u64 p10 = (u64)a1 * b0 + (u32)p01;     // < 2^64
    hi = (u64)a1 * b1 + (p01 >> 32) + (p10 >> 32);
    lo = (p10 << 32) | (u32)p00;
#else
    unsigned __int128 m = (unsigned __int128)a * b;
    lo = (u64)m;
    hi = (u64)(m >> 64);
#endif
}

// Reduce lo + 2^64 * hi (any 128-bit value) to its canonical residue.
// With hi = hh * 2^32 + hl: 2^64 = 2^32 - 1 and 2^96 = -1 (mod p), so
// x = lo - hh + hl * (2^32 - 1) (mod p).
MAM_HD u64 gl_reduce128(u64 lo, u64 hi) {
    u64 hh = hi >> 32;
    u64 hl = hi & GL_EPS;
    u64 t0 = lo - hh;
    if (lo < hh) t0 -= GL_EPS;  // borrowed 2^64: add back 2^64 - p... i.e. subtract 2^32 - 1
    u64 t1 = hl * GL_EPS;       // < 2^64
    u64 t2 = t0 + t1;
    if (t2 < t1) t2 += GL_EPS;  // carried 2^64: add 2^32 - 1 (cannot carry again)
    return gl_canon(t2);
}

MAM_HD u64 gl_mul(u64 a, u64 b) {
    u64 lo, hi;
    mul_wide(a, b, lo, hi);
    return gl_reduce128(lo, hi);
}

MAM_HD u64 gl_sqr(u64 a) { return gl_mul(a, a); }

// 7 * a for canonical a: 7a < 2^67.
MAM_HD u64 gl_mul7(u64 a) {
    u64 lo = a << 3;
    u64 hi = a >> 61;
    // 8a - a
    u64 lo2 = lo - a;
    hi -= (lo < a);
    return gl_reduce128(lo2, hi);
}

MAM_HD u64 gl_pow(u64 a, u64 e) {
    u64 r = 1;
    while (e) {
        if (e & 1) r = gl_mul(r, a);
        a = gl_mul(a, a);
        e >>= 1;
    }
    return r;
}

// Fermat inverse a^(p-2); maps 0 to 0 (as cuda/goldilocks.cuh:363-367 does).
MAM_HD u64 gl_inv(u64 a) { return gl_pow(a, GL_P - 2); }

// ---------------------------------------------------------------------------
// Ext2 = F_p[X]/(X^2 - 7), stored as two u64 words [c0, c1] (interleaved in memory).
// ---------------------------------------------------------------------------
struct ext2 {
    u64 c0, c1;
};

MAM_HD ext2 e2(u64 a, u64 b) { return ext2{a, b}; }
MAM_HD ext2 e2_zero() { return ext2{0, 0}; }
MAM_HD ext2 e2_one() { return ext2{1, 0}; }
MAM_HD ext2 e2_base(u64 a) { return ext2{a, 0}; }
MAM_HD ext2 e2_add(ext2 a, ext2 b) { return ext2{gl_add(a.c0, b.c0), gl_add(a.c1, b.c1)}; }
MAM_HD ext2 e2_sub(ext2 a, ext2 b) { return ext2{gl_sub(a.c0, b.c0), gl_sub(a.c1, b.c1)}; }
MAM_HD ext2 e2_neg(ext2 a) { return ext2{gl_neg(a.c0), gl_neg(a.c1)}; }
MAM_HD bool e2_eq(ext2 a, ext2 b) { return a.c0 == b.c0 && a.c1 == b.c1; }

// (a0 + a1 X)(b0 + b1 X) = (a0 b0 + 7 a1 b1) + (a0 b1 + a1 b0) X
MAM_HD ext2 e2_mul(ext2 a, ext2 b) {
    u64 m00 = gl_mul(a.c0, b.c0);
    u64 m11 = gl_mul(a.c1, b.c1);
    u64 m01 = gl_mul(a.c0, b.c1);
    u64 m10 = gl_mul(a.c1, b.c0);
    return ext2{gl_add(m00, gl_mul7(m11)), gl_add(m01, m10)};
}

// Product with a value k fixed for a whole loop, whose 7 k.c1 (k7 = gl_mul7(k.c1)) was
// formed once: (k0 + k1 X)(x0 + x1 X) = (k0 x0 + k7 x1) + (k0 x1 + k1 x0) X.
MAM_HD ext2 e2_mul_fixed(ext2 k, u64 k7, ext2 x) {
    return ext2{gl_add(gl_mul(k.c0, x.c0), gl_mul(k7, x.c1)), gl_add(gl_mul(k.c0, x.c1), gl_mul(k.c1, x.c0))};
}

MAM_HD ext2 e2_sqr(ext2 a) {
    u64 m00 = gl_mul(a.c0, a.c0);
    u64 m11 = gl_mul(a.c1, a.c1);
    u64 m01 = gl_mul(a.c0, a.c1);
    return ext2{gl_add(m00, gl_mul7(m11)), gl_double(m01)};
}

// base * ext2
MAM_HD ext2 e2_scale(u64 s, ext2 a) { return ext2{gl_mul(s, a.c0), gl_mul(s, a.c1)}; }

// (a0 - a1 X) / (a0^2 - 7 a1^2)   (cuda/extension.cuh:167-183)
MAM_HD ext2 e2_inv(ext2 a) {
    u64 norm = gl_sub(gl_sqr(a.c0), gl_mul7(gl_sqr(a.c1)));
    u64 ni = gl_inv(norm);
    return ext2{gl_mul(a.c0, ni), gl_neg(gl_mul(a.c1, ni))};
}

MAM_HD ext2 e2_pow(ext2 a, u64 e) {
    ext2 r = e2_one();
    while (e) {
        if (e & 1) r = e2_mul(r, a);
        a = e2_sqr(a);
        e >>= 1;
    }
    return r;
}

// Frobenius x -> x^p: conjugation (a0, -a1), since 7^((p-1)/2) = -1.
MAM_HD ext2 e2_frobenius(ext2 a) { return ext2{a.c0, gl_neg(a.c1)}; }

// Load an Ext2 from interleaved memory, canonicalizing each word.
MAM_HD ext2 e2_load(const u64 *p, size_t i) { return ext2{gl_canon(p[2 * i]), gl_canon(p[2 * i + 1])}; }
MAM_HD void e2_store(u64 *p, size_t i, ext2 v) {
    p[2 * i] = v.c0;
    p[2 * i + 1] = v.c1;
}
