// Host-side Poseidon2 width-8 permutation for the Fiat-Shamir transcript, tuned for a
// CPU core: the same function as p2::permute (same constants, same round schedule), but
// every intermediate value is an unreduced 64-bit representative, sums and products are
// formed in 128 bits, and a single reduction follows each linear-layer output or product.
// Only the final state is brought to canonical residues. The transcript spends its time in
// this function (every label byte is a field element), so it sits on the critical path of
// every host-driven proving round.
#pragma once

#include "poseidon2.h"

namespace p2h {

typedef unsigned __int128 u128;
static constexpr u64 EPS = 0xFFFFFFFFull;  // 2^64 mod p

// Any 128-bit value -> a congruent value below 2^64 (2^64 = EPS, 2^96 = -1 mod p).
static inline u64 red(u128 v) {
    const u64 lo = (u64)v, hi = (u64)(v >> 64);
    const u64 hh = hi >> 32, hl = hi & EPS;
    u64 t0 = lo - hh;
    if (lo < hh) t0 -= EPS;
    const u64 t1 = hl * EPS;
    u64 t2 = t0 + t1;
    if (t2 < t1) t2 += EPS;
    return t2;
}

static inline u64 addc(u64 x, u64 c) {  // x + c for any x < 2^64 and c < p
    u64 t = x + c;
    if (t < c) t += EPS;
    return t;
}

static inline u64 mul(u64 a, u64 b) { return red((u128)a * b); }

static inline u64 sbox(u64 x) {
    const u64 x2 = mul(x, x);
    const u64 x3 = mul(x2, x);
    const u64 x4 = mul(x2, x2);
    return mul(x3, x4);
}

// circ(2 M4, M4) with M4 = circ(2, 3, 1, 1), all sums in 128 bits
static inline void external_linear(u64 *x) {
    u128 a[4], b[4];
    {
        const u128 s = (u128)x[0] + x[1] + x[2] + x[3];
        a[0] = s + x[0] + 2 * (u128)x[1];
        a[1] = s + x[1] + 2 * (u128)x[2];
        a[2] = s + x[2] + 2 * (u128)x[3];
        a[3] = s + x[3] + 2 * (u128)x[0];
    }
    {
        const u128 s = (u128)x[4] + x[5] + x[6] + x[7];
        b[0] = s + x[4] + 2 * (u128)x[5];
        b[1] = s + x[5] + 2 * (u128)x[6];
        b[2] = s + x[6] + 2 * (u128)x[7];
        b[3] = s + x[7] + 2 * (u128)x[4];
    }
    for (int i = 0; i < 4; i++) {
        const u128 t = a[i] + b[i];
        x[i] = red(t + a[i]);
        x[i + 4] = red(t + b[i]);
    }
}

// x_i d_i + sum: the sum (< 2^67) is added to each 128-bit product unreduced, so one
// reduction per output follows the S-box on the round's critical path
static inline void internal_linear(u64 *x, const u64 *diag) {
    u128 s = 0;
    for (int i = 0; i < 8; i++) s += x[i];
    for (int i = 0; i < 8; i++) x[i] = red((u128)x[i] * diag[i] + s);
}

// Accepts any 64-bit words; returns canonical residues.
static inline void permute(u64 *x) {
    const p2::Consts &k = p2::consts();
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < 8; i++) x[i] = sbox(addc(x[i], k.ext[r][i]));
        external_linear(x);
    }
    for (int r = 0; r < 22; r++) {
        x[0] = sbox(addc(x[0], k.in[r]));
        internal_linear(x, k.diag);
    }
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < 8; i++) x[i] = sbox(addc(x[i], k.ext[4 + r][i]));
        external_linear(x);
    }
    for (int i = 0; i < 8; i++) x[i] = gl_canon(x[i]);
}

}  // namespace p2h
