// Lane-parallel Monolith-w12 permutation for latency-bound Merkle levels.
//
// A group of 16 consecutive lanes (one DPP row) computes one permutation: lane j < 12 holds
// state word x_j; lanes 12..15 compute on copies and are ignored. Per round each lane
// computes one row of the circulant MDS from the 12 words it reads back from an LDS slot,
// one brick (x_j += x_{j-1}^2, the neighbour's word by DPP row_shr), and lanes 0..3 one bar.
// A permutation is ~960 wave instructions here (one wave issues one every ~4.7 cycles:
// ~4.6k cycles, ~1.9 us); gathering the words with 24 lane shuffles per round instead took
// ~6.8k cycles, and one lane doing the whole permutation ~33k. The top levels of a Merkle
// tree are a serial chain of permutations, so their latency is this figure per level.
// Same function as mono::permute (monolith.h); see there for the reference citations.
#pragma once

#include "monolith.h"

namespace mono_wide {

static constexpr int GROUP = 16;

__device__ __forceinline__ u64 rc_lane(int round, int j) {
    // round in 1..5, j in 0..11
    u64 r = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) r = (j == k) ? mono::rc(round, k) : r;
    return r;
}

// This lane's row of the circulant MDS: coef[c] = ROW0[(c - j) mod 12] (lanes 12..15 take
// row 0 and are ignored). Computed once per kernel.
__device__ __forceinline__ void lane_coeffs(int j, u32 *coef) {
    const int r = j < 12 ? j : 0;
#pragma unroll
    for (int c = 0; c < 12; c++) {
        int k = c - r;
        k = k < 0 ? k + 12 : k;
        u32 v = 0;
#pragma unroll
        for (int q = 0; q < 12; q++) v = (k == q) ? mono::row0(q) : v;
        coef[c] = v;
    }
}

// Lane j - 1's word within the 16-lane row (DPP row_shr:1); lane 0 of the row keeps its own.
__device__ __forceinline__ u64 prev_lane(u64 x) {
    u32 lo = (u32)x, hi = (u32)(x >> 32);
    lo = (u32)__builtin_amdgcn_update_dpp((int)lo, (int)lo, 0x111, 0xF, 0xF, false);
    hi = (u32)__builtin_amdgcn_update_dpp((int)hi, (int)hi, 0x111, 0xF, 0xF, false);
    return ((u64)hi << 32) | lo;
}

// y_j = sum_c coef[c] x_c + rc: each lane writes its word to the group's 16-word LDS slot
// and reads all twelve back (broadcast reads, 6 x 16 bytes); the slot is reused every round.
// LDS operations of one wave complete in program order, so the wave needs no barrier; the
// wave barriers only keep the compiler from moving the accesses across each other.
// Lanes j < 4 return the canonical residue (the next bar inputs, or the digest).
template <int ROUND>
__device__ __forceinline__ u64 concrete(u64 x, int j, u64 *st, const u32 *coef) {
    st[j] = x;
    __builtin_amdgcn_wave_barrier();
    u64 v[12];
#pragma unroll
    for (int c = 0; c < 12; c++) v[c] = st[c];
    __builtin_amdgcn_wave_barrier();
    u64 a = 0, b = 0;
#pragma unroll
    for (int c = 0; c < 12; c++) {
        a += (u64)(u32)v[c] * coef[c];
        b += (u64)(u32)(v[c] >> 32) * coef[c];
    }
    const u64 y = mono::combine(a, b, (ROUND >= 1 && ROUND <= 5) ? rc_lane(ROUND, j < 12 ? j : 0) : 0);
    return j < 4 ? gl_canon(y) : y;
}

template <int ROUND>
__device__ __forceinline__ u64 round_fn(u64 x, int j, u64 *st, const u32 *coef) {
    if (j < 4) x = mono::bar(x);
    const u64 prev = prev_lane(x);
    if (j >= 1) x = mono::add_nc(x, mono::sqr_nc(prev));
    return concrete<ROUND>(x, j, st, coef);
}

// x: this lane's state word (any 64-bit representative); st: the group's 16-word LDS slot
// (16-byte aligned); coef from lane_coeffs. All 16 lanes of the group must be active.
// Returns the permuted word, canonical on lanes 0..3 (the digest lanes).
__device__ __forceinline__ u64 permute(u64 x, int j, u64 *st, const u32 *coef) {
    x = concrete<0>(x, j, st, coef);
    x = round_fn<1>(x, j, st, coef);
    x = round_fn<2>(x, j, st, coef);
    x = round_fn<3>(x, j, st, coef);
    x = round_fn<4>(x, j, st, coef);
    x = round_fn<5>(x, j, st, coef);
    x = round_fn<6>(x, j, st, coef);
    return x;
}

}  // namespace mono_wide
