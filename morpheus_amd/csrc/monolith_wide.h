// Lane-parallel Monolith-w12 permutation for latency-bound Merkle levels.
//
// A group of 16 consecutive lanes (one DPP row) computes one permutation: lane j < 12 holds
// state word x_j; lanes 12..15 compute on copies and are ignored. Per round each lane
// computes one row of the circulant MDS from the 12 words it reads back from an LDS slot,
// one brick (x_j += x_{j-1}^2, the neighbour's word by DPP row_shr), and lanes 0..3 one bar.
// A permutation is ~645 wave instructions here; a launch that builds six small tree levels
// in one pass of its workgroups takes ~14 us, about 2 us per level. Gathering the words with
// 24 lane shuffles per round instead of the LDS slot was ~1.5x slower, and one lane doing the
// whole permutation ~7x. The top levels of a Merkle tree are a serial chain of
// permutations, so their latency is this figure per level.
// Same function as mono::permute (monolith.h); see there for the reference citations.
#pragma once

#include "monolith.h"

namespace mono_wide {

static constexpr int GROUP = 16;

// Per-lane constants: this lane's row of the circulant MDS, coef[c] = ROW0[(c - j) mod 12],
// and its round constants for rounds 1..5 (lanes 12..15 take row 0 and are ignored). One
// table entry per lane of a group, read once per kernel.
struct LaneConsts {
    u32 coef[12];
    u64 rc[5];
};
struct LaneTable {
    LaneConsts l[GROUP];
};
constexpr LaneTable make_lane_table() {
    LaneTable t{};
    constexpr mono::RcTable rc = {{MONO_RC_TABLE}};
    for (int j = 0; j < GROUP; j++) {
        const int r = j < 12 ? j : 0;
        for (int c = 0; c < 12; c++) t.l[j].coef[c] = mono::row0((c - r + 12) % 12);
        for (int k = 0; k < 5; k++) t.l[j].rc[k] = rc.v[k][r];
    }
    return t;
}
static __constant__ LaneTable kLaneTable = make_lane_table();

__device__ __forceinline__ void lane_consts(int j, LaneConsts &lc) { lc = kLaneTable.l[j]; }

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
__device__ __forceinline__ u64 concrete(u64 x, int j, u64 *st, const LaneConsts &lc) {
    st[j] = x;
    __builtin_amdgcn_wave_barrier();
    u64 v[12];
#pragma unroll
    for (int c = 0; c < 12; c++) v[c] = st[c];
    __builtin_amdgcn_wave_barrier();
    // the round constant's halves start the accumulators (as in mono::concrete); even and odd
    // columns go to separate accumulators, which halves the dependent multiply-add chain
    constexpr bool HAS_RC = ROUND >= 1 && ROUND <= 5;
    const u64 k = HAS_RC ? lc.rc[HAS_RC ? ROUND - 1 : 0] : 0;
    u64 a0 = (u32)k, b0 = k >> 32, a1 = 0, b1 = 0;
#pragma unroll
    for (int c = 0; c < 12; c += 2) {
        a0 += (u64)(u32)v[c] * lc.coef[c];
        b0 += (u64)(u32)(v[c] >> 32) * lc.coef[c];
        a1 += (u64)(u32)v[c + 1] * lc.coef[c + 1];
        b1 += (u64)(u32)(v[c + 1] >> 32) * lc.coef[c + 1];
    }
    const u64 y = mono::fold_halves(a0 + a1, b0 + b1);
    return j < 4 ? gl_canon(y) : y;
}

template <int ROUND>
__device__ __forceinline__ u64 round_fn(u64 x, int j, u64 *st, const LaneConsts &lc) {
    if (j < 4) x = mono::bar(x);
    const u64 prev = prev_lane(x);
    if (j >= 1) x = mono::sqr_add(x, prev);
    return concrete<ROUND>(x, j, st, lc);
}

// x: this lane's state word (any 64-bit representative); st: the group's 16-word LDS slot
// (16-byte aligned); lc from lane_consts. All 16 lanes of the group must be active.
// Returns the permuted word, canonical on lanes 0..3 (the digest lanes).
__device__ __forceinline__ u64 permute(u64 x, int j, u64 *st, const LaneConsts &lc) {
    x = concrete<0>(x, j, st, lc);
    x = round_fn<1>(x, j, st, lc);
    x = round_fn<2>(x, j, st, lc);
    x = round_fn<3>(x, j, st, lc);
    x = round_fn<4>(x, j, st, lc);
    x = round_fn<5>(x, j, st, lc);
    x = round_fn<6>(x, j, st, lc);
    return x;
}

}  // namespace mono_wide
