// Lane-parallel Monolith-w12 permutation for latency-bound Merkle levels.
//
// A group of 16 consecutive lanes computes one permutation: lane j < 12 holds state word
// x_j; lanes 12..15 compute on copies and are ignored. Per round each lane computes one
// row of the circulant MDS from the 12 words it gathers with lane shuffles (ds_bpermute
// within the 16-lane group), one brick (x_j += x_{j-1}^2, one shuffle), and lanes 0..3 one
// bar. This spreads a permutation's ~9k VALU instructions over 12 lanes, so the
// permutation finishes in roughly a tenth of the single-lane time: the top levels of a
// Merkle tree, which are a serial chain of permutations, become short.
// Same function as mono::permute (monolith.h); see there for the reference citations.
#pragma once

#include "monolith.h"

namespace mono_wide {

static constexpr int GROUP = 16;

__device__ __forceinline__ u64 shfl64(u64 v, int src) {
    u32 lo = (u32)v, hi = (u32)(v >> 32);
    lo = (u32)__shfl((int)lo, src, GROUP);
    hi = (u32)__shfl((int)hi, src, GROUP);
    return ((u64)hi << 32) | lo;
}

__device__ __forceinline__ u64 rc_lane(int round, int j) {
    // round in 1..5, j in 0..11
    u64 r = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) r = (j == k) ? mono::rc(round, k) : r;
    return r;
}

// y_j = sum_k ROW0[k] x_{(j + k) mod 12} + rc  (row j of the circulant M[j][c] = ROW0[(c - j) mod 12]),
// on 32-bit halves with 64-bit multiply-accumulate (as mono::concrete). Lanes j < 4 return
// the canonical residue (the next bar inputs, or the digest after the last layer).
template <int ROUND>
__device__ __forceinline__ u64 concrete(u64 x, int j) {
    u64 a = 0, b = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
        int src = j + k;
        src = src >= 12 ? src - 12 : src;
        src = src >= 12 ? src - 12 : src;
        const u64 v = shfl64(x, src);
        const u64 m = mono::row0(k);
        a += (u64)(u32)v * m;
        b += (u64)(u32)(v >> 32) * m;
    }
    u64 s = a + (b << 32);
    u64 h = (b >> 32) + (s < a);
    if (ROUND >= 1 && ROUND <= 5) {
        const u64 k = rc_lane(ROUND, j < 12 ? j : 0);
        u64 s2 = s + k;
        h += (s2 < k);
        s = s2;
    }
    const u64 y = mono::fold_hi_nc(s, h);
    return j < 4 ? gl_canon(y) : y;
}

template <int ROUND>
__device__ __forceinline__ u64 round_fn(u64 x, int j) {
    if (j < 4) x = mono::bar(x);
    const u64 prev = shfl64(x, j >= 1 ? j - 1 : 0);
    if (j >= 1) x = mono::add_nc(x, mono::sqr_nc(prev));
    return concrete<ROUND>(x, j);
}

// x: this lane's state word (any 64-bit representative); returns the permuted word,
// canonical on lanes 0..3 (the digest lanes).
__device__ __forceinline__ u64 permute(u64 x, int j) {
    x = concrete<0>(x, j);
    x = round_fn<1>(x, j);
    x = round_fn<2>(x, j);
    x = round_fn<3>(x, j);
    x = round_fn<4>(x, j);
    x = round_fn<5>(x, j);
    x = round_fn<6>(x, j);
    return x;
}

}  // namespace mono_wide
