// Monolith-w12 spread over A lanes per permutation (A = 2, 4, 6 or 12), for Merkle levels
// too narrow to keep the chip busy with one permutation per lane.
//
// Layout: a group of G consecutive lanes (G = 2, 4, 8, 16: the power of two >= A) holds one
// state; lane q < A holds the words w = q + A k, k < W = 12 / A, in its registers x[k];
// lanes q >= A compute on copies and their results are ignored. Per round, lanes exchange
// words through the group's slot in LDS, which holds the 12 words twice (positions w and
// w + 12), so that every run of 12 words starting below 12 is contiguous:
//   bars     on words 0..3 (slot 0 of lanes 0..3; for A = 2 slots 0 and 1 of lanes 0, 1)
//   bricks   y_w = x_w + x_{w-1}^2 for w >= 1: each lane reads the post-bar word w - 1 from
//            position w + 11 of the slot
//   concrete y_r = sum_d ROW0[d] x_{(r+d) mod 12} (+ RC[round][r]): each lane reads the run of
//            24 - A words starting at position q and takes, for its word r = q + A k, the 12
//            words at offsets A k .. A k + 11. The multipliers ROW0[d] are the same on every
//            lane (inline constants) because the run starts at the lane's own first word.
// The permutation's latency falls about as 12 / A (fewer words per lane) while the work per
// permutation stays within a few percent of the one-lane form for A <= 4, so narrow Merkle
// levels run on A lanes at about the same total cost and a fraction of the latency.
// Same function as mono::permute (monolith.h), with the reference citations there.
#pragma once

#include "monolith.h"

namespace mono_lp {

template <int A>
struct Form {
    static_assert(A == 2 || A == 4 || A == 6 || A == 12, "lanes per permutation");
    static constexpr int W = 12 / A;                              // words per lane
    static constexpr int G = A == 6 ? 8 : A == 12 ? 16 : A;       // lanes per group
    // u64 per group slot: the doubled state (24), the idle lanes' spare positions and their
    // copies 12 above, rounded up to an odd count so that the slots of the groups of a wave
    // start on different LDS banks
    static constexpr int SLOT = A == 12 ? 41 : A == 6 ? 39 : 25;
};

// Round constants of rounds 1..5, one table per workgroup in LDS (filled by load_rc).
struct RcLds {
    u64 v[5][12];
};

__device__ __forceinline__ void load_rc(RcLds &t, int tid, int nthreads) {
    constexpr mono::RcTable rc = {{MONO_RC_TABLE}};
    for (int i = tid; i < 60; i += nthreads) t.v[i / 12][i % 12] = rc.v[i / 12][i % 12];
}

// Slot position of lane q's word in slot k (idle lanes q >= A write to a private spare
// position above the doubled state, so they never disturb the words of the active lanes).
template <int A>
__device__ __forceinline__ int wpos(int q, int k) {
    return q < A ? q + A * k : 24 + (q - A);
}

// Writes this lane's words at their positions p and p + 12 (one two-address store each).
template <int A>
__device__ __forceinline__ void put(const u64 (&x)[Form<A>::W], int q, u64 *st) {
#pragma unroll
    for (int k = 0; k < Form<A>::W; k++) {
        const int p = wpos<A>(q, k);
        st[p] = x[k];
        st[p + 12] = x[k];
    }
}

// Post-bar words to the slot (both copies), then each lane's w - 1 neighbours back.
template <int A>
__device__ __forceinline__ void bricks(u64 (&x)[Form<A>::W], int q, u64 *st) {
    constexpr int W = Form<A>::W;
    put<A>(x, q, st);
    __builtin_amdgcn_wave_barrier();
    u64 prev[W];
#pragma unroll
    for (int k = 0; k < W; k++) prev[k] = st[(q < A ? q : 0) + A * k + 11];
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int k = 0; k < W; k++) {
        const u64 y = mono::sqr_add(x[k], prev[k]);
        // word 0 (lane 0, slot 0) has no brick
        x[k] = (k == 0 && q == 0) ? x[k] : y;
    }
}

// y_r = sum_d ROW0[d] x_{(r+d) mod 12} + rc on 32-bit halves (as mono::concrete); own words
// below 4 are made canonical (the next bar inputs, or the digest).
template <int A, int ROUND>
__device__ __forceinline__ void concrete(u64 (&x)[Form<A>::W], int q, u64 *st, const RcLds &rc) {
    constexpr int W = Form<A>::W;
    constexpr int RUN = 24 - A;
    put<A>(x, q, st);
    __builtin_amdgcn_wave_barrier();
    const int base = q < A ? q : 0;
    u64 v[RUN];
#pragma unroll
    for (int e = 0; e < RUN; e++) v[e] = st[base + e];
    __builtin_amdgcn_wave_barrier();
    const u32 m8 = mono::opaque_u32(8);
#pragma unroll
    for (int k = 0; k < W; k++) {
        constexpr bool HAS_RC = ROUND >= 1 && ROUND <= 5;
        const u64 kc = HAS_RC ? rc.v[HAS_RC ? ROUND - 1 : 0][base + A * k] : 0;
        u64 a0 = (u32)kc, b0 = kc >> 32, a1 = 0, b1 = 0;
#pragma unroll
        for (int d = 0; d < 12; d += 2) {
            u32 m0 = mono::row0(d), m1 = mono::row0(d + 1);
            if (m0 == 8) m0 = m8;
            if (m1 == 8) m1 = m8;
            a0 += (u64)(u32)v[A * k + d] * m0;
            b0 += (u64)(u32)(v[A * k + d] >> 32) * m0;
            a1 += (u64)(u32)v[A * k + d + 1] * m1;
            b1 += (u64)(u32)(v[A * k + d + 1] >> 32) * m1;
        }
        const u64 y = mono::fold_halves(a0 + a1, b0 + b1);
        // own word q + A k is below 4 only in slot 0 (A >= 4) or slots 0, 1 (A = 2)
        const bool canon = A == 2 ? k < 2 : (k == 0 && q < 4);
        x[k] = canon ? gl_canon(y) : y;
    }
}

template <int A, int ROUND>
__device__ __forceinline__ void round_fn(u64 (&x)[Form<A>::W], int q, u64 *st, const RcLds &rc) {
    // bars on words 0..3: slot 0 of lanes 0..3 (A >= 4), slots 0 and 1 (A = 2); lanes whose
    // slot 0 is not a bar word (q >= 4) keep their word
    if (A == 2) {
        x[0] = mono::bar(x[0]);
        x[1] = mono::bar(x[1]);
    } else {
        const u64 b = mono::bar(x[0]);
        x[0] = q < 4 ? b : x[0];
    }
    bricks<A>(x, q, st);
    concrete<A, ROUND>(x, q, st, rc);
}

// Full permutation of the group's state. x: this lane's words (any 64-bit representatives);
// q: lane index within the group; st: the group's slot (Form<A>::SLOT words, reused every
// round); rc: the round-constant table. All G lanes of the group must be active. On return
// the digest words 0..3 are canonical.
template <int A>
__device__ __forceinline__ void permute(u64 (&x)[Form<A>::W], int q, u64 *st, const RcLds &rc) {
    concrete<A, 0>(x, q, st, rc);
    round_fn<A, 1>(x, q, st, rc);
    round_fn<A, 2>(x, q, st, rc);
    round_fn<A, 3>(x, q, st, rc);
    round_fn<A, 4>(x, q, st, rc);
    round_fn<A, 5>(x, q, st, rc);
    round_fn<A, 6>(x, q, st, rc);
}

}  // namespace mono_lp
