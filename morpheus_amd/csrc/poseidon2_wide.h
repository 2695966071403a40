// Lane-parallel Poseidon2-w8 over Goldilocks: a 16-lane DPP row computes one permutation,
// lane i holding state word x_{i & 7} (lanes i and i + 8 hold the same word). The same
// function as p2::permute in poseidon2.h, which carries the reference citations and constants.
//
// Every cross-lane move is a DPP move inside the row: xor 1 and xor 2 (quad_perm), the next
// word of a 4-word block (quad_perm), and word j ^ 4, which is the word 4 lanes away
// (row_ror:4) because the row holds the 8 words twice. External rounds: every lane adds its
// round constant and applies the S-box x^7, then circ(2 M4, M4): y = s + x_k + 2 x_{k+1}
// with s the block sum, x' = 2 y + y_{j^4}. Internal rounds: the S-box of word 0 (x^2 on
// lane 0, then x^3 on lane 0 and x^4 on lane 1 in one instruction stream, then x^7), then
// x_j d_j + sum on every lane. A permutation is about 30 dependent round latencies rather
// than a whole permutation's worth of serial field products on one lane.
#pragma once

#include "poseidon2.h"

namespace p2w {

static constexpr int GROUP = 16;

template <int CTRL>
__device__ __forceinline__ u64 dpp(u64 v) {
    const u32 lo = (u32)__builtin_amdgcn_update_dpp((int)(u32)v, (int)(u32)v, CTRL, 0xF, 0xF, false);
    const u32 hi = (u32)__builtin_amdgcn_update_dpp((int)(u32)(v >> 32), (int)(u32)(v >> 32), CTRL, 0xF, 0xF, false);
    return ((u64)hi << 32) | lo;
}
__device__ __forceinline__ u64 xor1(u64 v) { return dpp<0xB1>(v); }    // quad_perm [1, 0, 3, 2]
__device__ __forceinline__ u64 xor2(u64 v) { return dpp<0x4E>(v); }    // quad_perm [2, 3, 0, 1]
__device__ __forceinline__ u64 next4(u64 v) { return dpp<0x39>(v); }   // quad_perm [1, 2, 3, 0]
__device__ __forceinline__ u64 rev4(u64 v) { return dpp<0x1B>(v); }    // quad_perm [3, 2, 1, 0]
__device__ __forceinline__ u64 xor4(u64 v) { return dpp<0x124>(v); }   // row_ror:4 (word j ^ 4)

// This lane's constants: external round constants ext[r][j] and the internal diagonal d_j.
struct Lane {
    u64 ext[8];
    u64 diag;
};

__device__ __forceinline__ void lane_consts(int j, Lane &L) {
    const p2::Consts &k = p2::consts();
#pragma unroll
    for (int r = 0; r < 8; r++) L.ext[r] = k.ext[r][j];
    L.diag = k.diag[j];
}

__device__ __forceinline__ u64 external_linear(u64 x) {
    u64 s = gl_add(x, xor1(x));
    s = gl_add(s, xor2(s));
    const u64 y = gl_add(gl_add(s, x), gl_double(next4(x)));
    return gl_add(gl_double(y), xor4(y));
}

// x^7 of lane 0's word, the two middle products on lanes 0 and 1 at once
__device__ __forceinline__ u64 sbox0(u64 x, int j) {
    const u64 x2 = gl_sqr(x);
    const u64 a2 = xor1(x2);                     // lane 1 sees lane 0's x^2
    const u64 m = j == 1 ? gl_mul(a2, a2) : gl_mul(x2, x);   // lane 1: x^4, lane 0: x^3
    return gl_mul(m, xor1(m));                   // lane 0: x^3 x^4
}

// x: this lane's word (any 64-bit representative); j = lane & 7. All 16 lanes of the row must
// be active and lanes j, j + 8 must hold the same word. Returns the canonical permuted word.
__device__ __forceinline__ u64 permute(u64 x, int j, const Lane &L) {
    const p2::Consts &k = p2::consts();
    x = gl_canon(x);
#pragma unroll
    for (int r = 0; r < 4; r++) x = external_linear(p2::sbox(gl_add(x, L.ext[r])));
#pragma unroll 2
    for (int r = 0; r < 22; r++) {
        const u64 t = sbox0(gl_add(x, k.in[r]), j);
        if (j == 0) x = t;
        u64 s = gl_add(x, xor1(x));
        s = gl_add(s, xor2(s));
        s = gl_add(s, xor4(s));
        x = gl_add(gl_mul(x, L.diag), s);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) x = external_linear(p2::sbox(gl_add(x, L.ext[4 + r])));
    return x;
}

}  // namespace p2w
