// Ring-Ajtai shared pieces: R = F_q[X]/(X^64 + 1), KAPPA = 42 rows, and the matrix generator
// (cuda_almost_goldilocks/ajtai_chacha8.cuh:115-178): coefficient 8 b + k of M_{i,j} is the k-th
// accepted u64 (< q) of ChaCha8(key, counter = b (+ 8 per retry), nonce = (i, j_lo, j_hi)).
#pragma once

#include "chacha.h"
#include "common.h"
#include "fieldq.h"

static constexpr int RING_D = 64;
static constexpr int RING_KAPPA = 42;
static constexpr int RING_RG = 7;  // rows per group
static constexpr int RING_GROUPS = RING_KAPPA / RING_RG;

// Coefficients 8 blk .. 8 blk + 7 of M_{row, j}: the accepted 64-bit words (< q, in order) of
// ChaCha8 blocks with counters blk, blk + 8, blk + 16, ... (16 tries; unfilled slots 0).
__host__ __device__ __forceinline__ void ring_prg_block(const u32 key[8], u32 row, u64 j, u32 blk, u64 out[8]) {
    const u32 nonce[3] = {row, (u32)j, (u32)(j >> 32)};
    u32 w[16];
    chacha::block<8>(key, blk, nonce, w);
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        out[k] = ((u64)w[2 * k + 1] << 32) | w[2 * k];
        ok = ok && out[k] < fq::Q;
    }
    if (ok) return;
    // rare path: collect accepted words into a shift register (oldest in out[0])
    int written = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) out[k] = 0;
    for (u32 retry = 0; retry < 16 && written < 8; retry++) {
        chacha::block<8>(key, blk + 8 * retry, nonce, w);
        for (int k = 0; k < 8 && written < 8; k++) {
            const u64 s = ((u64)w[2 * k + 1] << 32) | w[2 * k];
            if (s < fq::Q) {
#pragma unroll
                for (int t = 0; t < 7; t++) out[t] = out[t + 1];
                out[7] = s;
                written++;
            }
        }
    }
    // fewer than 8 accepted (never observed): move them to the front, zeros after
    for (int pad = 8 - written; pad > 0; pad--) {
        const u64 first = out[0];
#pragma unroll
        for (int t = 0; t < 7; t++) out[t] = out[t + 1];
        out[7] = first;
    }
}

