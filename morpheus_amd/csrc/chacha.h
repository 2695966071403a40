// ChaCha block function (Bernstein; RFC 8439 section 2.3 layout), host and device:
// state = 4 constants "expand 32-byte k", 8 key words, block counter (word 12), 3 nonce
// words; ROUNDS/2 double rounds of column + diagonal quarter rounds; output = state +
// input. ROUNDS = 8 is the ChaCha8 the reference's Ajtai matrices use
// (cuda_almost_goldilocks/ajtai_chacha8.cuh:60-106); ROUNDS = 20 is the RFC's ChaCha20,
// whose published test vector pins this code.
#pragma once

#include "field.h"

namespace chacha {

MAM_HD u32 rotl(u32 x, int n) { return (x << n) | (x >> (32 - n)); }

MAM_HD void qr(u32 &a, u32 &b, u32 &c, u32 &d) {
    a += b; d ^= a; d = rotl(d, 16);
    c += d; b ^= c; b = rotl(b, 12);
    a += b; d ^= a; d = rotl(d, 8);
    c += d; b ^= c; b = rotl(b, 7);
}

template <int ROUNDS>
MAM_HD void block(const u32 key[8], u32 counter, const u32 nonce[3], u32 out[16]) {
    u32 s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u, key[0], key[1], key[2], key[3],
                 key[4],      key[5],      key[6],      key[7],      counter, nonce[0], nonce[1], nonce[2]};
    u32 x[16];
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = s[i];
#pragma unroll
    for (int r = 0; r < ROUNDS / 2; r++) {
        qr(x[0], x[4], x[8], x[12]);
        qr(x[1], x[5], x[9], x[13]);
        qr(x[2], x[6], x[10], x[14]);
        qr(x[3], x[7], x[11], x[15]);
        qr(x[0], x[5], x[10], x[15]);
        qr(x[1], x[6], x[11], x[12]);
        qr(x[2], x[7], x[8], x[13]);
        qr(x[3], x[4], x[9], x[14]);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) out[i] = x[i] + s[i];
}

}  // namespace chacha
