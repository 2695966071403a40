/*
 * oracle/ajtai.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * ChaCha block function (RFC 8439 section 2.3 layout, round count as a parameter) and
 * the int8-matrix Ajtai commitment c = M z (mod p) with
 *   M_ij = (int8) byte (j mod 64) of ChaCha8(key, counter = j / 64, nonce = (i, 0, 0)).
 * The block function follows cuda_almost_goldilocks/ajtai_chacha8.cuh:60-106 (the
 * reference's ChaCha8); it is pinned by the RFC 8439 section 2.3.2 ChaCha20 vector.
 * The commitment has no reference function (the benchmark's third configuration); it is
 * computed here as an exact integer matrix-vector product reduced mod p at the end.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "gl.h"
#include "oracle.h"

static uint32_t rotl32(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }
#define QR(a, b, c, d) \
    a += b; d ^= a; d = rotl32(d, 16); c += d; b ^= c; b = rotl32(b, 12); \
    a += b; d ^= a; d = rotl32(d, 8);  c += d; b ^= c; b = rotl32(b, 7);

void or_chacha_block(int rounds, const uint32_t *key8, uint32_t counter, const uint32_t *nonce3, uint32_t *out16) {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int i = 0; i < 8; i++) s[4 + i] = key8[i];
    s[12] = counter;
    s[13] = nonce3[0]; s[14] = nonce3[1]; s[15] = nonce3[2];
    uint32_t x[16];
    memcpy(x, s, sizeof(x));
    for (int r = 0; r < rounds / 2; r++) {
        QR(x[0], x[4], x[8], x[12]); QR(x[1], x[5], x[9], x[13]);
        QR(x[2], x[6], x[10], x[14]); QR(x[3], x[7], x[11], x[15]);
        QR(x[0], x[5], x[10], x[15]); QR(x[1], x[6], x[11], x[12]);
        QR(x[2], x[7], x[8], x[13]); QR(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) out16[i] = x[i] + s[i];
}

static void key_words(const uint8_t *key32, uint32_t *kw) {
    for (int k = 0; k < 8; k++)
        kw[k] = (uint32_t)key32[4 * k] | ((uint32_t)key32[4 * k + 1] << 8) | ((uint32_t)key32[4 * k + 2] << 16) |
                ((uint32_t)key32[4 * k + 3] << 24);
}

void or_ajtai_i8_row(const uint8_t *key32, uint32_t row, uint64_t n, int8_t *out) {
    uint32_t kw[8], w[16], nonce[3] = {row, 0, 0};
    key_words(key32, kw);
    for (uint64_t b = 0; 64 * b < n; b++) {
        or_chacha_block(8, kw, (uint32_t)b, nonce, w);
        for (uint64_t t = 0; t < 64 && 64 * b + t < n; t++) out[64 * b + t] = (int8_t)(w[t / 4] >> (8 * (t % 4)));
    }
}

/* out[w * rows + i] = sum_j M_ij z_w[j] mod p for batch witnesses of n words */
void or_ajtai_i8_commit(const uint8_t *key32, int rows, uint64_t n, const uint64_t *z, int batch, uint64_t *out) {
#pragma omp parallel
    {
    int8_t *m = (int8_t *)malloc(n);
#pragma omp for schedule(dynamic)
    for (int i = 0; i < rows; i++) {
        or_ajtai_i8_row(key32, (uint32_t)i, n, m);
        for (int w = 0; w < batch; w++) {
            __int128 acc = 0;
            for (uint64_t j = 0; j < n; j++) acc += (__int128)m[j] * (__int128)gl_canon(z[(uint64_t)w * n + j]);
            __int128 r = acc % (__int128)GL_P;
            if (r < 0) r += GL_P;
            out[(uint64_t)w * rows + i] = (uint64_t)r;
        }
    }
    free(m);
    }
}
