// Poseidon2 permutation, width 8, over Goldilocks (Grassi, Khovratovich, Schofnegger,
// "Poseidon2: A Faster Version of the Poseidon Hash Function", AFRICACRYPT 2023),
// host and device: S-box x^7, 4 + 4 external rounds with M_E = circ(2 M4, M4) where
// M4 = circ(2, 3, 1, 1), 22 internal rounds with M_I = 1 + diag(D).
//
// This is the instance the reference's Fiat-Shamir transcript and optional Merkle hash use
// (zk-torch-3/src/transcript.rs:12-161, cuda/poseidon2.cuh:322-358): its schedule starts
// directly with the first external round (no initial linear layer), and its constants
// (external/internal round constants and the internal diagonal) are data that must be
// identical to the reference's.
#pragma once

#include "field.h"

namespace p2 {

static constexpr int WIDTH = 8;

struct Consts {
    u64 ext[8][8];
    u64 in[22];
    u64 diag[8];
};

__host__ __device__ __forceinline__ const Consts &consts() {
    static constexpr Consts c = {
        {{0xdd5743e7f2a5a5d9ULL, 0xcb3a864e58ada44bULL, 0xd5a3726889dedc0dULL, 0x79365a7f967583c1ULL,
          0xb7820c177b0a3c30ULL, 0x68a60479943b7240ULL, 0x0a22eeab67b97c41ULL, 0xbee9d90e037fa7d4ULL},
         {0xf1dda5b9259dfcb4ULL, 0x27515210be112d59ULL, 0x654a4db8f15f6c3aULL, 0x5a9bfff8c29db0c4ULL,
          0xb72f2628c95a885bULL, 0xffcb54fc166beaddULL, 0xe32f91d59c463772ULL, 0x9e0ff2a9bbea79dbULL},
         {0xce57d6245ddca6b2ULL, 0xb1fc8d402bba1eb1ULL, 0x64df974fb666d528ULL, 0x59e4e5b237e52e78ULL,
          0x5e255e2742caa8fcULL, 0xc4a17304e330ef9aULL, 0x60e3a513cdd9f0b1ULL, 0xbf29a6a5a0c9c8baULL},
         {0xcea721cce82fb11bULL, 0xe5b55eb8098ece81ULL, 0x4ec66bfb89f7c380ULL, 0xc33be12b6ef7e4faULL,
          0xa41b5f5b84c57d7dULL, 0xb526206b2138a936ULL, 0x9f7a5ac62f16eb4eULL, 0x7de77f405f683aa5ULL},
         {0x014ef1197d341346ULL, 0x9725e20825d07394ULL, 0xc8ff2a22516f3604ULL, 0xde1b51f7cf493493ULL,
          0xa7e3af1a53eadb78ULL, 0x9b7b7ee0ddf9f229ULL, 0xf0a460f2dc649e7dULL, 0xdc4448d02c2cb823ULL},
         {0xaa62c88e0b294011ULL, 0x058eb9d810ce9f74ULL, 0x4e3d7a5403566d24ULL, 0xe8ae66da2f8b7a63ULL,
          0x3fdad3b08dae2e0bULL, 0xec81e61303dd409bULL, 0x3cbaa1cd35d31fc4ULL, 0x847f7806eec88ffaULL},
         {0x98ae09a325893803ULL, 0xf8a6475077968838ULL, 0xdf55b7419443de5bULL, 0xdbe15699e696ca70ULL,
          0x8e54a27d8db00424ULL, 0x7679a45cd9d1e12aULL, 0x2844f52be73e0e2fULL, 0xd41f8eb31ada34f8ULL},
         {0xe9dd318bae1f3961ULL, 0xf7462137299efe1aULL, 0x6dbbe06779e1d573ULL, 0xfe35b05cbe707632ULL,
          0x5d8896b12654fd8cULL, 0x6f96ef47c32d4ae2ULL, 0xb0caa221dbbfc0daULL, 0x0bc2a5bf1f238d3fULL}},
        {0x488897d85ff51f56ULL, 0x1140737ccb162218ULL, 0xa7eeb9215866ed35ULL, 0x9bd2976fee49fcc9ULL,
         0xc0c8f0de580a3fccULL, 0x4fb2dae6ee8fc793ULL, 0x343a89f35f37395bULL, 0x223b525a77ca72c8ULL,
         0x56ccb62574aaa918ULL, 0xc4d507d8027af9edULL, 0xa080673cf0b7e95cULL, 0xf0184884eb70dcf8ULL,
         0x044f10b0cb3d5c69ULL, 0xe9e3f7993938f186ULL, 0x1b761c80e772f459ULL, 0x606cec607a1b5facULL,
         0x14a0c2e1d45f03cdULL, 0x4eace8855398574fULL, 0xf905ca7103eff3e6ULL, 0xf8c8f8d20862c059ULL,
         0xb524fe8bdd678e5aULL, 0xfbb7865901a1ec41ULL},
        {0xa98811a1fed4e3a5ULL, 0x1cc48b54f377e2a0ULL, 0xe40cd4f6c5609a26ULL, 0x11de79ebca97a4a3ULL,
         0x9177c73d8b7e929cULL, 0x2a6fe8085797e791ULL, 0x3de6e93329f8d5adULL, 0x3f7af9125da962feULL},
    };
    return c;
}

MAM_HD u64 sbox(u64 x) {
    u64 x2 = gl_sqr(x);
    u64 x3 = gl_mul(x2, x);
    u64 x4 = gl_sqr(x2);
    return gl_mul(x3, x4);
}

// M4 = circ(2, 3, 1, 1) applied to x[0..4)
MAM_HD void m4(u64 *x) {
    u64 s = gl_add(gl_add(x[0], x[1]), gl_add(x[2], x[3]));
    u64 y0 = gl_add(gl_add(s, x[0]), gl_double(x[1]));  // 2 x0 + 3 x1 + x2 + x3
    u64 y1 = gl_add(gl_add(s, x[1]), gl_double(x[2]));  // x0 + 2 x1 + 3 x2 + x3
    u64 y2 = gl_add(gl_add(s, x[2]), gl_double(x[3]));  // x0 + x1 + 2 x2 + 3 x3
    u64 y3 = gl_add(gl_add(s, x[3]), gl_double(x[0]));  // 3 x0 + x1 + x2 + 2 x3
    x[0] = y0;
    x[1] = y1;
    x[2] = y2;
    x[3] = y3;
}

// External layer circ(2 M4, M4): y_i = M4 x_block(i) + sum of both blocks' M4 outputs at i mod 4
MAM_HD void external_linear(u64 *x) {
    m4(x);
    m4(x + 4);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        u64 t = gl_add(x[i], x[i + 4]);
        x[i] = gl_add(x[i], t);
        x[i + 4] = gl_add(x[i + 4], t);
    }
}

MAM_HD void internal_linear(u64 *x) {
    const Consts &k = consts();
    u64 s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s = gl_add(s, x[i]);
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = gl_add(gl_mul(x[i], k.diag[i]), s);
}

// Accepts any 64-bit words; inputs are reduced first.
MAM_HD void permute(u64 *x) {
    const Consts &k = consts();
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = gl_canon(x[i]);
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = sbox(gl_add(x[i], k.ext[r][i]));
        external_linear(x);
    }
#pragma unroll
    for (int r = 0; r < 22; r++) {
        x[0] = sbox(gl_add(x[0], k.in[r]));
        internal_linear(x);
    }
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = sbox(gl_add(x[i], k.ext[4 + r][i]));
        external_linear(x);
    }
}

MAM_HD void hash_gl_leaf(u64 a, u64 b, u64 *out4) {
    u64 s[8] = {a, b, 0, 0, 0, 0, 0, 0};
    permute(s);
    for (int i = 0; i < 4; i++) out4[i] = s[i];
}

MAM_HD void hash_ext2_leaf(const u64 *ab, u64 *out4) {
    u64 s[8] = {ab[0], ab[1], ab[2], ab[3], 0, 0, 0, 0};
    permute(s);
    for (int i = 0; i < 4; i++) out4[i] = s[i];
}

MAM_HD void compress(const u64 *l4, const u64 *r4, u64 *out4) {
    u64 s[8] = {l4[0], l4[1], l4[2], l4[3], r4[0], r4[1], r4[2], r4[3]};
    permute(s);
    for (int i = 0; i < 4; i++) out4[i] = s[i];
}

}  // namespace p2
