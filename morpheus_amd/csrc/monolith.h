// Monolith-64 permutation, width 12, 6 rounds, 8-bit bars (Grassi et al., "Monolith:
// Circuit-Friendly Hash Functions with New Nonlinear Layers for Fast and Constant-Time
// Implementations", ToSC 2024), over Goldilocks, host and device.
//
// The reference's instance: goldilocks-cuda-rs/src/cpu_monolith.rs:18-167 (constants
// :18-60, bars :100-112, bricks :115-121, concrete :124-135, permutation :141-148) and
// cuda/monolith.cuh:94-216; its Merkle leaf / node hashing: cpu_monolith.rs:155-210,
// cuda/monolith.cuh:228-314. The round-constant table and the circulant MDS row are data
// that must be identical to the reference's; they are the published Monolith-64 values.
//
// Layout of the computation here (all integer, exact):
//   concrete: y_r = sum_c ROW0[(c - r) mod 12] * x_c + RC[round][r]  (mod p)
//             computed on 32-bit halves with 32x32+64 multiply-adds
//   bars:     per byte b: b ^ (rotl1(~b) & rotl2(b) & rotl3(b)), then rotl1, on x_0..x_3
//   bricks:   x_i += x_{i-1}^2 for i = 11..1 (old values)
#pragma once

#include "field.h"

namespace mono {

static constexpr int WIDTH = 12;
static constexpr int ROUNDS = 6;

// First row of the circulant MDS matrix: M[r][c] = ROW0[(c - r) mod 12].
__host__ __device__ constexpr u32 row0(int k) {
    return k == 0 ? 7 : k == 1 ? 23 : k == 2 ? 8 : k == 3 ? 26 : k == 4 ? 13 : k == 5 ? 10
         : k == 6 ? 9 : k == 7 ? 7 : k == 8 ? 6 : k == 9 ? 22 : k == 10 ? 21 : 8;
}

// Round constants for rounds 1..5 (round 0 and round 6 add zero).
#define MONO_RC_TABLE                                                                           \
    {13596126580325903823ULL, 5676126986831820406ULL, 11349149288412960427ULL,                 \
     3368797843020733411ULL, 16240671731749717664ULL, 9273190757374900239ULL,                  \
     14446552112110239438ULL, 4033077683985131644ULL, 4291229347329361293ULL,                  \
     13231607645683636062ULL, 1383651072186713277ULL, 8898815177417587567ULL},                 \
    {2383619671172821638ULL, 6065528368924797662ULL, 16737578966352303081ULL,                  \
     2661700069680749654ULL, 7414030722730336790ULL, 18124970299993404776ULL,                  \
     9169923000283400738ULL, 15832813151034110977ULL, 16245117847613094506ULL,                 \
     11056181639108379773ULL, 10546400734398052938ULL, 8443860941261719174ULL},                \
    {15799082741422909885ULL, 13421235861052008152ULL, 15448208253823605561ULL,                \
     2540286744040770964ULL, 2895626806801935918ULL, 8644593510196221619ULL,                   \
     17722491003064835823ULL, 5166255496419771636ULL, 1015740739405252346ULL,                  \
     4400043467547597488ULL, 5176473243271652644ULL, 4517904634837939508ULL},                  \
    {18341030605319882173ULL, 13366339881666916534ULL, 6291492342503367536ULL,                 \
     10004214885638819819ULL, 4748655089269860551ULL, 1520762444865670308ULL,                  \
     8393589389936386108ULL, 11025183333304586284ULL, 5993305003203422738ULL,                  \
     458912836931247573ULL, 5947003897778655410ULL, 17184667486285295106ULL},                  \
    {15710528677110011358ULL, 8929476121507374707ULL, 2351989866172789037ULL,                  \
     11264145846854799752ULL, 14924075362538455764ULL, 10107004551857451916ULL,                \
     18325221206052792232ULL, 16751515052585522105ULL, 15305034267720085905ULL,                \
     15639149412312342017ULL, 14624541102106656564ULL, 3542311898554959098ULL }

struct RcTable {
    u64 v[5][12];
};

__host__ __device__ __forceinline__ u64 rc(int round, int r) {
    // round in 1..5
    constexpr RcTable t = {{MONO_RC_TABLE}};
    return t.v[round - 1][r];
}

// Values inside the permutation are kept as arbitrary 64-bit representatives of their
// residues; only the four bar inputs (bars act on the bits of the canonical value) and the
// outputs are reduced to [0, p).

// y = a + 2^32 b (mod p) as a 64-bit representative, for a, b < 2^42. With
// b = bh 2^32 + bl and 2^64 = 2^32 - 1 (mod p): y = (a + bh (2^32 - 1)) + bl 2^32. The first
// sum stays below 2^43 (one 32x32+64 multiply-add); the second may wrap 2^64 once, and a
// wrap-around is worth 2^32 - 1.
MAM_HD u64 fold_halves(u64 a, u64 b) {
    const u64 t = a + (b >> 32) * GL_EPS;
    const u64 s = t + (b << 32);
    return s < t ? s + GL_EPS : s;
}

// A multiplier the compiler has to treat as unknown: x * 8 then stays one 32x32+64
// multiply-add instead of becoming a shift of a zero-extended word plus an add.
MAM_HD u32 opaque_u32(u32 v) {
#ifdef __HIP_DEVICE_COMPILE__
    asm("" : "+s"(v));
#endif
    return v;
}

// y_r = sum_c ROW0[(c - r) mod 12] * x_c + rc_r on 32-bit halves of any representatives:
// the round constant's halves start the two accumulators, and sum_c m * lo_c, sum_c m * hi_c
// (one 32x32+64 multiply-add per term) stay below 161 * 2^32 < 2^40 (the row sums to 160),
// then y = A + 2^32 B is folded back below 2^64. CANON_ROWS leading rows are made canonical
// (the next layer's bar inputs).
template <int ROUND, int CANON_ROWS>
MAM_HD void concrete(u64 *x) {
    u32 lo[WIDTH], hi[WIDTH];
    const u32 m8 = opaque_u32(8);
#pragma unroll
    for (int c = 0; c < WIDTH; c++) {
        lo[c] = (u32)x[c];
        hi[c] = (u32)(x[c] >> 32);
    }
#pragma unroll
    for (int r = 0; r < WIDTH; r++) {
        const u64 k = (ROUND >= 1 && ROUND <= 5) ? rc(ROUND, r) : 0;
        u64 a = (u32)k, b = k >> 32;
#pragma unroll
        for (int c = 0; c < WIDTH; c++) {
            u32 m = row0((c - r + WIDTH) % WIDTH);
            if (m == 8) m = m8;
            a += (u64)lo[c] * m;
            b += (u64)hi[c] * m;
        }
        x[r] = fold_halves(a, b);
        if (r < CANON_ROWS) x[r] = gl_canon(x[r]);
    }
}

// Each byte of w rotated left by k: the byte's top 8 - k bits move up (mask m keeps them),
// its low k bits come round from the top (a bit-insert of the two shifts).
MAM_HD u32 rotl_bytes(u32 w, int k, u32 m) { return ((w << k) & m) | ((w >> (8 - k)) & ~m); }

// One bar on a 32-bit half: per byte b ^ (rotl1(~b) & rotl2(b) & rotl3(b)), then rotl1;
// rotl1(~b) = ~rotl1(b). Bytes never straddle the halves, so each half is done on its own.
MAM_HD u32 bar_half(u32 v) {
    const u32 r1 = rotl_bytes(v, 1, 0xFEFEFEFEu), r2 = rotl_bytes(v, 2, 0xFCFCFCFCu);
    const u32 r3 = rotl_bytes(v, 3, 0xF8F8F8F8u);
    return rotl_bytes(v ^ (~r1 & r2 & r3), 1, 0xFEFEFEFEu);
}

MAM_HD u64 bar(u64 v) { return ((u64)bar_half((u32)(v >> 32)) << 32) | bar_half((u32)v); }

// y + x^2 (mod p) as a 64-bit representative, for any representatives x, y.
// With x = l + 2^32 h: y + x^2 = y + l^2 + 2^33 lh + 2^64 h^2 = lo + 2^64 hi (hi < 2^64, three
// 32x32 products and two carries). Then 2^64 = 2^32 - 1 and 2^96 = -1 (mod p) give, with
// hi = H1 2^32 + H0, lo + 2^64 hi = lo + H0 (2^32 - 1) - H1. t = H0 (2^32 - 1) + (2^32 - 1 - H1)
// lies in [0, 2^64 - 2^32], so the result is lo + t - (2^32 - 1): a wrap-around of lo + t
// (worth 2^32 - 1) cancels the subtraction, and a borrow of the subtraction (worth
// -(2^32 - 1)) is paid by subtracting 2^32 - 1 once more.
MAM_HD u64 sqr_add(u64 y, u64 x) {
    const u64 l = (u32)x, h = x >> 32;
    const u64 lh = l * h, hh = h * h;
    const u64 a = l * l + y;
    u64 hi = hh + (lh >> 31) + (a < y);
    const u64 lo = a + (lh << 33);
    hi += (lo < a);
    const u64 t = (hi & GL_EPS) * GL_EPS + ((~(hi >> 32)) & GL_EPS);
    const u64 r = lo + t;
    if (r < lo) return r;
    const u64 s = r - GL_EPS;
    return r < GL_EPS ? s - GL_EPS : s;
}

MAM_HD void bricks(u64 *x) {
#pragma unroll
    for (int i = WIDTH - 1; i >= 1; i--) x[i] = sqr_add(x[i], x[i - 1]);
}

// bar inputs x[0..3] are canonical (made so by the previous concrete layer)
template <int ROUND, int CANON_ROWS>
MAM_HD void round_fn(u64 *x) {
#pragma unroll
    for (int i = 0; i < 4; i++) x[i] = bar(x[i]);
    bricks(x);
    concrete<ROUND, CANON_ROWS>(x);
}

// Full permutation. Accepts any 64-bit words (the first concrete layer is linear, so the
// result depends only on the residues). OUT_WORDS leading output words are canonical:
// 12 for the full permutation, 4 where only a digest is kept.
template <int OUT_WORDS = WIDTH>
MAM_HD void permute(u64 *x) {
    concrete<0, 4>(x);
    round_fn<1, 4>(x);
    round_fn<2, 4>(x);
    round_fn<3, 4>(x);
    round_fn<4, 4>(x);
    round_fn<5, 4>(x);
    round_fn<6, OUT_WORDS>(x);
}

// Leaf of a base-field codeword: permute([a, b, 0, ...])[0..4]
MAM_HD void hash_gl_leaf(u64 a, u64 b, u64 *out4) {
    u64 s[WIDTH] = {a, b, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    permute<4>(s);
    for (int i = 0; i < 4; i++) out4[i] = s[i];
}

// Leaf of an Ext2 codeword: permute([a0, a1, b0, b1, 0, ...])[0..4]
MAM_HD void hash_ext2_leaf(const u64 *ab, u64 *out4) {
    u64 s[WIDTH] = {ab[0], ab[1], ab[2], ab[3], 0, 0, 0, 0, 0, 0, 0, 0};
    permute<4>(s);
    for (int i = 0; i < 4; i++) out4[i] = s[i];
}

// Two-to-one node: permute([L0..L3, R0..R3, 0, 0, 0, 0])[0..4]
MAM_HD void compress(const u64 *l4, const u64 *r4, u64 *out4) {
    u64 s[WIDTH] = {l4[0], l4[1], l4[2], l4[3], r4[0], r4[1], r4[2], r4[3], 0, 0, 0, 0};
    permute<4>(s);
    for (int i = 0; i < 4; i++) out4[i] = s[i];
}

}  // namespace mono
