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
//             computed on 32-bit halves with 64-bit multiply-accumulate
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

// a + b for any 64-bit representatives: a wrap-around of 2^64 is worth 2^32 - 1.
MAM_HD u64 add_nc(u64 a, u64 b) {
    u64 s = a + b;
    if (s < a) {
        s += GL_EPS;
        if (s < GL_EPS) s += GL_EPS;
    }
    return s;
}

// x^2 reduced to a 64-bit representative (not necessarily canonical).
// With x = l + 2^32 h: x^2 = l^2 + 2^33 l h + 2^64 h^2 (three 32x32 products).
MAM_HD u64 sqr_nc(u64 x) {
    const u64 l = (u32)x, h = x >> 32;
    const u64 ll = l * l, lh = l * h, hh2 = h * h;
    u64 lo = ll + (lh << 33);
    u64 hi = hh2 + (lh >> 31) + (lo < ll);
    u64 hh = hi >> 32, hl = hi & GL_EPS;
    u64 t0 = lo - hh;
    if (lo < hh) t0 -= GL_EPS;
    u64 t1 = hl * GL_EPS;
    u64 t2 = t0 + t1;
    if (t2 < t1) t2 += GL_EPS;
    return t2;
}

// lo + 2^64 * hi with hi < 2^32, reduced to a 64-bit representative: 2^64 = 2^32 - 1 (mod p)
MAM_HD u64 fold_hi_nc(u64 lo, u64 hi) {
    u64 t = (hi << 32) - hi;
    u64 s = lo + t;
    if (s < t) s += GL_EPS;
    return s;
}

// Canonical residue of lo + 2^64 * hi with hi < 2^32.
MAM_HD u64 reduce_small_hi(u64 lo, u64 hi) { return gl_canon(fold_hi_nc(lo, hi)); }

// y = a + 2^32 b + k (mod p) for a, b < 2^41 and the 64-bit constant k, as a 64-bit
// representative, with 32-bit add-with-carry steps:
//   y = (a0 + k0) + 2^32 (a1 + b0 + k1) + 2^64 b1 = (w1:w0) + 2^64 w2,
//   then 2^64 w2 = w2 (2^32 - 1) and one wrap-around of the final 64-bit add is 2^32 - 1.
MAM_HD u32 addc32(u32 x, u32 y, u32 cin, u32 &cout) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_addc(x, y, cin, &cout);
#else
    u64 s = (u64)x + y + cin;
    cout = (u32)(s >> 32);
    return (u32)s;
#endif
}

MAM_HD u64 combine(u64 a, u64 b, u64 k) {
    u32 c0, c1, c2;
    const u32 w0 = addc32((u32)a, (u32)k, 0, c0);
    u32 w1 = addc32((u32)b, (u32)(k >> 32), c0, c1);
    w1 = addc32(w1, (u32)(a >> 32), 0, c2);
    const u64 w2 = (u64)((u32)(b >> 32) + c1 + c2);  // < 2^10
    return fold_hi_nc(((u64)w1 << 32) | w0, w2);
}

template <int ROUND>
MAM_HD u64 combine_row(u64 a, u64 b, int r) {
    return combine(a, b, (ROUND >= 1 && ROUND <= 5) ? rc(ROUND, r) : 0);
}

// y_r = sum_c ROW0[(c - r) mod 12] * x_c + rc_r on 32-bit halves of any representatives:
// sum_c m * lo_c and sum_c m * hi_c are each below 12 * 26 * 2^32 < 2^41 (one 32x32+64
// multiply-add per term), then y = A + 2^32 B + rc is folded back below 2^64.
// CANON_ROWS leading rows are made canonical (the next layer's bar inputs).
template <int ROUND, int CANON_ROWS>
MAM_HD void concrete(u64 *x) {
    u32 lo[WIDTH], hi[WIDTH];
#pragma unroll
    for (int c = 0; c < WIDTH; c++) {
        lo[c] = (u32)x[c];
        hi[c] = (u32)(x[c] >> 32);
    }
#pragma unroll
    for (int r = 0; r < WIDTH; r++) {
        u64 a = 0, b = 0;
#pragma unroll
        for (int c = 0; c < WIDTH; c++) {
            const u64 m = row0((c - r + WIDTH) % WIDTH);
            a += (u64)lo[c] * m;
            b += (u64)hi[c] * m;
        }
        x[r] = combine_row<ROUND>(a, b, r);
        if (r < CANON_ROWS) x[r] = gl_canon(x[r]);
    }
}

MAM_HD u64 bar(u64 v) {
    const u64 m1h = 0x8080808080808080ULL, m1l = 0x7F7F7F7F7F7F7F7FULL;
    const u64 m2h = 0xC0C0C0C0C0C0C0C0ULL, m2l = 0x3F3F3F3F3F3F3F3FULL;
    const u64 m3h = 0xE0E0E0E0E0E0E0E0ULL, m3l = 0x1F1F1F1F1F1F1F1FULL;
    u64 nv = ~v;
    u64 r1 = ((nv & m1h) >> 7) | ((nv & m1l) << 1);
    u64 r2 = ((v & m2h) >> 6) | ((v & m2l) << 2);
    u64 r3 = ((v & m3h) >> 5) | ((v & m3l) << 3);
    u64 t = v ^ (r1 & r2 & r3);
    return ((t & m1h) >> 7) | ((t & m1l) << 1);
}

MAM_HD void bricks(u64 *x) {
#pragma unroll
    for (int i = WIDTH - 1; i >= 1; i--) x[i] = add_nc(x[i], sqr_nc(x[i - 1]));
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
