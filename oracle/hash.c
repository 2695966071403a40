/*
 * oracle/hash.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Poseidon2 width 8 (the zk-torch-3 Fiat-Shamir permutation) and Monolith
 * width 12 (the default Merkle hash of goldilocks-cuda-rs), restated in scalar C
 * on canonical residues.  Constants are the reference's published tables.
 *
 *   Poseidon2-w8 permutation  zk-torch-3/src/transcript.rs:12-161
 *   Monolith-w12 permutation  goldilocks-cuda-rs/src/cpu_monolith.rs:18-167
 *   Monolith leaf / compress  goldilocks-cuda-rs/src/cpu_monolith.rs:173-230
 *   Poseidon2 Merkle leaf / compress (GPU semantics, one RC add per round)
 *                             goldilocks-cuda-rs/cuda/wrapper.cu:1604-1672, cuda/poseidon2.cuh:322-358,500
 */
#include <string.h>
#include "gl.h"
#include "oracle.h"

/* ------------------------------------------------------------------ */
/* Poseidon2, width 8: 4 external + 22 internal + 4 external rounds    */
/* ------------------------------------------------------------------ */

static const u64 P2_RC_EXT[8][8] = {
    {0xdd5743e7f2a5a5d9ULL, 0xcb3a864e58ada44bULL, 0xd5a3726889dedc0dULL, 0x79365a7f967583c1ULL,
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
     0x5d8896b12654fd8cULL, 0x6f96ef47c32d4ae2ULL, 0xb0caa221dbbfc0daULL, 0x0bc2a5bf1f238d3fULL},
};

static const u64 P2_RC_INT[22] = {
    0x488897d85ff51f56ULL, 0x1140737ccb162218ULL, 0xa7eeb9215866ed35ULL, 0x9bd2976fee49fcc9ULL,
    0xc0c8f0de580a3fccULL, 0x4fb2dae6ee8fc793ULL, 0x343a89f35f37395bULL, 0x223b525a77ca72c8ULL,
    0x56ccb62574aaa918ULL, 0xc4d507d8027af9edULL, 0xa080673cf0b7e95cULL, 0xf0184884eb70dcf8ULL,
    0x044f10b0cb3d5c69ULL, 0xe9e3f7993938f186ULL, 0x1b761c80e772f459ULL, 0x606cec607a1b5facULL,
    0x14a0c2e1d45f03cdULL, 0x4eace8855398574fULL, 0xf905ca7103eff3e6ULL, 0xf8c8f8d20862c059ULL,
    0xb524fe8bdd678e5aULL, 0xfbb7865901a1ec41ULL,
};

static const u64 P2_DIAG[8] = {
    0xa98811a1fed4e3a5ULL, 0x1cc48b54f377e2a0ULL, 0xe40cd4f6c5609a26ULL, 0x11de79ebca97a4a3ULL,
    0x9177c73d8b7e929cULL, 0x2a6fe8085797e791ULL, 0x3de6e93329f8d5adULL, 0x3f7af9125da962feULL,
};

static inline u64 p2_sbox(u64 x) {
    u64 x2 = gl_mul(x, x), x4 = gl_mul(x2, x2), x3 = gl_mul(x2, x);
    return gl_mul(x4, x3);
}

/* circ(2,3,1,1) on 4 lanes */
static void p2_mds4(u64 *s) {
    u64 t01 = gl_add(s[0], s[1]), t23 = gl_add(s[2], s[3]);
    u64 t0123 = gl_add(t01, t23);
    u64 t01123 = gl_add(t0123, s[1]), t01233 = gl_add(t0123, s[3]);
    u64 n0 = gl_add(t01123, t01);
    u64 n1 = gl_add(t01123, gl_add(s[2], s[2]));
    u64 n2 = gl_add(t01233, t23);
    u64 n3 = gl_add(t01233, gl_add(s[0], s[0]));
    s[0] = n0; s[1] = n1; s[2] = n2; s[3] = n3;
}

static void p2_mds_light(u64 *s) {
    p2_mds4(s);
    p2_mds4(s + 4);
    for (int i = 0; i < 4; i++) {
        u64 sum = gl_add(s[i], s[i + 4]);
        s[i] = gl_add(s[i], sum);
        s[i + 4] = gl_add(s[i + 4], sum);
    }
}

void or_poseidon2_permute(u64 *s, int initial_linear_layer) {
    for (int i = 0; i < 8; i++) s[i] = gl_canon(s[i]);
    if (initial_linear_layer) p2_mds_light(s); /* Plonky3 schedule; the reference omits it */
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < 8; i++) s[i] = p2_sbox(gl_add(s[i], P2_RC_EXT[r][i]));
        p2_mds_light(s);
    }
    for (int r = 0; r < 22; r++) {
        s[0] = p2_sbox(gl_add(s[0], P2_RC_INT[r]));
        u64 sum = 0;
        for (int i = 0; i < 8; i++) sum = gl_add(sum, s[i]);
        for (int i = 0; i < 8; i++) s[i] = gl_add(gl_mul(s[i], P2_DIAG[i]), sum);
    }
    for (int r = 0; r < 4; r++) {
        for (int i = 0; i < 8; i++) s[i] = p2_sbox(gl_add(s[i], P2_RC_EXT[4 + r][i]));
        p2_mds_light(s);
    }
}

/* ------------------------------------------------------------------ */
/* Monolith, width 12, 6 rounds, 8-bit lookup bars                     */
/* ------------------------------------------------------------------ */

static const u64 MONO_RC[7][12] = {
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    {13596126580325903823ULL, 5676126986831820406ULL, 11349149288412960427ULL, 3368797843020733411ULL,
     16240671731749717664ULL, 9273190757374900239ULL, 14446552112110239438ULL, 4033077683985131644ULL,
     4291229347329361293ULL, 13231607645683636062ULL, 1383651072186713277ULL, 8898815177417587567ULL},
    {2383619671172821638ULL, 6065528368924797662ULL, 16737578966352303081ULL, 2661700069680749654ULL,
     7414030722730336790ULL, 18124970299993404776ULL, 9169923000283400738ULL, 15832813151034110977ULL,
     16245117847613094506ULL, 11056181639108379773ULL, 10546400734398052938ULL, 8443860941261719174ULL},
    {15799082741422909885ULL, 13421235861052008152ULL, 15448208253823605561ULL, 2540286744040770964ULL,
     2895626806801935918ULL, 8644593510196221619ULL, 17722491003064835823ULL, 5166255496419771636ULL,
     1015740739405252346ULL, 4400043467547597488ULL, 5176473243271652644ULL, 4517904634837939508ULL},
    {18341030605319882173ULL, 13366339881666916534ULL, 6291492342503367536ULL, 10004214885638819819ULL,
     4748655089269860551ULL, 1520762444865670308ULL, 8393589389936386108ULL, 11025183333304586284ULL,
     5993305003203422738ULL, 458912836931247573ULL, 5947003897778655410ULL, 17184667486285295106ULL},
    {15710528677110011358ULL, 8929476121507374707ULL, 2351989866172789037ULL, 11264145846854799752ULL,
     14924075362538455764ULL, 10107004551857451916ULL, 18325221206052792232ULL, 16751515052585522105ULL,
     15305034267720085905ULL, 15639149412312342017ULL, 14624541102106656564ULL, 3542311898554959098ULL},
    {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
};

/* circulant: row r, column c holds MONO_ROW0[(c - r) mod 12] */
static const u64 MONO_ROW0[12] = {7, 23, 8, 26, 13, 10, 9, 7, 6, 22, 21, 8};

static inline u64 mono_bar(u64 x) {
    /* per byte: y = x ^ (rotl1(~x) & rotl2(x) & rotl3(x)), then rotl1 */
    u64 l1 = ((~x & 0x8080808080808080ULL) >> 7) | ((~x & 0x7F7F7F7F7F7F7F7FULL) << 1);
    u64 l2 = ((x & 0xC0C0C0C0C0C0C0C0ULL) >> 6) | ((x & 0x3F3F3F3F3F3F3F3FULL) << 2);
    u64 l3 = ((x & 0xE0E0E0E0E0E0E0E0ULL) >> 5) | ((x & 0x1F1F1F1F1F1F1F1FULL) << 3);
    u64 t = x ^ (l1 & l2 & l3);
    return ((t & 0x8080808080808080ULL) >> 7) | ((t & 0x7F7F7F7F7F7F7F7FULL) << 1);
}

static void mono_concrete(u64 *s, int round) {
    u64 out[12];
    for (int r = 0; r < 12; r++) {
        u128 acc = MONO_RC[round][r];
        for (int c = 0; c < 12; c++) acc += (u128)s[c] * MONO_ROW0[(c - r + 12) % 12];
        out[r] = gl_reduce128(acc);
    }
    memcpy(s, out, sizeof(out));
}

void or_monolith_permute(u64 *s) {
    mono_concrete(s, 0);               /* canonical from here on */
    for (int r = 1; r <= 6; r++) {
        for (int i = 0; i < 4; i++) s[i] = mono_bar(s[i]);  /* bars act on canonical bits */
        for (int i = 11; i >= 1; i--) s[i] = gl_add(s[i], gl_mul(s[i - 1], s[i - 1]));
        mono_concrete(s, r);
    }
}

void or_monolith_hash_gl_leaf(u64 a, u64 b, u64 *out4) {
    u64 s[12] = {0};
    s[0] = a; s[1] = b;
    or_monolith_permute(s);
    memcpy(out4, s, 4 * sizeof(u64));
}

void or_monolith_hash_ext2_leaf(const u64 *a2, const u64 *b2, u64 *out4) {
    u64 s[12] = {0};
    s[0] = a2[0]; s[1] = a2[1]; s[2] = b2[0]; s[3] = b2[1];
    or_monolith_permute(s);
    memcpy(out4, s, 4 * sizeof(u64));
}

void or_monolith_compress(const u64 *l4, const u64 *r4, u64 *out4) {
    u64 s[12] = {0};
    memcpy(s, l4, 4 * sizeof(u64));
    memcpy(s + 4, r4, 4 * sizeof(u64));
    or_monolith_permute(s);
    memcpy(out4, s, 4 * sizeof(u64));
}

/* Poseidon2-w8 Merkle hashing as the GPU tree kernels compute it */
void or_poseidon2_hash_gl_leaf(u64 a, u64 b, u64 *out4) {
    u64 s[8] = {0};
    s[0] = a; s[1] = b;
    or_poseidon2_permute(s, 0);
    memcpy(out4, s, 4 * sizeof(u64));
}

void or_poseidon2_hash_ext2_leaf(const u64 *a2, const u64 *b2, u64 *out4) {
    u64 s[8] = {0};
    s[0] = a2[0]; s[1] = a2[1]; s[2] = b2[0]; s[3] = b2[1];
    or_poseidon2_permute(s, 0);
    memcpy(out4, s, 4 * sizeof(u64));
}

void or_poseidon2_compress(const u64 *l4, const u64 *r4, u64 *out4) {
    u64 s[8];
    memcpy(s, l4, 4 * sizeof(u64));
    memcpy(s + 4, r4, 4 * sizeof(u64));
    or_poseidon2_permute(s, 0);
    memcpy(out4, s, 4 * sizeof(u64));
}

/* ------------------------------------------------------------------ */
/* Merkle trees: layout [N leaves][N/2]...[1] x 4 u64                  */
/* (goldilocks-cuda-rs/src/merkle.rs:1-25, 267-280)                    */
/* ------------------------------------------------------------------ */

static void tree_upper(u64 *tree, long num_leaves, int hash) {
    long off = 0, n = num_leaves;
    while (n > 1) {
        u64 *prev = tree + off * 4, *next = tree + (off + n) * 4;
#pragma omp parallel for schedule(static) if (n > 8192)
        for (long i = 0; i < n / 2; i++) {
            if (hash == OR_HASH_MONOLITH)
                or_monolith_compress(prev + 8 * i, prev + 8 * i + 4, next + 4 * i);
            else
                or_poseidon2_compress(prev + 8 * i, prev + 8 * i + 4, next + 4 * i);
        }
        off += n;
        n /= 2;
    }
}

void or_merkle_tree_gl(const u64 *cw, long num_leaves, int hash, u64 *tree) {
#pragma omp parallel for schedule(static) if (num_leaves > 4096)
    for (long i = 0; i < num_leaves; i++) {
        if (hash == OR_HASH_MONOLITH) or_monolith_hash_gl_leaf(cw[2 * i], cw[2 * i + 1], tree + 4 * i);
        else or_poseidon2_hash_gl_leaf(cw[2 * i], cw[2 * i + 1], tree + 4 * i);
    }
    tree_upper(tree, num_leaves, hash);
}

void or_merkle_tree_ext2(const u64 *cw, long num_leaves, int hash, u64 *tree) {
#pragma omp parallel for schedule(static) if (num_leaves > 4096)
    for (long i = 0; i < num_leaves; i++) {
        if (hash == OR_HASH_MONOLITH) or_monolith_hash_ext2_leaf(cw + 4 * i, cw + 4 * i + 2, tree + 4 * i);
        else or_poseidon2_hash_ext2_leaf(cw + 4 * i, cw + 4 * i + 2, tree + 4 * i);
    }
    tree_upper(tree, num_leaves, hash);
}

long or_merkle_root_offset(long num_leaves) { return (2 * num_leaves - 2) * 4; }

/* Walk an authentication path to the root (cpu_monolith.rs:213-230). */
int or_merkle_verify_path(const u64 *leaf4, const u64 *path, int path_len, long index,
                          const u64 *root4, int hash) {
    u64 cur[4], nxt[4];
    memcpy(cur, leaf4, sizeof(cur));
    for (int l = 0; l < path_len; l++) {
        const u64 *sib = path + 4 * l;
        if (hash == OR_HASH_MONOLITH) {
            if (index & 1) or_monolith_compress(sib, cur, nxt);
            else or_monolith_compress(cur, sib, nxt);
        } else {
            if (index & 1) or_poseidon2_compress(sib, cur, nxt);
            else or_poseidon2_compress(cur, sib, nxt);
        }
        memcpy(cur, nxt, sizeof(cur));
        index >>= 1;
    }
    for (int i = 0; i < 4; i++)
        if (gl_canon(cur[i]) != gl_canon(root4[i])) return 0;
    return 1;
}
