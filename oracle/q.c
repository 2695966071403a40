/*
 * oracle/q.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * The almost-Goldilocks field q = 2^64 - 2^32 - 31 and its extension F_q[X]/(X^2 - 3),
 * computed with 128-bit remainders (independent of the library's Solinas reduction);
 * q-twins of the eq table, partial evaluation and reference-style sumcheck round message
 * and fold; and the ring Ajtai commitment over R = F_q[X]/(X^64 + 1) with KAPPA = 42.
 *
 * Followed files:
 *   field / Ext2: almost-goldilocks-cuda-rs/src/field.rs:11-133, extension.rs:55-86,
 *     cuda_almost_goldilocks/almost_extension.cuh:31-170
 *   eq: almost_eq_lagrange.cuh:25-120 (index bit i <-> r_i)
 *   partial eval: almost_partial_eval.cuh (fold pairs (2y, 2y+1) at r_i, LSB first)
 *   sumcheck: almost_sumcheck_prover.cuh:19-224 (values at c = 0..d), :233-330 (batched)
 *   ring Ajtai: ajtai_cpu_reference.cuh:58-242 (negacyclic shift, naive product, PRG,
 *     full / sparse commit), ajtai_chacha8.cuh:115-178 (rejection sampling order),
 *     almost-goldilocks-cuda-rs/src/ajtai.rs:138-386 (batched, col_offset, wide layout)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

#define QP 0xFFFFFFFEFFFFFFE1ULL
typedef unsigned __int128 u128;

static uint64_t cq(uint64_t a) { return a % QP; }
uint64_t or_q_add(uint64_t a, uint64_t b) { return (uint64_t)(((u128)cq(a) + cq(b)) % QP); }
uint64_t or_q_sub(uint64_t a, uint64_t b) { return (uint64_t)(((u128)cq(a) + QP - cq(b)) % QP); }
uint64_t or_q_mul(uint64_t a, uint64_t b) { return (uint64_t)(((u128)cq(a) * cq(b)) % QP); }
static uint64_t q_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    a = cq(a);
    while (e) {
        if (e & 1) r = or_q_mul(r, a);
        a = or_q_mul(a, a);
        e >>= 1;
    }
    return r;
}
uint64_t or_q_inv(uint64_t a) { return q_pow(a, QP - 2); }

void or_q2_mul(const uint64_t *a, const uint64_t *b, uint64_t *out) {
    /* (a0 + a1 X)(b0 + b1 X), X^2 = 3: schoolbook */
    uint64_t c0 = or_q_add(or_q_mul(a[0], b[0]), or_q_mul(3, or_q_mul(a[1], b[1])));
    uint64_t c1 = or_q_add(or_q_mul(a[0], b[1]), or_q_mul(a[1], b[0]));
    out[0] = c0;
    out[1] = c1;
}
void or_q2_inv(const uint64_t *a, uint64_t *out) {
    uint64_t n = or_q_sub(or_q_mul(a[0], a[0]), or_q_mul(3, or_q_mul(a[1], a[1])));
    uint64_t ni = or_q_inv(n);
    out[0] = or_q_mul(a[0], ni);
    out[1] = or_q_mul(or_q_sub(0, a[1]), ni);
}

static void e_add(uint64_t *r, const uint64_t *a, const uint64_t *b) {
    r[0] = or_q_add(a[0], b[0]);
    r[1] = or_q_add(a[1], b[1]);
}
static void e_sub(uint64_t *r, const uint64_t *a, const uint64_t *b) {
    r[0] = or_q_sub(a[0], b[0]);
    r[1] = or_q_sub(a[1], b[1]);
}

/* eq table, Ext2 (ext = 1, r and out interleaved) or base (ext = 0) */
void or_q_eq_table(int ext, const uint64_t *r, int nv, uint64_t *out) {
    uint64_t n = 1ULL << nv;
    for (uint64_t x = 0; x < n; x++) {
        uint64_t acc[2] = {1, 0};
        for (int i = 0; i < nv; i++) {
            uint64_t ri[2] = {ext ? cq(r[2 * i]) : cq(r[i]), ext ? cq(r[2 * i + 1]) : 0};
            uint64_t f[2];
            if ((x >> i) & 1) {
                f[0] = ri[0];
                f[1] = ri[1];
            } else {
                f[0] = or_q_sub(1, ri[0]);
                f[1] = or_q_sub(0, ri[1]);
            }
            or_q2_mul(acc, f, acc);
        }
        if (ext) {
            out[2 * x] = acc[0];
            out[2 * x + 1] = acc[1];
        } else {
            out[x] = acc[0];
        }
    }
}

/* fold the first m variables of an nv-variable table: in_ext selects base/Ext2 input,
 * r_ext base/Ext2 points; output Ext2 (interleaved) of 2^(nv-m) elements */
void or_q_partial_eval(const uint64_t *in, int in_ext, int nv, const uint64_t *r, int r_ext, int m, uint64_t *out) {
    uint64_t n = 1ULL << nv;
    uint64_t *cur = (uint64_t *)malloc(2 * n * sizeof(uint64_t));
    for (uint64_t x = 0; x < n; x++) {
        cur[2 * x] = in_ext ? cq(in[2 * x]) : cq(in[x]);
        cur[2 * x + 1] = in_ext ? cq(in[2 * x + 1]) : 0;
    }
    for (int i = 0; i < m; i++) {
        uint64_t ri[2] = {r_ext ? cq(r[2 * i]) : cq(r[i]), r_ext ? cq(r[2 * i + 1]) : 0};
        n /= 2;
        for (uint64_t y = 0; y < n; y++) {
            uint64_t d[2], t[2];
            e_sub(d, cur + 4 * y + 2, cur + 4 * y);
            or_q2_mul(ri, d, t);
            e_add(cur + 2 * y, cur + 4 * y, t);
        }
    }
    memcpy(out, cur, 2 * n * sizeof(uint64_t));
    free(cur);
}

/* Sum over y < half of prod_i (a_i + c (b_i - a_i)), c = 0..d, for d polynomials of
 * original_size elements back to back (base: ext = 0; Ext2 interleaved: ext = 1).
 * out: (d + 1) Ext2 values (c1 = 0 for base inputs). */
void or_q_sumcheck_msg(int ext, const uint64_t *polys, int d, uint64_t original_size, uint64_t half, uint64_t *out) {
    for (int c = 0; c <= d; c++) out[2 * c] = out[2 * c + 1] = 0;
    uint64_t stride = ext ? 2 * original_size : original_size;
    for (uint64_t y = 0; y < half; y++) {
        for (int c = 0; c <= d; c++) {
            uint64_t prod[2] = {1, 0};
            for (int i = 0; i < d; i++) {
                const uint64_t *p = polys + i * stride;
                uint64_t a[2] = {ext ? cq(p[4 * y]) : cq(p[2 * y]), ext ? cq(p[4 * y + 1]) : 0};
                uint64_t b[2] = {ext ? cq(p[4 * y + 2]) : cq(p[2 * y + 1]), ext ? cq(p[4 * y + 3]) : 0};
                uint64_t df[2], v[2], cc[2] = {(uint64_t)c, 0};
                e_sub(df, b, a);
                or_q2_mul(cc, df, v);
                e_add(v, a, v);
                or_q2_mul(prod, v, prod);
            }
            e_add(out + 2 * c, out + 2 * c, prod);
        }
    }
}

/* out[i][y] = a + ch (b - a) for pairs (2y, 2y + 1) of each of d polynomials */
void or_q_sumcheck_fold(int ext, const uint64_t *in, uint64_t *out, const uint64_t *ch2, int d,
                        uint64_t original_size, uint64_t half) {
    uint64_t stride = ext ? 2 * original_size : original_size;
    for (int i = 0; i < d; i++) {
        const uint64_t *p = in + i * stride;
        uint64_t *o = out + i * stride;
        for (uint64_t y = 0; y < half; y++) {
            uint64_t a[2] = {ext ? cq(p[4 * y]) : cq(p[2 * y]), ext ? cq(p[4 * y + 1]) : 0};
            uint64_t b[2] = {ext ? cq(p[4 * y + 2]) : cq(p[2 * y + 1]), ext ? cq(p[4 * y + 3]) : 0};
            uint64_t df[2], v[2], chv[2] = {cq(ch2[0]), ext ? cq(ch2[1]) : 0};
            e_sub(df, b, a);
            or_q2_mul(chv, df, v);
            e_add(v, a, v);
            if (ext) {
                o[2 * y] = v[0];
                o[2 * y + 1] = v[1];
            } else {
                o[y] = v[0];
            }
        }
    }
}

/* ---- ring Ajtai ---- */
#define RD 64
#define RKAPPA 42

/* ajtai_chacha8.cuh:115-178: coefficients 8 blk .. 8 blk + 7 */
static void prg_block(const uint32_t *key8, uint32_t row, uint64_t j, uint32_t blk, uint64_t *out8) {
    uint32_t nonce[3] = {row, (uint32_t)j, (uint32_t)(j >> 32)};
    uint32_t w[16];
    int written = 0;
    for (int k = 0; k < 8; k++) out8[k] = 0;
    for (uint32_t retry = 0; retry < 16 && written < 8; retry++) {
        or_chacha_block(8, key8, blk + retry * 8u, nonce, w);
        for (int k = 0; k < 8 && written < 8; k++) {
            uint64_t s = ((uint64_t)w[2 * k + 1] << 32) | w[2 * k];
            if (s < QP) out8[written++] = s;
        }
    }
}

void or_ring_matrix_entry(const uint32_t *key8, uint32_t row, uint64_t j, uint64_t *out64) {
    for (uint32_t b = 0; b < 8; b++) prg_block(key8, row, j, b, out64 + 8 * b);
}

/* out += a * b in F_q[X]/(X^64 + 1) (naive, ajtai_cpu_reference.cuh:84-101) */
static void ring_mul_acc(const uint64_t *a, const uint64_t *b, uint64_t *out) {
    for (int i = 0; i < RD; i++) {
        if (!a[i]) continue;
        for (int k = 0; k < RD; k++) {
            uint64_t v = or_q_mul(a[i], b[k]);
            int r = i + k;
            if (r < RD) out[r] = or_q_add(out[r], v);
            else out[r - RD] = or_q_sub(out[r - RD], v);
        }
    }
}

/* out[b][i][r] for B binary witnesses z[b * N + j] (bit l = coefficient l) */
void or_ring_commit_binary(const uint32_t *key8, const uint64_t *z, uint64_t N, int B, uint64_t col_offset,
                           uint64_t *out) {
    memset(out, 0, (size_t)B * RKAPPA * RD * sizeof(uint64_t));
    uint64_t m[RD], zr[RD];
    for (uint64_t j = 0; j < N; j++) {
        for (int i = 0; i < RKAPPA; i++) {
            or_ring_matrix_entry(key8, (uint32_t)i, j + col_offset, m);
            for (int b = 0; b < B; b++) {
                uint64_t bits = z[(uint64_t)b * N + j];
                if (!bits) continue;
                for (int l = 0; l < RD; l++) zr[l] = (bits >> l) & 1;
                ring_mul_acc(zr, m, out + ((uint64_t)b * RKAPPA + i) * RD);
            }
        }
    }
}

/* z_wide[j * 64 + l]: field coefficients */
void or_ring_commit_wide(const uint32_t *key8, const uint64_t *z_wide, uint64_t N, uint64_t col_offset,
                         uint64_t *out) {
    memset(out, 0, (size_t)RKAPPA * RD * sizeof(uint64_t));
    uint64_t m[RD], zr[RD];
    for (uint64_t j = 0; j < N; j++) {
        for (int l = 0; l < RD; l++) zr[l] = cq(z_wide[j * RD + l]);
        for (int i = 0; i < RKAPPA; i++) {
            or_ring_matrix_entry(key8, (uint32_t)i, j + col_offset, m);
            ring_mul_acc(zr, m, out + (uint64_t)i * RD);
        }
    }
}

/* positions p = 64 j + l of a binary witness (duplicates add) */
void or_ring_commit_sparse(const uint32_t *key8, const uint64_t *pos, uint64_t K, uint64_t *out) {
    memset(out, 0, (size_t)RKAPPA * RD * sizeof(uint64_t));
    uint64_t m[RD], zr[RD];
    for (uint64_t k = 0; k < K; k++) {
        uint64_t j = pos[k] >> 6;
        int l = (int)(pos[k] & 63);
        memset(zr, 0, sizeof(zr));
        zr[l] = 1;
        for (int i = 0; i < RKAPPA; i++) {
            or_ring_matrix_entry(key8, (uint32_t)i, j, m);
            ring_mul_acc(zr, m, out + (uint64_t)i * RD);
        }
    }
}

/* ---- folding, splitting and ternary commits of ring witnesses
 * (cuda_almost_goldilocks/ajtai.cuh:700-1280, ajtai_premat.cuh:35-200, restated with a generic
 * negacyclic product instead of the kernels' rotated look-ups) ---- */

/* out += a * b in Z[X]/(X^64 + 1) on signed integers */
static void ring_mul_int_acc(const int64_t *a, const int64_t *b, int64_t *out) {
    for (int i = 0; i < RD; i++) {
        if (!a[i]) continue;
        for (int k = 0; k < RD; k++) {
            int r = i + k;
            if (r < RD) out[r] += a[i] * b[k];
            else out[r - RD] -= a[i] * b[k];
        }
    }
}

static uint64_t lift_int(int64_t v) {
    if (v >= 0) return (uint64_t)v % QP;
    const uint64_t m = (uint64_t)(-v) % QP;
    return m ? QP - m : 0;
}

static void bits_to_int(uint64_t bits, int64_t *z) {
    for (int l = 0; l < RD; l++) z[l] = (int64_t)((bits >> l) & 1);
}

/* fold_witness: out[j * 64 + k] = (z1_j + z2_j * r)[k] mod q (canonical), z1, z2 binary */
void or_ring_fold_witness(const uint64_t *z1, const uint64_t *z2, const int8_t *r, uint64_t N, uint64_t *out) {
    int64_t rr[RD], a[RD], acc[RD];
    for (int l = 0; l < RD; l++) rr[l] = r[l];
    for (uint64_t j = 0; j < N; j++) {
        bits_to_int(z1[j], acc);
        bits_to_int(z2[j], a);
        ring_mul_int_acc(a, rr, acc);
        for (int k = 0; k < RD; k++) out[j * RD + k] = lift_int(acc[k]);
    }
}

/* fold_commitment: out = c1 + r * c2 in R^KAPPA (r's small coefficients lifted into F_q) */
/* The commitment folds take a challenge coefficient only when it lies in [-2, 2]; others add
 * nothing (cuda_almost_goldilocks/ajtai.cuh:808-826 and 1120-1140 branch on +-1 and +-2 only). */
static uint64_t commit_coeff(int8_t r) {
    return (r >= -2 && r <= 2) ? lift_int(r) : 0;
}

void or_ring_fold_commitment(const uint64_t *c1, const uint64_t *c2, const int8_t *r, uint64_t *out) {
    uint64_t rf[RD];
    for (int l = 0; l < RD; l++) rf[l] = commit_coeff(r[l]);
    for (int i = 0; i < RKAPPA; i++) {
        for (int k = 0; k < RD; k++) out[i * RD + k] = cq(c1[i * RD + k]);
        ring_mul_acc(rf, c2 + (uint64_t)i * RD, out + (uint64_t)i * RD);
    }
}

/* multifold_mixed_witness: out[j * 64 + k] = sum_b (bin_b,j * r_b)[k] + sum_t ((pos_t,j - neg_t,j) * r_(K+t))[k]
 * as 16-bit integers; multifold_witness is the case T = 0 */
void or_ring_multifold_mixed_witness(const uint64_t *bin, const uint64_t *pos, const uint64_t *neg, const int8_t *r_all,
                                     int K, int T, uint64_t N, int16_t *out) {
    int64_t z[RD], rr[RD], acc[RD];
    for (uint64_t j = 0; j < N; j++) {
        memset(acc, 0, sizeof(acc));
        for (int m = 0; m < K + T; m++) {
            for (int l = 0; l < RD; l++) rr[l] = r_all[m * RD + l];
            if (m < K) {
                bits_to_int(bin[(uint64_t)m * N + j], z);
            } else {
                uint64_t p = pos[(uint64_t)(m - K) * N + j], n = neg[(uint64_t)(m - K) * N + j];
                for (int l = 0; l < RD; l++) z[l] = (int64_t)((p >> l) & 1) - (int64_t)((n >> l) & 1);
            }
            ring_mul_int_acc(z, rr, acc);
        }
        for (int k = 0; k < RD; k++) out[j * RD + k] = (int16_t)acc[k];
    }
}

/* multifold_commitment: out = sum_m r_m * c_m in R^KAPPA */
void or_ring_multifold_commitment(const uint64_t *c_all, const int8_t *r_all, int M, uint64_t *out) {
    uint64_t rf[RD];
    memset(out, 0, (size_t)RKAPPA * RD * sizeof(uint64_t));
    for (int m = 0; m < M; m++) {
        for (int l = 0; l < RD; l++) rf[l] = commit_coeff(r_all[m * RD + l]);
        for (int i = 0; i < RKAPPA; i++)
            ring_mul_acc(rf, c_all + ((uint64_t)m * RKAPPA + i) * RD, out + (uint64_t)i * RD);
    }
}

/* split_witness (SPLIT_K = 13 chunks): v = sum_i 2^i (pos_i - neg_i) with the bits of |v| on the
 * side of v's sign; pos[i * N + j] bit k, neg likewise */
void or_ring_split_witness(const int16_t *z_wide, uint64_t N, uint64_t *pos, uint64_t *neg) {
    memset(pos, 0, (size_t)13 * N * sizeof(uint64_t));
    memset(neg, 0, (size_t)13 * N * sizeof(uint64_t));
    for (uint64_t j = 0; j < N; j++)
        for (int k = 0; k < RD; k++) {
            int v = z_wide[j * RD + k], a = v < 0 ? -v : v;
            for (int i = 0; i < 13; i++)
                if ((a >> i) & 1) (v < 0 ? neg : pos)[(uint64_t)i * N + j] |= 1ull << k;
        }
}

/* commit_ternary: out[b][i][r] for the 13 ternary chunks z_b = pos_b - neg_b (wide commits) */
void or_ring_commit_ternary(const uint32_t *key8, const uint64_t *pos, const uint64_t *neg, uint64_t N, uint64_t *out) {
    uint64_t *zw = (uint64_t *)malloc((N ? N : 1) * RD * sizeof(uint64_t));
    for (int b = 0; b < 13; b++) {
        for (uint64_t j = 0; j < N; j++) {
            uint64_t p = pos[(uint64_t)b * N + j], n = neg[(uint64_t)b * N + j];
            for (int l = 0; l < RD; l++)
                zw[j * RD + l] = lift_int((int64_t)((p >> l) & 1) - (int64_t)((n >> l) & 1));
        }
        or_ring_commit_wide(key8, zw, N, 0, out + (uint64_t)b * RKAPPA * RD);
    }
    free(zw);
}

/* materialized matrix M[j][row][k] (ajtai_premat.cuh:41-70) */
void or_ring_materialize(const uint32_t *key8, uint64_t N, uint64_t *m) {
    for (uint64_t j = 0; j < N; j++)
        for (int i = 0; i < RKAPPA; i++) or_ring_matrix_entry(key8, (uint32_t)i, j, m + (j * RKAPPA + i) * RD);
}

/* selective add over shared eq (almost_sumcheck_prover.cuh:924-966), totals per plane:
 * out[p] = sum of the Ext2 eq[i] over set bits i of plane p */
void or_q_selective_add_planes(const uint64_t *eq2, const uint64_t *planes, uint64_t total, int n_planes,
                               uint64_t packed, uint64_t *out2) {
    for (int p = 0; p < n_planes; p++) {
        uint64_t s[2] = {0, 0};
        for (uint64_t i = 0; i < total; i++)
            if ((planes[(uint64_t)p * packed + (i >> 6)] >> (i & 63)) & 1) e_add(s, s, eq2 + 2 * i);
        out2[2 * p] = s[0];
        out2[2 * p + 1] = s[1];
    }
}

/* fused permute + partial evaluation (cuda/fused_permute_peval.cuh:31-120) over the prime
 * `mod` (p or q): out[j] = sum_b evals[lo[x & (2^half-1)] | hi[x >> half]] * eq[b], x = j 2^m + b */
void or_fused_permute_peval(uint64_t mod, const uint64_t *evals, const uint64_t *eq, const uint32_t *lo,
                            const uint32_t *hi, int n, int m, int half, long out_size, uint64_t *out) {
    uint64_t lo_mask = (1ULL << half) - 1;
    for (long j = 0; j < out_size; j++) {
        u128 a0 = 0, a1 = 0;
        for (uint64_t b = 0; b < (1ULL << m); b++) {
            uint64_t x = ((uint64_t)j << m) + b;
            uint64_t old = (uint64_t)(lo[x & lo_mask] | hi[x >> half]);
            uint64_t v = evals[old] % mod;
            a0 = (a0 + (u128)v * (eq[2 * b] % mod)) % mod;
            a1 = (a1 + (u128)v * (eq[2 * b + 1] % mod)) % mod;
        }
        out[2 * j] = (uint64_t)a0;
        out[2 * j + 1] = (uint64_t)a1;
    }
    (void)n;
}

/* ---- zk-torch-4 transcript: Monolith-w12 over q and the duplex sponge ----
 * monolith.rs:33-213 (permutation; the round constants are the caller's, derived from
 * SHA-256 in pyoracle.py as monolith.rs:67-110 does) and transcript.rs:21-240 (sponge). */
static const uint64_t MQ_ROW0[12] = {7, 23, 8, 26, 13, 10, 9, 7, 6, 22, 21, 8};

static uint8_t rotl8(uint8_t b, int k) { return (uint8_t)((b << k) | (b >> (8 - k))); }

/* bar_64 byte by byte: b ^ (rotl1(~b) & rotl2(b) & rotl3(b)), then rotl1 (monolith.rs:146-153) */
static uint64_t mq_bar(uint64_t v) {
    uint64_t r = 0;
    for (int k = 0; k < 8; k++) {
        uint8_t b = (uint8_t)(v >> (8 * k));
        uint8_t t = b ^ (rotl8((uint8_t)~b, 1) & rotl8(b, 2) & rotl8(b, 3));
        r |= (uint64_t)rotl8(t, 1) << (8 * k);
    }
    return r;
}

/* y[row] = sum_col M[row][col] x[col] + rc[row], M[row][col] = ROW0[(col - row) mod 12] */
static void mq_concrete(uint64_t *x, const uint64_t *rc) {
    uint64_t y[12];
    for (int row = 0; row < 12; row++) {
        u128 acc = rc ? rc[row] : 0;
        for (int col = 0; col < 12; col++) acc += (u128)cq(x[col]) * MQ_ROW0[(col - row + 12) % 12];
        y[row] = (uint64_t)(acc % QP);
    }
    memcpy(x, y, sizeof(y));
}

void or_monoq_permute(uint64_t *x, const uint64_t *rc60) {
    mq_concrete(x, NULL);
    for (int r = 1; r <= 6; r++) {
        for (int i = 0; i < 4; i++) x[i] = mq_bar(x[i]);
        for (int i = 11; i >= 1; i--) x[i] = or_q_add(x[i], or_q_mul(x[i - 1], x[i - 1]));
        mq_concrete(x, r <= 5 ? rc60 + 12 * (r - 1) : NULL);
    }
}

void or_qtr_init(or_qtr *t, const uint8_t *label, long len, const uint64_t *rc60) {
    memset(t, 0, sizeof(*t));
    memcpy(t->rc, rc60, sizeof(t->rc));
    for (long i = 0; i < len; i++) or_qtr_absorb(t, label[i]);
}

static void qtr_duplex(or_qtr *t) {
    for (int i = 0; i < 8; i++) t->st[i] = i < t->nin ? t->in[i] : 0;
    t->nin = 0;
    or_monoq_permute(t->st, t->rc);
    for (int i = 0; i < 8; i++) t->out[7 - i] = t->st[i];
    t->nout = 8;
}

void or_qtr_absorb(or_qtr *t, uint64_t v) {
    t->nout = 0;
    t->in[t->nin++] = cq(v);
    if (t->nin == 8) qtr_duplex(t);
}

static void qtr_label(or_qtr *t, const uint8_t *label, long len) {
    if (len == 0) return;
    or_qtr_absorb(t, (uint64_t)len);
    for (long i = 0; i < len; i++) or_qtr_absorb(t, label[i]);
}

void or_qtr_append(or_qtr *t, const uint8_t *label, long len, const uint64_t *v, long n) {
    qtr_label(t, label, len);
    for (long i = 0; i < n; i++) or_qtr_absorb(t, v[i]);
}

void or_qtr_challenge(or_qtr *t, const uint8_t *label, long len, uint64_t *out, long n) {
    qtr_label(t, label, len);
    for (long i = 0; i < n; i++) {
        if (t->nout == 0) qtr_duplex(t);
        out[i] = t->out[--t->nout];
    }
}

/* CpuLinearSumcheckProverExt2::new + prove (zk-torch-4/src/sumcheck/cpu_ext2_prover.rs:23-100):
 * polys [np][2^nv] base words (ext = 0, lifted to Ext2 as GpuSumcheckStateExt2::new_from_base
 * does) or Ext2 pairs; outputs msgs [nv][np + 1][2], chs [nv][2], finals [np][2], fe [2]. */
void or_q_sumcheck_prove(or_qtr *t, int ext, const uint64_t *polys, int np, int nv, uint64_t *msgs, uint64_t *chs,
                         uint64_t *finals, uint64_t *fe) {
    const uint64_t n = 1ULL << nv;
    uint64_t *w = (uint64_t *)malloc(sizeof(uint64_t) * 2 * n * (uint64_t)np);
    for (int i = 0; i < np; i++)
        for (uint64_t j = 0; j < n; j++) {
            w[2 * (i * n + j)] = cq(ext ? polys[2 * (i * n + j)] : polys[i * n + j]);
            w[2 * (i * n + j) + 1] = ext ? cq(polys[2 * (i * n + j) + 1]) : 0;
        }
    uint64_t v = (uint64_t)nv;
    or_qtr_append(t, (const uint8_t *)"num_var", 7, &v, 1);
    v = (uint64_t)np;
    or_qtr_append(t, (const uint8_t *)"num_poly", 8, &v, 1);
    uint64_t size = n;
    for (int r = 0; r < nv; r++) {
        uint64_t *m = msgs + (uint64_t)r * 2 * (np + 1);
        /* arrays of `size` elements at stride n: view them as polynomials of original size n */
        or_q_sumcheck_msg(1, w, np, n, size / 2, m);
        for (int c = 0; c <= np; c++) or_qtr_append(t, (const uint8_t *)"round_message", 13, m + 2 * c, 2);
        or_qtr_challenge(t, (const uint8_t *)"challenge", 9, chs + 2 * r, 2);
        or_q_sumcheck_fold(1, w, w, chs + 2 * r, np, n, size / 2);
        size /= 2;
    }
    uint64_t acc[2] = {1, 0};
    for (int i = 0; i < np; i++) {
        finals[2 * i] = w[2 * i * n];
        finals[2 * i + 1] = w[2 * i * n + 1];
        or_q2_mul(acc, finals + 2 * i, acc);
    }
    fe[0] = acc[0];
    fe[1] = acc[1];
    free(w);
}
