/*
 * oracle/core.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Field exports, the zk-torch-3 Fiat-Shamir transcript, eq tables, partial
 * evaluation and the CPU linear sumcheck prover/verifier over Ext2.
 *
 *   transcript        zk-torch-3/src/transcript.rs:167-330
 *   eq tables         zk-torch-3/src/poly/mod.rs:106-163
 *   sumcheck prover   zk-torch-3/src/sumcheck/cpu_ext2_prover.rs:21-148
 *   sumcheck verifier zk-torch-3/src/sumcheck/verifier.rs:17-65,116-140
 */
#include <stdlib.h>
#include <string.h>
#include "gl.h"
#include "oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the parallel loops (OpenMP); every result is independent of the count:
 * parallel sums add exact residues mod p, whose order does not matter. */
void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}
int or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

u64 or_gl_add(u64 a, u64 b) { return gl_add(a, b); }
u64 or_gl_sub(u64 a, u64 b) { return gl_sub(a, b); }
u64 or_gl_mul(u64 a, u64 b) { return gl_mul(a, b); }
u64 or_gl_inv(u64 a) { return gl_inv(a); }
void or_ext2_mul(const u64 *a, const u64 *b, u64 *out) {
    ext2 r = e2_mul(e2(a[0], a[1]), e2(b[0], b[1]));
    out[0] = r.c0; out[1] = r.c1;
}
void or_ext2_inv(const u64 *a, u64 *out) {
    ext2 r = e2_inv(e2(a[0], a[1]));
    out[0] = r.c0; out[1] = r.c1;
}

/* ------------------------------------------------------------------ */
/* Duplex transcript: width 8, rate 4, overwrite mode                  */
/* ------------------------------------------------------------------ */

static void tr_duplex(or_transcript *t) {
    for (int i = 0; i < t->in_count; i++) t->state[i] = t->in_buf[i];
    for (int i = (int)t->in_count; i < 4; i++) t->state[i] = 0;
    t->in_count = 0;
    or_poseidon2_permute(t->state, 0);
    for (int i = 0; i < 4; i++) t->out_buf[3 - i] = t->state[i];
    t->out_count = 4;
}

static void tr_absorb(or_transcript *t, u64 v) {
    t->out_count = 0;
    t->in_buf[t->in_count++] = gl_canon(v);
    if (t->in_count == 4) tr_duplex(t);
}

static u64 tr_squeeze(or_transcript *t) {
    if (t->out_count == 0) tr_duplex(t);
    t->out_count--;
    return t->out_buf[t->out_count];
}

static void tr_label(or_transcript *t, const uint8_t *label, long len) {
    if (len <= 0) return;
    tr_absorb(t, (u64)len);
    for (long i = 0; i < len; i++) tr_absorb(t, label[i]);
}

void or_tr_new(or_transcript *t, const uint8_t *label, long len) {
    memset(t, 0, sizeof(*t));
    for (long i = 0; i < len; i++) tr_absorb(t, label[i]);
}

void or_tr_append_u64(or_transcript *t, const uint8_t *label, long len, u64 v) {
    tr_label(t, label, len);
    tr_absorb(t, v);
}

void or_tr_append_scalars(or_transcript *t, const uint8_t *label, long len, const u64 *v, long n) {
    tr_label(t, label, len);
    for (long i = 0; i < n; i++) tr_absorb(t, v[i]);
}

void or_tr_append_ext2(or_transcript *t, const uint8_t *label, long len, const u64 *v2) {
    tr_label(t, label, len);
    tr_absorb(t, v2[0]);
    tr_absorb(t, v2[1]);
}

u64 or_tr_challenge_scalar(or_transcript *t, const uint8_t *label, long len) {
    tr_label(t, label, len);
    return tr_squeeze(t);
}

void or_tr_challenge_ext2(or_transcript *t, const uint8_t *label, long len, u64 *out2) {
    tr_label(t, label, len);
    out2[0] = tr_squeeze(t);
    out2[1] = tr_squeeze(t);
}

/* ------------------------------------------------------------------ */
/* eq tables: eq[x] = prod_i (bit i of x ? r_i : 1 - r_i)              */
/* ------------------------------------------------------------------ */

void or_eq_table(const u64 *r, int nv, u64 *out) {
    long size = 1L << nv;
    memset(out, 0, size * sizeof(u64));
    out[0] = 1;
    for (int i = 0; i < nv; i++) {
        long half = 1L << i;
        u64 ri = gl_canon(r[i]), omr = gl_sub(1, ri);
#pragma omp parallel for schedule(static) if (half > 8192)
        for (long j = half - 1; j >= 0; j--) {
            out[j | half] = gl_mul(out[j], ri);
            out[j] = gl_mul(out[j], omr);
        }
    }
}

void or_eq_table_ext2(const u64 *r2, int nv, u64 *out2) {
    long size = 1L << nv;
    ext2 *o = (ext2 *)out2;
    for (long j = 0; j < size; j++) o[j] = e2_zero();
    o[0] = e2_one();
    for (int i = 0; i < nv; i++) {
        long half = 1L << i;
        ext2 ri = e2(r2[2 * i], r2[2 * i + 1]), omr = e2_sub(e2_one(), ri);
#pragma omp parallel for schedule(static) if (half > 8192)
        for (long j = half - 1; j >= 0; j--) {
            o[j | half] = e2_mul(o[j], ri);
            o[j] = e2_mul(o[j], omr);
        }
    }
}

/* fold the first m variables (LSB first): out[j] = in[2j] + r (in[2j+1] - in[2j]) */
void or_partial_eval_ext2(const u64 *in2, int nv, const u64 *r2, int m, u64 *out2) {
    long n = 1L << nv;
    ext2 *buf = (ext2 *)malloc(n * sizeof(ext2));
    for (long j = 0; j < n; j++) buf[j] = e2(in2[2 * j], in2[2 * j + 1]);
    for (int k = 0; k < m; k++) {
        ext2 r = e2(r2[2 * k], r2[2 * k + 1]);
        n /= 2;
        /* in place: chunks of j only read entries 2j, 2j + 1 >= j, so a sequential pass is
         * required; a second buffer makes the fold parallel */
        ext2 *nb = (ext2 *)malloc((n > 0 ? n : 1) * sizeof(ext2));
#pragma omp parallel for schedule(static) if (n > 8192)
        for (long j = 0; j < n; j++) nb[j] = e2_add(buf[2 * j], e2_mul(r, e2_sub(buf[2 * j + 1], buf[2 * j])));
        memcpy(buf, nb, (n > 0 ? n : 1) * sizeof(ext2));
        free(nb);
    }
    memcpy(out2, buf, n * sizeof(ext2));
    free(buf);
}

/* ------------------------------------------------------------------ */
/* Linear sumcheck over Ext2                                           */
/* ------------------------------------------------------------------ */

static const uint8_t L_NUM_VAR[] = "num_var";
static const uint8_t L_NUM_POLY[] = "num_poly";
static const uint8_t L_ROUND_MSG[] = "round_message";
static const uint8_t L_CHALLENGE[] = "challenge";

void or_sumcheck_prove_ext2(or_transcript *t, u64 *polys, int np, int nv, u64 *round_msgs,
                            u64 *challenges, u64 *final_evals, u64 *final_eval) {
    long size = 1L << nv;
    ext2 *p = (ext2 *)polys;
    for (long k = 0; k < (long)np * size; k++) p[k] = e2(p[k].c0, p[k].c1);
    or_tr_append_u64(t, L_NUM_VAR, 7, (u64)nv);
    or_tr_append_u64(t, L_NUM_POLY, 8, (u64)np);
    long cur = size;
    for (int round = 0; round < nv; round++) {
        long half = cur / 2;
        ext2 *msg = (ext2 *)(round_msgs + (long)round * (np + 1) * 2);
        for (int e = 0; e <= np; e++) {
            ext2 c = e2_base((u64)e), sum = e2_zero();
#pragma omp parallel if (half > 8192)
            {
                ext2 part = e2_zero();
#pragma omp for schedule(static)
                for (long j = 0; j < half; j++) {
                    ext2 prod = e2_one();
                    for (int i = 0; i < np; i++) {
                        ext2 a = p[i * size + 2 * j], b = p[i * size + 2 * j + 1];
                        prod = e2_mul(prod, e2_add(a, e2_mul(c, e2_sub(b, a))));
                    }
                    part = e2_add(part, prod);
                }
#pragma omp critical
                sum = e2_add(sum, part);
            }
            msg[e] = sum;
        }
        for (int e = 0; e <= np; e++) or_tr_append_ext2(t, L_ROUND_MSG, 13, (u64 *)&msg[e]);
        u64 ch[2];
        or_tr_challenge_ext2(t, L_CHALLENGE, 9, ch);
        challenges[2 * round] = ch[0];
        challenges[2 * round + 1] = ch[1];
        ext2 r = e2(ch[0], ch[1]);
        ext2 *tmp = (ext2 *)malloc((half > 0 ? half : 1) * sizeof(ext2));
        for (int i = 0; i < np; i++) {
            ext2 *q = p + (long)i * size;
#pragma omp parallel for schedule(static) if (half > 8192)
            for (long j = 0; j < half; j++) tmp[j] = e2_add(q[2 * j], e2_mul(r, e2_sub(q[2 * j + 1], q[2 * j])));
            memcpy(q, tmp, (half > 0 ? half : 1) * sizeof(ext2));
        }
        free(tmp);
        cur = half;
    }
    ext2 fe = e2_one();
    for (int i = 0; i < np; i++) {
        ext2 v = p[(long)i * size];
        final_evals[2 * i] = v.c0;
        final_evals[2 * i + 1] = v.c1;
        fe = e2_mul(fe, v);
    }
    final_eval[0] = fe.c0;
    final_eval[1] = fe.c1;
}

/* Lagrange interpolation through (0..d-1, evals) evaluated at x (verifier.rs:116-140) */
static ext2 interp_eval(const ext2 *ev, int d, ext2 x) {
    ext2 res = e2_zero();
    for (int i = 0; i < d; i++) {
        ext2 basis = e2_one();
        for (int j = 0; j < d; j++) {
            if (j == i) continue;
            ext2 num = e2_sub(x, e2_base((u64)j));
            ext2 den = e2_sub(e2_base((u64)i), e2_base((u64)j));
            basis = e2_mul(basis, e2_mul(num, e2_inv(den)));
        }
        res = e2_add(res, e2_mul(ev[i], basis));
    }
    return res;
}

int or_sumcheck_verify_ext2(or_transcript *t, const u64 *claim2, int np, int nv,
                            const u64 *round_msgs, const u64 *final_eval) {
    or_tr_append_u64(t, L_NUM_VAR, 7, (u64)nv);
    or_tr_append_u64(t, L_NUM_POLY, 8, (u64)np);
    ext2 cur = e2(claim2[0], claim2[1]);
    ext2 *ev = (ext2 *)malloc((np + 1) * sizeof(ext2));
    int ok = 1;
    for (int round = 0; round < nv && ok; round++) {
        const u64 *m = round_msgs + (long)round * (np + 1) * 2;
        for (int e = 0; e <= np; e++) ev[e] = e2(m[2 * e], m[2 * e + 1]);
        if (!e2_eq(e2_add(ev[0], ev[1]), cur)) { ok = 0; break; }
        for (int e = 0; e <= np; e++) or_tr_append_ext2(t, L_ROUND_MSG, 13, m + 2 * e);
        u64 ch[2];
        or_tr_challenge_ext2(t, L_CHALLENGE, 9, ch);
        cur = interp_eval(ev, np + 1, e2(ch[0], ch[1]));
    }
    if (ok) ok = e2_eq(cur, e2(final_eval[0], final_eval[1]));
    free(ev);
    return ok;
}
