/*
 * oracle/basefold.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Basefold PCS over Goldilocks with Ext2 opening, restated literally:
 *   BasefoldTable::generate      goldilocks-cuda-rs/src/basefold.rs:281-321
 *   commit: BHC interpolation    goldilocks-cuda-rs/cuda/wrapper.cu:1242-1274, cuda/basefold.cuh:103-147
 *           encode               goldilocks-cuda-rs/cuda/wrapper.cu:1277-1312, cuda/basefold.cuh:157-205
 *           Merkle tree          goldilocks-cuda-rs/src/merkle.rs:34-59
 *   open_ext2 (CPU)              zk-torch-3/src/commit/cpu_basefold.rs:235-425
 *                                (query sampling goldilocks-cuda-rs/src/basefold.rs:2310-2402)
 *   verify_ext2                  goldilocks-cuda-rs/src/basefold.rs:1759-1925
 *
 * Flat proof layout (u64, canonical), shared with the HIP library:
 *   eval[2] | oracles[nv][3][2] | folded_roots[nv-1][4] | final_cw[2^lr][2] |
 *   nq x { index | values[nv][2][2] | initial path[nv+lr-1][4] |
 *          folded path i=0..nv-2: [nv+lr-2-i][4] }
 */
#include <stdlib.h>
#include <string.h>
#include "gl.h"
#include "oracle.h"

long or_basefold_table_len(int nv, int lr, int nrounds) {
    long total = 0;
    for (int i = 0; i < nrounds; i++) total += 1L << (nv + lr - i - 1);
    return total;
}

static long table_level_offset(int table_nv, int lr, int level) {
    long off = 0;
    for (int i = 0; i < level; i++) off += 1L << (table_nv + lr - i - 1);
    return off;
}

/* xorshift(13,7,17); each draw is the raw state reduced mod p */
void or_basefold_table_generate(int nv, int lr, int nrounds, u64 seed, u64 *entries) {
    u64 st = seed;
#define DRAW() (st ^= st << 13, st ^= st >> 7, st ^= st << 17, st % GL_P)
    long off = 0;
    for (int i = 0; i < nrounds; i++) {
        long lsize = 1L << (nv + lr - i - 1);
        u64 *e = entries + 2 * off;
        for (long k = 0; k < lsize; k++) {
            u64 x0 = DRAW();
            u64 x1 = DRAW();
            while (x1 == x0) x1 = DRAW();
            e[2 * k] = x0;
            e[2 * k + 1] = gl_sub(x1, x0); /* inverted below */
        }
        /* batch inversion of the level's differences (same values as Fermat per entry) */
        u64 *pref = (u64 *)malloc(lsize * sizeof(u64));
        u64 acc = 1;
        for (long k = 0; k < lsize; k++) { pref[k] = acc; acc = gl_mul(acc, e[2 * k + 1]); }
        u64 inv = gl_inv(acc);
        for (long k = lsize - 1; k >= 0; k--) {
            u64 d = e[2 * k + 1];
            e[2 * k + 1] = gl_mul(inv, pref[k]);
            inv = gl_mul(inv, d);
        }
        free(pref);
        off += lsize;
    }
#undef DRAW
}

static long bitrev(long x, int bits) {
    long y = 0;
    for (int i = 0; i < bits; i++) { y = (y << 1) | (x & 1); x >>= 1; }
    return y;
}

static void bitrev_permute(u64 *v, int bits, int width) {
    long n = 1L << bits;
    for (long i = 0; i < n; i++) {
        long r = bitrev(i, bits);
        if (i < r)
            for (int w = 0; w < width; w++) {
                u64 t = v[i * width + w];
                v[i * width + w] = v[r * width + w];
                v[r * width + w] = t;
            }
    }
}

void or_basefold_commit(const u64 *evals, int nv, int lr, int hash, u64 *codeword, u64 *bh_evals, u64 *tree) {
    long n = 1L << nv, cwn = 1L << (nv + lr);
    u64 *coeffs = (u64 *)malloc(n * sizeof(u64));
    /* BHC interpolation: pairs (a, b - a), then coeffs[upper] -= coeffs[lower] per layer */
    for (long i = 0; i < n / 2; i++) {
        coeffs[2 * i] = gl_canon(evals[2 * i]);
        coeffs[2 * i + 1] = gl_sub(evals[2 * i + 1], evals[2 * i]);
    }
    for (int k = 1; k < nv; k++) {
        long hc = 1L << k;
#pragma omp parallel for schedule(static) if (n > 16384)
        for (long idx = 0; idx < n / 2; idx++) {
            long base = (idx / hc) * 2 * hc, pos = idx % hc;
            coeffs[base + hc + pos] = gl_sub(coeffs[base + hc + pos], coeffs[base + pos]);
        }
    }
    for (long i = 0; i < n; i++) bh_evals[i] = gl_canon(evals[i]);
    bitrev_permute(bh_evals, nv, 1);
    /* encode: repetition, nv butterfly layers (lo+hi, lo-hi), bit reversal */
    for (long i = 0; i < cwn; i++) codeword[i] = coeffs[i >> lr];
    for (int i = 0; i < nv; i++) {
        long hc = 1L << (i + lr);
#pragma omp parallel for schedule(static) if (cwn > 16384)
        for (long idx = 0; idx < cwn / 2; idx++) {
            long lo = (idx / hc) * 2 * hc + idx % hc, hi = lo + hc;
            u64 a = codeword[lo], b = codeword[hi];
            codeword[lo] = gl_add(a, b);
            codeword[hi] = gl_sub(a, b);
        }
    }
    bitrev_permute(codeword, nv + lr, 1);
    free(coeffs);
    if (tree) {
        if (hash == OR_HASH_MONOLITH || hash == OR_HASH_POSEIDON2) or_merkle_tree_gl(codeword, cwn / 2, hash, tree);
    }
}

long or_basefold_proof_len(int nv, int lr, int nq) {
    long q = 1 + 4L * nv + 4L * (nv + lr - 1);
    for (int i = 0; i <= nv - 2; i++) q += 4L * (nv + lr - 2 - i);
    return 2 + 6L * nv + 4L * (nv - 1) + 2L * (1L << lr) + (long)nq * q;
}

/* degree-2 oracle coefficients from P(0), P(1), P(2) (cpu_basefold.rs:157-221) */
static void oracle_coeffs(ext2 p0, ext2 p1, ext2 p2, ext2 *c) {
    ext2 two = e2_base(2), inv2 = e2_base(9223372034707292161ULL);
    c[0] = p0;
    c[2] = e2_mul(e2_add(e2_sub(p0, e2_mul(two, p1)), p2), inv2);
    c[1] = e2_sub(e2_sub(p1, p0), c[2]);
}

/* a[j] = a[2j] + ch (a[2j+1] - a[2j]) for j < half, through a temporary so that the
 * loop can run in parallel */
static void fold_pairs_par(ext2 *a, long half, ext2 ch) {
    ext2 *t = (ext2 *)malloc((half > 0 ? half : 1) * sizeof(ext2));
#pragma omp parallel for schedule(static) if (half > 16384)
    for (long j = 0; j < half; j++) t[j] = e2_add(a[2 * j], e2_mul(ch, e2_sub(a[2 * j + 1], a[2 * j])));
    memcpy(a, t, (half > 0 ? half : 1) * sizeof(ext2));
    free(t);
}

static void compute_oracle(const ext2 *bh, const ext2 *eq, long size, ext2 *c) {
    ext2 two = e2_base(2), p0 = e2_zero(), p1 = e2_zero(), p2 = e2_zero();
#pragma omp parallel if (size > 16384)
    {
        ext2 q0 = e2_zero(), q1 = e2_zero(), q2 = e2_zero();
#pragma omp for schedule(static)
        for (long j = 0; j < size / 2; j++) {
            q0 = e2_add(q0, e2_mul(eq[2 * j], bh[2 * j]));
            q1 = e2_add(q1, e2_mul(eq[2 * j + 1], bh[2 * j + 1]));
            ext2 ei = e2_sub(e2_mul(two, eq[2 * j + 1]), eq[2 * j]);
            ext2 bi = e2_sub(e2_mul(two, bh[2 * j + 1]), bh[2 * j]);
            q2 = e2_add(q2, e2_mul(ei, bi));
        }
#pragma omp critical
        {
            p0 = e2_add(p0, q0);
            p1 = e2_add(p1, q1);
            p2 = e2_add(p2, q2);
        }
    }
    oracle_coeffs(p0, p1, p2, c);
}

static ext2 fold_one(ext2 v0, ext2 v1, u64 x0, u64 w, ext2 ch) {
    ext2 diffw = e2_mul(e2_sub(v1, v0), e2_base(w));
    ext2 cx = e2_sub(ch, e2_base(x0));
    return e2_add(v0, e2_mul(cx, diffw));
}

static void observe_hash(or_transcript *t, const u64 *h4) {
    for (int i = 0; i < 4; i++) or_tr_append_scalars(t, NULL, 0, h4 + i, 1);
}
static void observe_ext2(or_transcript *t, ext2 v) {
    u64 a[2] = {v.c0, v.c1};
    or_tr_append_ext2(t, NULL, 0, a);
}
static ext2 sample_ext2(or_transcript *t) {
    u64 a[2];
    or_tr_challenge_ext2(t, NULL, 0, a);
    return e2(a[0], a[1]);
}

static void tree_path(const u64 *tree, long num_leaves, long leaf, u64 *out) {
    long off = 0, n = num_leaves, idx = leaf;
    int l = 0;
    while (n > 1) {
        memcpy(out + 4 * l, tree + (off + (idx ^ 1)) * 4, 4 * sizeof(u64));
        off += n;
        n /= 2;
        idx /= 2;
        l++;
    }
}

/* tree0: the commitment's Merkle tree over the codeword (the reference's CPU opening takes it
 * from the commitment's host cache, cpu_basefold.rs:235-260); NULL rebuilds it here. */
void or_basefold_open_ext2_tree(const u64 *codeword, const u64 *bh_evals, const u64 *root4, const u64 *tree0,
                                int nv, int lr, int hash, const u64 *point2, const u64 *table, int table_nv,
                                or_transcript *t, int nq, u64 *proof) {
    long n = 1L << nv, cwn = 1L << (nv + lr);
    u64 *pc = proof;
    /* transcript preamble */
    observe_hash(t, root4);
    for (int i = 0; i < nv; i++) observe_ext2(t, e2(point2[2 * i], point2[2 * i + 1]));
    /* eq table and bh in Type2 order */
    ext2 *eq = (ext2 *)malloc(n * sizeof(ext2));
    or_eq_table_ext2(point2, nv, (u64 *)eq);
    ext2 *bh = (ext2 *)malloc(n * sizeof(ext2));
#pragma omp parallel for schedule(static) if (n > 16384)
    for (long i = 0; i < n; i++) bh[bitrev(i, nv)] = e2_base(bh_evals[i]);
    ext2 eval = e2_zero();
#pragma omp parallel if (n > 16384)
    {
        ext2 part = e2_zero();
#pragma omp for schedule(static)
        for (long i = 0; i < n; i++) part = e2_add(part, e2_mul(bh[i], eq[i]));
#pragma omp critical
        eval = e2_add(eval, part);
    }
    *pc++ = eval.c0;
    *pc++ = eval.c1;
    u64 *oracles = pc;
    pc += 6L * nv;
    u64 *roots = pc;
    pc += 4L * (nv - 1);
    u64 *final_cw = pc;
    pc += 2L * (1L << lr);

    ext2 c[3];
    compute_oracle(bh, eq, n, c);
    for (int k = 0; k < 3; k++) { oracles[2 * k] = c[k].c0; oracles[2 * k + 1] = c[k].c1; observe_ext2(t, c[k]); }
    ext2 ch = sample_ext2(t);
    long sc = n / 2;
    fold_pairs_par(bh, sc, ch);
    fold_pairs_par(eq, sc, ch);
    compute_oracle(bh, eq, sc, c);
    for (int k = 0; k < 3; k++) { oracles[6 + 2 * k] = c[k].c0; oracles[6 + 2 * k + 1] = c[k].c1; observe_ext2(t, c[k]); }

    /* folded codewords (kept for queries) and their trees */
    ext2 **fcw = (ext2 **)calloc(nv, sizeof(ext2 *));
    u64 **ftree = (u64 **)calloc(nv, sizeof(u64 *));
    long fsize = cwn / 2;
    const u64 *lvl = table + 2 * table_level_offset(table_nv, lr, 0);
    fcw[0] = (ext2 *)malloc(fsize * sizeof(ext2));
#pragma omp parallel for schedule(static) if (fsize > 16384)
    for (long i = 0; i < fsize; i++)
        fcw[0][i] = fold_one(e2_base(codeword[2 * i]), e2_base(codeword[2 * i + 1]), lvl[2 * i], lvl[2 * i + 1], ch);
    ftree[0] = (u64 *)malloc((2 * (fsize / 2) - 1) * 4 * sizeof(u64));
    or_merkle_tree_ext2((u64 *)fcw[0], fsize / 2, hash, ftree[0]);
    memcpy(roots, ftree[0] + or_merkle_root_offset(fsize / 2), 4 * sizeof(u64));
    observe_hash(t, roots);

    for (int round = 1; round < nv - 1; round++) {
        ch = sample_ext2(t);
        long half = sc / 2;
        fold_pairs_par(eq, half, ch);
        fold_pairs_par(bh, half, ch);
        sc = half;
        compute_oracle(bh, eq, sc, c);
        for (int k = 0; k < 3; k++) {
            oracles[6L * (round + 1) + 2 * k] = c[k].c0;
            oracles[6L * (round + 1) + 2 * k + 1] = c[k].c1;
            observe_ext2(t, c[k]);
        }
        lvl = table + 2 * table_level_offset(table_nv, lr, round);
        long nsize = fsize / 2;
        fcw[round] = (ext2 *)malloc(nsize * sizeof(ext2));
#pragma omp parallel for schedule(static) if (nsize > 16384)
        for (long i = 0; i < nsize; i++)
            fcw[round][i] = fold_one(fcw[round - 1][2 * i], fcw[round - 1][2 * i + 1], lvl[2 * i], lvl[2 * i + 1], ch);
        ftree[round] = (u64 *)malloc((2 * (nsize / 2) - 1) * 4 * sizeof(u64));
        or_merkle_tree_ext2((u64 *)fcw[round], nsize / 2, hash, ftree[round]);
        memcpy(roots + 4L * round, ftree[round] + or_merkle_root_offset(nsize / 2), 4 * sizeof(u64));
        observe_hash(t, roots + 4L * round);
        fsize = nsize;
    }
    /* last fold, no tree */
    ch = sample_ext2(t);
    lvl = table + 2 * table_level_offset(table_nv, lr, nv - 1);
    for (long i = 0; i < fsize / 2; i++) {
        ext2 v = fold_one(fcw[nv - 2][2 * i], fcw[nv - 2][2 * i + 1], lvl[2 * i], lvl[2 * i + 1], ch);
        final_cw[2 * i] = v.c0;
        final_cw[2 * i + 1] = v.c1;
    }
    /* queries */
    u64 *own = NULL;
    const u64 *itree = tree0;
    if (!itree) {
        own = (u64 *)malloc((2 * (cwn / 2) - 1) * 4 * sizeof(u64));
        or_merkle_tree_gl(codeword, cwn / 2, hash, own);
        itree = own;
    }
    long *qidx = (long *)malloc((nq > 0 ? nq : 1) * sizeof(long));
    for (int q = 0; q < nq; q++) qidx[q] = (long)(or_tr_challenge_scalar(t, NULL, 0) % (u64)(cwn / 2));
    for (int q = 0; q < nq; q++) {
        long leaf = qidx[q];
        u64 *qp = pc;
        *qp++ = (u64)leaf;
        u64 *vals = qp;
        qp += 4L * nv;
        vals[0] = gl_canon(codeword[2 * leaf]); vals[1] = 0;
        vals[2] = gl_canon(codeword[2 * leaf + 1]); vals[3] = 0;
        long idx = leaf / 2;
        for (int i = 0; i <= nv - 2; i++) {
            ext2 a = fcw[i][2 * idx], b = fcw[i][2 * idx + 1];
            u64 *v = vals + 4L * (i + 1);
            v[0] = a.c0; v[1] = a.c1; v[2] = b.c0; v[3] = b.c1;
            idx /= 2;
        }
        tree_path(itree, cwn / 2, leaf, qp);
        qp += 4L * (nv + lr - 1);
        idx = leaf / 2;
        long leaves = cwn / 4;
        for (int i = 0; i <= nv - 2; i++) {
            tree_path(ftree[i], leaves, idx, qp);
            qp += 4L * (nv + lr - 2 - i);
            idx /= 2;
            leaves /= 2;
        }
        pc = qp;
    }
    for (int i = 0; i < nv; i++) { free(fcw[i]); free(ftree[i]); }
    free(fcw); free(ftree); free(own); free(qidx); free(eq); free(bh);
}

void or_basefold_open_ext2(const u64 *codeword, const u64 *bh_evals, const u64 *root4, int nv, int lr,
                           int hash, const u64 *point2, const u64 *table, int table_nv, or_transcript *t,
                           int nq, u64 *proof) {
    or_basefold_open_ext2_tree(codeword, bh_evals, root4, NULL, nv, lr, hash, point2, table, table_nv, t, nq, proof);
}

static ext2 ld2(const u64 *p) { return e2(p[0], p[1]); }

int or_basefold_verify_ext2(const u64 *root4, const u64 *point2, int nv, int lr, int hash, const u64 *proof,
                            int nq, const u64 *table, int table_nv, or_transcript *t) {
    const u64 *pc = proof;
    ext2 eval = ld2(pc);
    pc += 2;
    const u64 *oracles = pc;
    pc += 6L * nv;
    const u64 *roots = pc;
    pc += 4L * (nv - 1);
    const u64 *final_cw = pc;
    pc += 2L * (1L << lr);

    observe_hash(t, root4);
    for (int i = 0; i < nv; i++) observe_ext2(t, ld2(point2 + 2 * i));
    ext2 c0 = ld2(oracles), c1 = ld2(oracles + 2), c2 = ld2(oracles + 4);
    if (!e2_eq(e2_add(c0, e2_add(c0, e2_add(c1, c2))), eval)) return 0;
    observe_ext2(t, c0); observe_ext2(t, c1); observe_ext2(t, c2);
    ext2 *chs = (ext2 *)malloc(nv * sizeof(ext2));
    for (int round = 0; round < nv - 1; round++) {
        ext2 ch = sample_ext2(t);
        chs[round] = ch;
        const u64 *o = oracles + 6L * round, *on = oracles + 6L * (round + 1);
        ext2 val = e2_add(ld2(o), e2_add(e2_mul(ld2(o + 2), ch), e2_mul(ld2(o + 4), e2_mul(ch, ch))));
        ext2 n0 = ld2(on), n1 = ld2(on + 2), n2 = ld2(on + 4);
        if (!e2_eq(val, e2_add(n0, e2_add(n0, e2_add(n1, n2))))) { free(chs); return 0; }
        observe_ext2(t, n0); observe_ext2(t, n1); observe_ext2(t, n2);
        observe_hash(t, roots + 4L * round);
    }
    chs[nv - 1] = sample_ext2(t);
    int ok = 1;
    for (int q = 0; q < nq && ok; q++) {
        (void)or_tr_challenge_scalar(t, NULL, 0);
        long index = (long)pc[0];
        const u64 *vals = pc + 1;
        const u64 *paths = vals + 4L * nv;
        u64 leaf[4];
        if (hash == OR_HASH_MONOLITH) or_monolith_hash_gl_leaf(vals[0], vals[2], leaf);
        else or_poseidon2_hash_gl_leaf(vals[0], vals[2], leaf);
        if (!or_merkle_verify_path(leaf, paths, nv + lr - 1, index, root4, hash)) { ok = 0; break; }
        const u64 *path = paths + 4L * (nv + lr - 1);
        long fidx = index;
        for (int r = 0; r < nv && ok; r++) {
            const u64 *lvl = table + 2 * table_level_offset(table_nv, lr, r);
            ext2 v0 = ld2(vals + 4L * r), v1 = ld2(vals + 4L * r + 2);
            ext2 fr = fold_one(v0, v1, lvl[2 * fidx], lvl[2 * fidx + 1], chs[r]);
            if (r < nv - 1) {
                ext2 nv0 = ld2(vals + 4L * (r + 1)), nv1 = ld2(vals + 4L * (r + 1) + 2);
                if (!e2_eq(fr, (fidx & 1) ? nv1 : nv0)) { ok = 0; break; }
                u64 lh[4];
                u64 a[2] = {nv0.c0, nv0.c1}, b[2] = {nv1.c0, nv1.c1};
                if (hash == OR_HASH_MONOLITH) or_monolith_hash_ext2_leaf(a, b, lh);
                else or_poseidon2_hash_ext2_leaf(a, b, lh);
                int plen = nv + lr - 2 - r;
                if (!or_merkle_verify_path(lh, path, plen, fidx / 2, roots + 4L * r, hash)) { ok = 0; break; }
                path += 4L * plen;
            } else {
                if (!e2_eq(fr, ld2(final_cw + 2 * fidx))) { ok = 0; break; }
            }
            fidx /= 2;
        }
        pc = paths + 4L * (nv + lr - 1);
        for (int i = 0; i <= nv - 2; i++) pc += 4L * (nv + lr - 2 - i);
    }
    free(chs);
    return ok;
}

/* ---------------------------------------------------------------------------------------
 * Batch opening of several commitments (same nv, lr) at several Ext2 points:
 *   batch_open_ext2       goldilocks-cuda-rs/src/basefold.rs:1202-1525
 *   batch_verify_ext2     goldilocks-cuda-rs/src/basefold.rs:1931-2221
 *   helpers               eq_poly_host :3008-3025 (first challenge = most significant
 *                         index bit), eq_eval_host_ext2 :3218-3230
 * Flat layout (u64, canonical):
 *   outer_oracles[nv][3][2] | combined_root[4] | inner_eval[2] | inner_oracles[nv][3][2] |
 *   folded_roots[nv-1][4] | final_cw[2^lr][2] |
 *   nq x { index | values[nv][2][2] | combined path[nv+lr-1][4] | folded path i: [nv+lr-2-i][4] } |
 *   nq x num_evals x { values[2] | path[nv+lr-1][4] }
 * --------------------------------------------------------------------------------------- */
static long combined_query_len(int nv, int lr) {
    long q = 1 + 4L * nv + 4L * (nv + lr - 1);
    for (int i = 0; i <= nv - 2; i++) q += 4L * (nv + lr - 2 - i);
    return q;
}

long or_basefold_batch_proof_len(int nv, int lr, int nq, int num_evals) {
    return 6L * nv + 4 + 2 + 6L * nv + 4L * (nv - 1) + 2L * (1L << lr) + (long)nq * combined_query_len(nv, lr) +
           (long)nq * num_evals * (2 + 4L * (nv + lr - 1));
}

static int batch_ell(int num_evals) {
    if (num_evals <= 1) return 1;
    int l = 0;
    while ((1 << l) < num_evals) l++;
    return l;
}

/* eq_poly_host: weights[i] = eq(t, bits of i), t_0 on the most significant bit */
static void eq_poly_msb(const u64 *t, int ell, u64 *out) {
    out[0] = 1;
    for (int i = 0; i < ell; i++) {
        long half = 1L << i;
        u64 omr = gl_sub(1, t[i]);
        for (long j = half - 1; j >= 0; j--) {
            u64 v = out[j];
            out[2 * j] = gl_mul(v, omr);
            out[2 * j + 1] = gl_mul(v, t[i]);
        }
    }
}

static ext2 eq_eval_ext2(const ext2 *r, const u64 *pt2, int nv) {
    ext2 res = e2_base(1), one = e2_base(1);
    for (int i = 0; i < nv; i++) {
        ext2 p = ld2(pt2 + 2 * i);
        ext2 term = e2_add(e2_mul(r[i], p), e2_mul(e2_sub(one, r[i]), e2_sub(one, p)));
        res = e2_mul(res, term);
    }
    return res;
}

static void fold_pairs(ext2 *a, long half, ext2 ch) {
    for (long j = 0; j < half; j++) a[j] = e2_add(a[2 * j], e2_mul(ch, e2_sub(a[2 * j + 1], a[2 * j])));
}

static void put_oracle(u64 *o, const ext2 *c, or_transcript *t) {
    for (int k = 0; k < 3; k++) { o[2 * k] = c[k].c0; o[2 * k + 1] = c[k].c1; observe_ext2(t, c[k]); }
}

void or_basefold_batch_open_ext2(int num_comms, const u64 *const *codewords, const u64 *const *bh_list,
                                 const u64 *roots, int nv, int lr, int hash, int num_points, const u64 *points,
                                 int num_evals, const int *eval_poly, const int *eval_point, const u64 *eval_values,
                                 const u64 *table, int table_nv, or_transcript *t, int nq, u64 *proof) {
    long n = 1L << nv, cwn = 1L << (nv + lr);
    for (int c = 0; c < num_comms; c++) observe_hash(t, roots + 4L * c);
    for (int p = 0; p < num_points; p++)
        for (int i = 0; i < nv; i++) observe_ext2(t, ld2(points + 2L * (p * nv + i)));
    for (int e = 0; e < num_evals; e++) observe_ext2(t, ld2(eval_values + 2L * e));
    int ell = batch_ell(num_evals);
    u64 tch[32], *w = (u64 *)malloc(sizeof(u64) << ell);
    for (int i = 0; i < ell; i++) tch[i] = or_tr_challenge_scalar(t, NULL, 0);
    eq_poly_msb(tch, ell, w);
    /* g_p (natural order) and eq_p */
    ext2 **g = (ext2 **)malloc(num_points * sizeof(ext2 *)), **eq = (ext2 **)malloc(num_points * sizeof(ext2 *));
    for (int p = 0; p < num_points; p++) {
        g[p] = (ext2 *)calloc(n, sizeof(ext2));
        eq[p] = (ext2 *)malloc(n * sizeof(ext2));
        or_eq_table_ext2(points + 2L * p * nv, nv, (u64 *)eq[p]);
    }
    for (int e = 0; e < num_evals; e++) {
        const u64 *bh = bh_list[eval_poly[e]];
        ext2 *gp = g[eval_point[e]];
        for (long x = 0; x < n; x++) gp[x] = e2_add(gp[x], e2_base(gl_mul(w[e], gl_canon(bh[bitrev(x, nv)]))));
    }
    u64 *pc = proof;
    u64 *outer = pc;
    pc += 6L * nv;
    u64 *comb_root = pc;
    pc += 4;
    u64 *inner_eval = pc;
    pc += 2;
    u64 *inner = pc;
    pc += 6L * nv;
    u64 *froots = pc;
    pc += 4L * (nv - 1);
    u64 *final_cw = pc;
    pc += 2L * (1L << lr);
    /* outer sumcheck over sum_p eq_p g_p */
    ext2 *r = (ext2 *)malloc(nv * sizeof(ext2)), c[3], cp[3];
    long size = n;
    for (int round = 0; round < nv; round++) {
        if (round > 0) {
            for (int p = 0; p < num_points; p++) { fold_pairs(eq[p], size / 2, r[round - 1]); fold_pairs(g[p], size / 2, r[round - 1]); }
            size /= 2;
        }
        c[0] = c[1] = c[2] = e2_zero();
        for (int p = 0; p < num_points; p++) {
            compute_oracle(g[p], eq[p], size, cp);
            for (int k = 0; k < 3; k++) c[k] = e2_add(c[k], cp[k]);
        }
        put_oracle(outer + 6L * round, c, t);
        r[round] = sample_ext2(t);
    }
    /* combined codeword and bh: sum_e eq(r, point_e) w_e (codeword_e, bh_e) */
    ext2 *ccw = (ext2 *)calloc(cwn, sizeof(ext2)), *cbh = (ext2 *)calloc(n, sizeof(ext2));
    for (int e = 0; e < num_evals; e++) {
        ext2 sc = e2_mul(eq_eval_ext2(r, points + 2L * eval_point[e] * nv, nv), e2_base(w[e]));
        const u64 *cw = codewords[eval_poly[e]], *bh = bh_list[eval_poly[e]];
        for (long y = 0; y < cwn; y++) ccw[y] = e2_add(ccw[y], e2_mul(sc, e2_base(gl_canon(cw[y]))));
        for (long y = 0; y < n; y++) cbh[bitrev(y, nv)] = e2_add(cbh[bitrev(y, nv)], e2_mul(sc, e2_base(gl_canon(bh[y]))));
    }
    u64 *ctree = (u64 *)malloc((cwn - 1) * 4 * sizeof(u64));
    or_merkle_tree_ext2((u64 *)ccw, cwn / 2, hash, ctree);
    memcpy(comb_root, ctree + or_merkle_root_offset(cwn / 2), 4 * sizeof(u64));
    observe_hash(t, comb_root);
    /* inner commit phase */
    ext2 *eqr = (ext2 *)malloc(n * sizeof(ext2));
    u64 *r2 = (u64 *)malloc(2 * nv * sizeof(u64));
    for (int i = 0; i < nv; i++) { r2[2 * i] = r[i].c0; r2[2 * i + 1] = r[i].c1; }
    or_eq_table_ext2(r2, nv, (u64 *)eqr);
    compute_oracle(cbh, eqr, n, c);
    ext2 ie = e2_add(c[0], e2_add(c[0], e2_add(c[1], c[2])));
    inner_eval[0] = ie.c0;
    inner_eval[1] = ie.c1;
    put_oracle(inner, c, t);
    ext2 **fcw = (ext2 **)calloc(nv, sizeof(ext2 *));
    u64 **ftree = (u64 **)calloc(nv, sizeof(u64 *));
    const ext2 *cur = ccw;
    long fsize = cwn / 2;
    size = n;
    for (int round = 0; round < nv - 1; round++) {
        ext2 ch = sample_ext2(t);
        fold_pairs(eqr, size / 2, ch);
        fold_pairs(cbh, size / 2, ch);
        size /= 2;
        compute_oracle(cbh, eqr, size, c);
        put_oracle(inner + 6L * (round + 1), c, t);
        const u64 *lvl = table + 2 * table_level_offset(table_nv, lr, round);
        fcw[round] = (ext2 *)malloc(fsize * sizeof(ext2));
        for (long i = 0; i < fsize; i++) fcw[round][i] = fold_one(cur[2 * i], cur[2 * i + 1], lvl[2 * i], lvl[2 * i + 1], ch);
        ftree[round] = (u64 *)malloc((fsize - 1) * 4 * sizeof(u64));
        or_merkle_tree_ext2((u64 *)fcw[round], fsize / 2, hash, ftree[round]);
        memcpy(froots + 4L * round, ftree[round] + or_merkle_root_offset(fsize / 2), 4 * sizeof(u64));
        observe_hash(t, froots + 4L * round);
        cur = fcw[round];
        fsize /= 2;
    }
    {
        ext2 ch = sample_ext2(t);
        const u64 *lvl = table + 2 * table_level_offset(table_nv, lr, nv - 1);
        for (long i = 0; i < fsize; i++) {
            ext2 v = fold_one(cur[2 * i], cur[2 * i + 1], lvl[2 * i], lvl[2 * i + 1], ch);
            final_cw[2 * i] = v.c0;
            final_cw[2 * i + 1] = v.c1;
        }
    }
    /* queries */
    long *qidx = (long *)malloc((nq > 0 ? nq : 1) * sizeof(long));
    for (int q = 0; q < nq; q++) qidx[q] = (long)(or_tr_challenge_scalar(t, NULL, 0) % (u64)(cwn / 2));
    for (int q = 0; q < nq; q++) {
        long leaf = qidx[q];
        u64 *qp = pc;
        *qp++ = (u64)leaf;
        u64 *vals = qp;
        qp += 4L * nv;
        ext2 a = ccw[2 * leaf], b = ccw[2 * leaf + 1];
        vals[0] = a.c0; vals[1] = a.c1; vals[2] = b.c0; vals[3] = b.c1;
        long idx = leaf / 2;
        for (int i = 0; i <= nv - 2; i++) {
            a = fcw[i][2 * idx];
            b = fcw[i][2 * idx + 1];
            u64 *v = vals + 4L * (i + 1);
            v[0] = a.c0; v[1] = a.c1; v[2] = b.c0; v[3] = b.c1;
            idx /= 2;
        }
        tree_path(ctree, cwn / 2, leaf, qp);
        qp += 4L * (nv + lr - 1);
        idx = leaf / 2;
        long leaves = cwn / 4;
        for (int i = 0; i <= nv - 2; i++) {
            tree_path(ftree[i], leaves, idx, qp);
            qp += 4L * (nv + lr - 2 - i);
            idx /= 2;
            leaves /= 2;
        }
        pc = qp;
    }
    u64 **ptrees = (u64 **)calloc(num_comms, sizeof(u64 *));
    for (int q = 0; q < nq; q++) {
        long leaf = qidx[q];
        for (int e = 0; e < num_evals; e++) {
            int pi = eval_poly[e];
            if (!ptrees[pi]) {
                ptrees[pi] = (u64 *)malloc((cwn - 1) * 4 * sizeof(u64));
                or_merkle_tree_gl(codewords[pi], cwn / 2, hash, ptrees[pi]);
            }
            pc[0] = gl_canon(codewords[pi][2 * leaf]);
            pc[1] = gl_canon(codewords[pi][2 * leaf + 1]);
            tree_path(ptrees[pi], cwn / 2, leaf, pc + 2);
            pc += 2 + 4L * (nv + lr - 1);
        }
    }
    for (int c2 = 0; c2 < num_comms; c2++) free(ptrees[c2]);
    for (int i = 0; i < nv; i++) { free(fcw[i]); free(ftree[i]); }
    for (int p = 0; p < num_points; p++) { free(g[p]); free(eq[p]); }
    free(ptrees); free(fcw); free(ftree); free(qidx); free(eqr); free(r2); free(ctree); free(ccw); free(cbh);
    free(r); free(g); free(eq); free(w);
}

static int check_transition(const u64 *o, const u64 *on, ext2 ch) {
    ext2 val = e2_add(ld2(o), e2_add(e2_mul(ld2(o + 2), ch), e2_mul(ld2(o + 4), e2_mul(ch, ch))));
    ext2 n0 = ld2(on);
    return e2_eq(val, e2_add(n0, e2_add(n0, e2_add(ld2(on + 2), ld2(on + 4)))));
}

int or_basefold_batch_verify_ext2(int num_comms, const u64 *roots, int nv, int lr, int hash, int num_points,
                                  const u64 *points, int num_evals, const int *eval_poly, const int *eval_point,
                                  const u64 *eval_values, const u64 *proof, int nq, const u64 *table, int table_nv,
                                  or_transcript *t) {
    const u64 *pc = proof;
    const u64 *outer = pc;
    pc += 6L * nv;
    const u64 *comb_root = pc;
    pc += 4;
    ext2 inner_eval = ld2(pc);
    pc += 2;
    const u64 *inner = pc;
    pc += 6L * nv;
    const u64 *froots = pc;
    pc += 4L * (nv - 1);
    const u64 *final_cw = pc;
    pc += 2L * (1L << lr);
    for (int c = 0; c < num_comms; c++) observe_hash(t, roots + 4L * c);
    for (int p = 0; p < num_points; p++)
        for (int i = 0; i < nv; i++) observe_ext2(t, ld2(points + 2L * (p * nv + i)));
    for (int e = 0; e < num_evals; e++) observe_ext2(t, ld2(eval_values + 2L * e));
    int ell = batch_ell(num_evals);
    u64 tch[32], *w = (u64 *)malloc(sizeof(u64) << ell);
    for (int i = 0; i < ell; i++) tch[i] = or_tr_challenge_scalar(t, NULL, 0);
    eq_poly_msb(tch, ell, w);
    ext2 claimed = e2_zero();
    for (int e = 0; e < num_evals; e++) claimed = e2_add(claimed, e2_mul(e2_base(w[e]), ld2(eval_values + 2L * e)));
    int ok = 1;
    ext2 o0 = ld2(outer);
    if (!e2_eq(e2_add(o0, e2_add(o0, e2_add(ld2(outer + 2), ld2(outer + 4)))), claimed)) ok = 0;
    ext2 *r = (ext2 *)malloc(nv * sizeof(ext2)), *ri = (ext2 *)malloc(nv * sizeof(ext2));
    for (int k = 0; k < 3; k++) observe_ext2(t, ld2(outer + 2 * k));
    for (int round = 0; round < nv - 1 && ok; round++) {
        r[round] = sample_ext2(t);
        if (!check_transition(outer + 6L * round, outer + 6L * (round + 1), r[round])) ok = 0;
        for (int k = 0; k < 3; k++) observe_ext2(t, ld2(outer + 6L * (round + 1) + 2 * k));
    }
    if (ok) {
        r[nv - 1] = sample_ext2(t);
        const u64 *o = outer + 6L * (nv - 1);
        ext2 ch = r[nv - 1];
        ext2 fin = e2_add(ld2(o), e2_add(e2_mul(ld2(o + 2), ch), e2_mul(ld2(o + 4), e2_mul(ch, ch))));
        if (!e2_eq(fin, inner_eval)) ok = 0;
    }
    if (ok) {
        observe_hash(t, comb_root);
        ext2 i0 = ld2(inner);
        if (!e2_eq(e2_add(i0, e2_add(i0, e2_add(ld2(inner + 2), ld2(inner + 4)))), inner_eval)) ok = 0;
        for (int k = 0; k < 3; k++) observe_ext2(t, ld2(inner + 2 * k));
    }
    for (int round = 0; round < nv - 1 && ok; round++) {
        ri[round] = sample_ext2(t);
        if (!check_transition(inner + 6L * round, inner + 6L * (round + 1), ri[round])) ok = 0;
        for (int k = 0; k < 3; k++) observe_ext2(t, ld2(inner + 6L * (round + 1) + 2 * k));
        observe_hash(t, froots + 4L * round);
    }
    if (ok) ri[nv - 1] = sample_ext2(t);
    long cwn = 1L << (nv + lr);
    ext2 *sc = (ext2 *)malloc((num_evals > 0 ? num_evals : 1) * sizeof(ext2));
    for (int e = 0; e < num_evals && ok; e++)
        sc[e] = e2_mul(eq_eval_ext2(r, points + 2L * eval_point[e] * nv, nv), e2_base(w[e]));
    const long qlen = combined_query_len(nv, lr);
    const u64 *indiv = pc + (long)nq * qlen;
    const long ilen = 2 + 4L * (nv + lr - 1);
    for (int q = 0; q < nq && ok; q++) {
        long leaf = (long)(or_tr_challenge_scalar(t, NULL, 0) % (u64)(cwn / 2));
        const u64 *qp = pc + (long)q * qlen;
        if ((long)qp[0] != leaf) { ok = 0; break; }
        const u64 *vals = qp + 1;
        const u64 *path = vals + 4L * nv;
        u64 lh[4];
        if (hash == OR_HASH_MONOLITH) or_monolith_hash_ext2_leaf(vals, vals + 2, lh);
        else or_poseidon2_hash_ext2_leaf(vals, vals + 2, lh);
        if (!or_merkle_verify_path(lh, path, nv + lr - 1, leaf, comb_root, hash)) { ok = 0; break; }
        path += 4L * (nv + lr - 1);
        long fidx = leaf;
        for (int rr = 0; rr < nv && ok; rr++) {
            const u64 *lvl = table + 2 * table_level_offset(table_nv, lr, rr);
            ext2 fr = fold_one(ld2(vals + 4L * rr), ld2(vals + 4L * rr + 2), lvl[2 * fidx], lvl[2 * fidx + 1], ri[rr]);
            if (rr < nv - 1) {
                const u64 *nx = vals + 4L * (rr + 1);
                if (!e2_eq(fr, (fidx & 1) ? ld2(nx + 2) : ld2(nx))) { ok = 0; break; }
                if (hash == OR_HASH_MONOLITH) or_monolith_hash_ext2_leaf(nx, nx + 2, lh);
                else or_poseidon2_hash_ext2_leaf(nx, nx + 2, lh);
                int plen = nv + lr - 2 - rr;
                if (!or_merkle_verify_path(lh, path, plen, fidx / 2, froots + 4L * rr, hash)) { ok = 0; break; }
                path += 4L * plen;
            } else if (!e2_eq(fr, ld2(final_cw + 2 * fidx))) {
                ok = 0;
            }
            fidx /= 2;
        }
        ext2 cv0 = e2_zero(), cv1 = e2_zero();
        for (int e = 0; e < num_evals && ok; e++) {
            const u64 *ip = indiv + ((long)q * num_evals + e) * ilen;
            u64 leaf_h[4];
            if (hash == OR_HASH_MONOLITH) or_monolith_hash_gl_leaf(ip[0], ip[1], leaf_h);
            else or_poseidon2_hash_gl_leaf(ip[0], ip[1], leaf_h);
            if (!or_merkle_verify_path(leaf_h, ip + 2, nv + lr - 1, leaf, roots + 4L * eval_poly[e], hash)) ok = 0;
            cv0 = e2_add(cv0, e2_mul(sc[e], e2_base(gl_canon(ip[0]))));
            cv1 = e2_add(cv1, e2_mul(sc[e], e2_base(gl_canon(ip[1]))));
        }
        if (ok && (!e2_eq(cv0, ld2(vals)) || !e2_eq(cv1, ld2(vals + 2)))) ok = 0;
    }
    free(sc); free(r); free(ri); free(w);
    return ok;
}
