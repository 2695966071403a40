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

static void compute_oracle(const ext2 *bh, const ext2 *eq, long size, ext2 *c) {
    ext2 two = e2_base(2), p0 = e2_zero(), p1 = e2_zero(), p2 = e2_zero();
    for (long j = 0; j < size / 2; j++) {
        p0 = e2_add(p0, e2_mul(eq[2 * j], bh[2 * j]));
        p1 = e2_add(p1, e2_mul(eq[2 * j + 1], bh[2 * j + 1]));
        ext2 ei = e2_sub(e2_mul(two, eq[2 * j + 1]), eq[2 * j]);
        ext2 bi = e2_sub(e2_mul(two, bh[2 * j + 1]), bh[2 * j]);
        p2 = e2_add(p2, e2_mul(ei, bi));
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

void or_basefold_open_ext2(const u64 *codeword, const u64 *bh_evals, const u64 *root4, int nv, int lr,
                           int hash, const u64 *point2, const u64 *table, int table_nv, or_transcript *t,
                           int nq, u64 *proof) {
    long n = 1L << nv, cwn = 1L << (nv + lr);
    u64 *pc = proof;
    /* transcript preamble */
    observe_hash(t, root4);
    for (int i = 0; i < nv; i++) observe_ext2(t, e2(point2[2 * i], point2[2 * i + 1]));
    /* eq table and bh in Type2 order */
    ext2 *eq = (ext2 *)malloc(n * sizeof(ext2));
    or_eq_table_ext2(point2, nv, (u64 *)eq);
    ext2 *bh = (ext2 *)malloc(n * sizeof(ext2));
    for (long i = 0; i < n; i++) bh[bitrev(i, nv)] = e2_base(bh_evals[i]);
    ext2 eval = e2_zero();
    for (long i = 0; i < n; i++) eval = e2_add(eval, e2_mul(bh[i], eq[i]));
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
    for (long j = 0; j < sc; j++) {
        bh[j] = e2_add(bh[2 * j], e2_mul(ch, e2_sub(bh[2 * j + 1], bh[2 * j])));
        eq[j] = e2_add(eq[2 * j], e2_mul(ch, e2_sub(eq[2 * j + 1], eq[2 * j])));
    }
    compute_oracle(bh, eq, sc, c);
    for (int k = 0; k < 3; k++) { oracles[6 + 2 * k] = c[k].c0; oracles[6 + 2 * k + 1] = c[k].c1; observe_ext2(t, c[k]); }

    /* folded codewords (kept for queries) and their trees */
    ext2 **fcw = (ext2 **)calloc(nv, sizeof(ext2 *));
    u64 **ftree = (u64 **)calloc(nv, sizeof(u64 *));
    long fsize = cwn / 2;
    const u64 *lvl = table + 2 * table_level_offset(table_nv, lr, 0);
    fcw[0] = (ext2 *)malloc(fsize * sizeof(ext2));
    for (long i = 0; i < fsize; i++)
        fcw[0][i] = fold_one(e2_base(codeword[2 * i]), e2_base(codeword[2 * i + 1]), lvl[2 * i], lvl[2 * i + 1], ch);
    ftree[0] = (u64 *)malloc((2 * (fsize / 2) - 1) * 4 * sizeof(u64));
    or_merkle_tree_ext2((u64 *)fcw[0], fsize / 2, hash, ftree[0]);
    memcpy(roots, ftree[0] + or_merkle_root_offset(fsize / 2), 4 * sizeof(u64));
    observe_hash(t, roots);

    for (int round = 1; round < nv - 1; round++) {
        ch = sample_ext2(t);
        long half = sc / 2;
        for (long j = 0; j < half; j++) {
            eq[j] = e2_add(eq[2 * j], e2_mul(ch, e2_sub(eq[2 * j + 1], eq[2 * j])));
            bh[j] = e2_add(bh[2 * j], e2_mul(ch, e2_sub(bh[2 * j + 1], bh[2 * j])));
        }
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
    u64 *itree = (u64 *)malloc((2 * (cwn / 2) - 1) * 4 * sizeof(u64));
    or_merkle_tree_gl(codeword, cwn / 2, hash, itree);
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
    free(fcw); free(ftree); free(itree); free(qidx); free(eq); free(bh);
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
