// Internal interface of the Basefold kernels (k_basefold.hip, k_merkle.hip fold leaves).
#pragma once

#include <hip/hip_runtime.h>

#include "field.h"

// Workgroup cap of the in-launch reductions of the opening (sizes RedScratch::partials).
#define RED_BLOCKS 2048

struct RedScratch {
    u64 *partials;  // RED_BLOCKS * 6 words
    u32 *counter;   // zero-initialized ticket counter
    u64 *result;    // 6 words: oracle (c0, c1, c2) as Ext2
};

// v0 + cx * (v1 - v0) * w   (cuda/basefold.cuh:397-419, 739-761)
MAM_HD ext2 mam_fold_value(ext2 v0, ext2 v1, ext2 cx, u64 w) {
    return e2_add(v0, e2_mul(cx, e2_scale(w, e2_sub(v1, v0))));
}

// Degree-2 oracle terms of one pair in coefficient form:
// c0 += e0 b0, c1 += e0 b1 + e1 b0, c2 += e1 b1 (acc holds c0, c1, c2 as Ext2 word pairs).
MAM_HD void mam_oracle_accumulate(u64 *acc, ext2 e0, ext2 e1, ext2 b0, ext2 b1) {
    ext2 t0 = e2_mul(e0, b0);
    ext2 t2 = e2_mul(e1, b1);
    // e0 b1 + e1 b0 = (e0 + e1)(b0 + b1) - e0 b0 - e1 b1: three products instead of four
    ext2 t1 = e2_sub(e2_sub(e2_mul(e2_add(e0, e1), e2_add(b0, b1)), t0), t2);
    acc[0] = gl_add(acc[0], t0.c0);
    acc[1] = gl_add(acc[1], t0.c1);
    acc[2] = gl_add(acc[2], t1.c0);
    acc[3] = gl_add(acc[3], t1.c1);
    acc[4] = gl_add(acc[4], t2.c0);
    acc[5] = gl_add(acc[5], t2.c1);
}

// The same terms from the two end values of each factor (e(0) = e0, e(1) = e1, likewise b):
// c1 = e1 b1 - c0 - c2 needs no separate slope products.
MAM_HD void mam_oracle_accumulate_ends(u64 *acc, ext2 e0, ext2 e1, ext2 b0, ext2 b1) {
    ext2 t0 = e2_mul(e0, b0);
    ext2 t2 = e2_mul(e2_sub(e1, e0), e2_sub(b1, b0));
    ext2 t1 = e2_sub(e2_sub(e2_mul(e1, b1), t0), t2);
    acc[0] = gl_add(acc[0], t0.c0);
    acc[1] = gl_add(acc[1], t0.c1);
    acc[2] = gl_add(acc[2], t1.c0);
    acc[3] = gl_add(acc[3], t1.c1);
    acc[4] = gl_add(acc[4], t2.c0);
    acc[5] = gl_add(acc[5], t2.c1);
}

int mam_bitrev(const u64 *src, u64 *dst, int nb, int words, hipStream_t s);
int mam_commit_codeword(const u64 *d_evals, u64 *d_bh_evals, u64 *d_codeword, int nv, int lr, hipStream_t s);
// alpha_p (optional): the challenge is read from device memory there instead of alpha.
int mam_cw_fold(const u64 *cw, const u64 *table, ext2 alpha, u64 *out, size_t pairs, bool base_cw, hipStream_t s,
                const u64 *alpha_p = nullptr);
int mam_eq_table(const u64 *d_r, int nv, u64 *d_out, bool ext, hipStream_t s);
// s * eq(r, x) over nv variables (Ext2 table): variable 0 contributes s (1 - r_0) / s r_0.
int mam_eq_table_scaled(const u64 *d_r, int nv, ext2 s0, ext2 s1, u64 *d_out, hipStream_t s);
// Levels with fewer nodes than this use the lane-parallel Merkle kernel (and carry an
// opening round's fold job along).
size_t mam_merkle_narrow_min();
// One opening round in product form (single evaluation point): with FOLD, fold bh (4y .. 4y+3
// -> folded pair y, written to bh_out) at ch; then U0 = sum_y E(y) b(2y), U1 = sum_y E(y)
// b(2y+1) over `pairs` pairs y, E(y) = eq(r, y) over nb variables (r: device Ext2 point
// words). result: U0, U1 (4 words); partials: RedScratch partials. chp (optional): ch is read
// from device memory there. rev: bh in bit-reversed order and r reversed (UJob::rev).
int mam_open_u(const u64 *bh, bool bh_base, bool fold, ext2 ch, u64 *bh_out, size_t pairs, const u64 *r, int nb,
               const RedScratch &rs, hipStream_t s, const u64 *chp = nullptr, bool rev = false);
int mam_open_oracle0(const u64 *eq, const u64 *bh, bool bh_base, size_t pairs, const RedScratch &rs,
                     hipStream_t s);
int mam_open_fold_oracle(const u64 *eq, const u64 *bh, bool bh_base, ext2 ch, u64 *eq_out, u64 *bh_out,
                         size_t half_out, const RedScratch &rs, hipStream_t s);
// Fold a codeword (base or Ext2) at alpha into `fold_out` (Ext2) and build the Merkle tree
// over the folded codeword in the same pass (leaf i = folded elements 2i, 2i+1). `oj`
// (open_fold.h, may be null): the round's eq / bh fold and oracle job, run along with the
// tree's first launch where it can be, else launched before it. `mj` (may be null): a
// mailbox post of mj->a's words followed by the root, made by the tree's final launch when
// it can (*mail_posted true); otherwise the caller posts. `uj` (may be null): a product-form
// round's bh fold and U sums (open_fold.h), run on extra workgroups of the tree's first
// launch, which is a one-lane launch whenever the round is in product form (the result is in
// uj->result before any later launch of the stream starts). alpha_p (may be null): the fold
// reads alpha from device memory there (a value written on the device after the launch).
struct OracleJob;
struct MailJob;
struct UJob;
int mam_fold_and_tree(const u64 *cw, bool base_cw, const u64 *table, ext2 alpha, u64 *fold_out, u64 *tree,
                      size_t folded_len, int hash, hipStream_t s, const OracleJob *oj = nullptr,
                      const MailJob *mj = nullptr, bool *mail_posted = nullptr, const UJob *uj = nullptr,
                      const u64 *alpha_p = nullptr);
