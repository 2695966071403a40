// Linear sumcheck prover over the almost-Goldilocks Ext2 (q, X^2 - 3) on the device, driven
// by the zk-torch-4 transcript (Monolith over q).
//
// Mirrors GpuLinearSumcheckProver::new + prove / prove_ext2 of
// zk-torch-4/src/sumcheck/gpu_prover.rs:26-112 (labels "num_var", "num_poly",
// "round_message", "challenge"), whose proofs are those of the CPU prover
// zk-torch-4/src/sumcheck/cpu_ext2_prover.rs:23-100: per round the message at c = 0..num_poly
// of sum_y prod_i (p_i[2y] + c (p_i[2y+1] - p_i[2y])), each value appended, an Ext2 challenge
// drawn, all polynomials folded at it; final_eval = the product of the final values.
//
// Device work per round is one fused fold + next-message pass over the arrays (the same
// kernels as the Goldilocks prover, instantiated for q), and the message comes back through
// the thread's pinned mailbox. The input buffer is never written.
#include <string.h>

#include "common.h"
#include "qtranscript.h"
#include "sumcheck.h"

namespace {

const uint8_t L_NUM_VAR[] = "num_var";
const uint8_t L_NUM_POLY[] = "num_poly";
const uint8_t L_ROUND_MSG[] = "round_message";
const uint8_t L_CHALLENGE[] = "challenge";

size_t align_words(size_t w) { return (w + 31) & ~(size_t)31; }

}  // namespace

// d_polys: num_polys polynomials of 2^num_vars elements back to back, base-field words
// (input_is_base = 1, prove) or interleaved Ext2 (prove_ext2). Outputs (host): round_msgs
// [num_vars][num_polys + 1][2], challenges [num_vars][2], final_evals [num_polys][2],
// final_eval [2]. Returns 0, -1 (runtime error) or -2 (rejected argument).
MAM_API int mam_q_sumcheck_prove(mam_qtranscript *t, const u64 *d_polys, int num_polys, int num_vars,
                                 int input_is_base, u64 *round_msgs, u64 *challenges, u64 *final_evals,
                                 u64 *final_eval) {
    const int np = num_polys, nv = num_vars;
    if (!t || !d_polys || np < 1 || np > MAM_MAX_POLYS || nv < 1 || nv > 34 || !round_msgs || !challenges ||
        !final_evals || !final_eval)
        return -2;
    const hipStream_t st = mam_stream();
    const bool base = input_is_base != 0;
    const size_t n = (size_t)1 << nv;
    // buf[0] holds rounds 1, 3, ... (n / 2 Ext2 per polynomial), buf[1] rounds 2, 4, ...
    const size_t stride0 = n, stride1 = n / 2 > 1 ? n / 2 : 2;
    const size_t w0 = align_words(np * stride0), w1 = align_words(np * stride1);
    const size_t wp = align_words((size_t)MAM_RED_MAX_BLOCKS * 2 * (MAM_MAX_POLYS + 1)), wm = 32, wf = 32;
    u64 *arena = mam_dev_arena(w0 + w1 + wp + wm + wf, 1);
    if (!arena) return -1;
    u64 *buf[2] = {arena, arena + w0};
    const size_t stride[2] = {stride0, stride1};
    ScScratch s = {arena + w0 + w1, nullptr, arena + w0 + w1 + wp};
    u64 *finals = s.msg + wm;

    const u64 nv64 = (u64)nv, np64 = (u64)np;
    mam::qtr_append_u64s(t, L_NUM_VAR, 7, &nv64, 1);
    mam::qtr_append_u64s(t, L_NUM_POLY, 8, &np64, 1);

    const int dp1 = np + 1;
    const u64 *cur = d_polys;
    size_t cur_stride = base ? n : 2 * n, cur_size = n;
    bool cur_base = base;
    int nxt = 0;
    qe2 ch = q2_zero();
    u64 host[2 * (MAM_MAX_POLYS + 1)];
    for (int r = 0; r < nv; r++) {
        int rc;
        if (r == 0) {
            rc = mam_qsc_launch_msg(np, cur, cur_stride, cur_size / 2, cur_base, s, st);
        } else {
            rc = mam_qsc_launch_fold_msg(np, cur, cur_stride, buf[nxt], stride[nxt], cur_size / 4, cur_base, ch, s, st);
            cur = buf[nxt];
            cur_stride = stride[nxt];
            cur_size /= 2;
            cur_base = false;
            nxt ^= 1;
        }
        // the first message's label does not depend on the message: absorb it (its full block
        // runs the permutation) while the device computes the message; the sequence is the same
        if (!rc) mam::qtr_append_u64s(t, L_ROUND_MSG, 13, nullptr, 0);
        if (!rc) rc = mam_publish_wait(s.msg, 2 * dp1, nullptr, 0, host, st);
        if (rc) return rc;
        u64 *m = round_msgs + (size_t)r * dp1 * 2;
        memcpy(m, host, dp1 * 2 * sizeof(u64));
        for (int c = 0; c < dp1; c++) mam::qtr_append_ext2(t, c == 0 ? nullptr : L_ROUND_MSG, c == 0 ? 0 : 13, m + 2 * c);
        u64 chw[2];
        mam::qtr_challenge_scalars(t, L_CHALLENGE, 9, chw, 2);
        challenges[2 * r] = chw[0];
        challenges[2 * r + 1] = chw[1];
        ch = qe2{fq::canon(chw[0]), fq::canon(chw[1])};
    }
    int rc = mam_qsc_launch_fold_final(np, cur, cur_stride, cur_base, ch, finals, st);
    if (!rc) rc = mam_publish_wait(finals, 2 * np, nullptr, 0, host, st);
    if (rc) return rc;
    qe2 fe = q2_one();
    for (int i = 0; i < np; i++) {
        final_evals[2 * i] = host[2 * i];
        final_evals[2 * i + 1] = host[2 * i + 1];
        fe = q2_mul(fe, qe2{host[2 * i], host[2 * i + 1]});
    }
    final_eval[0] = fe.c0;
    final_eval[1] = fe.c1;
    return 0;
}
