// One round of the Basefold opening's sumcheck on the device: fold eq (Ext2) and bh (base
// or Ext2) at the round's challenge and reduce the next round's oracle in the same pass
// (cuda/basefold.cuh:654-735 fused with the folds of :397-761). Each output pair y comes
// from inputs 4y .. 4y + 3.
//
// The body runs either as its own launch (k_open_fold_oracle) or on extra workgroups of the
// first Merkle launch of the round's folded codeword (k_tree_wide): both depend only on the
// challenge, and in the small rounds each is a latency-bound launch of its own.
#pragma once

#include "basefold.h"
#include "reduce.h"

struct OracleJob {
    const u64 *eq, *bh;  // round inputs (bh base words when bh_base)
    ext2 ch;
    u64 *eq_out, *bh_out;  // folded outputs (Ext2), half_out pairs each
    size_t half_out;
    u64 *partials;  // reduction scratch (RedScratch)
    u32 *counter;
    u64 *result;    // 6 words: the next oracle (c0, c1, c2) as Ext2
    unsigned blocks;  // workgroups running the job (0: no job)
    int bh_base;
};

// `lds` holds 6 * (blockDim.x / 64) words; bid / nblocks number the job's workgroups.
// TICKET: finish the oracle with the one-launch ticketed reduction (the job rides along a
// tree launch); otherwise leave the workgroups' partials for k_sum_partials (a launch of its
// own, whose caller adds them in a second launch).
template <bool BH_BASE, bool TICKET>
__device__ __forceinline__ void open_fold_oracle_run(const OracleJob &j, unsigned bid, unsigned nblocks, u64 *lds) {
    u64 acc[6] = {0, 0, 0, 0, 0, 0};
    const ext2 ch = j.ch;
    const u64 ch7 = gl_mul7(ch.c1);
    for (size_t y = bid * (size_t)blockDim.x + threadIdx.x; y < j.half_out; y += (size_t)nblocks * blockDim.x) {
        ext2 q0 = e2_load(j.eq, 4 * y), q1 = e2_load(j.eq, 4 * y + 1), q2 = e2_load(j.eq, 4 * y + 2),
             q3 = e2_load(j.eq, 4 * y + 3);
        ext2 ea = e2_add(q0, e2_mul_fixed(ch, ch7, e2_sub(q1, q0)));
        ext2 eb = e2_add(q2, e2_mul_fixed(ch, ch7, e2_sub(q3, q2)));
        ext2 ba, bb;
        if (BH_BASE) {
            u64 a0 = gl_canon(j.bh[4 * y]), a1 = gl_canon(j.bh[4 * y + 1]);
            u64 a2 = gl_canon(j.bh[4 * y + 2]), a3 = gl_canon(j.bh[4 * y + 3]);
            ba = e2_add(e2_base(a0), e2_scale(gl_sub(a1, a0), ch));
            bb = e2_add(e2_base(a2), e2_scale(gl_sub(a3, a2), ch));
        } else {
            ext2 b0 = e2_load(j.bh, 4 * y), b1 = e2_load(j.bh, 4 * y + 1), b2 = e2_load(j.bh, 4 * y + 2),
                 b3 = e2_load(j.bh, 4 * y + 3);
            ba = e2_add(b0, e2_mul_fixed(ch, ch7, e2_sub(b1, b0)));
            bb = e2_add(b2, e2_mul_fixed(ch, ch7, e2_sub(b3, b2)));
        }
        e2_store(j.eq_out, 2 * y, ea);
        e2_store(j.eq_out, 2 * y + 1, eb);
        e2_store(j.bh_out, 2 * y, ba);
        e2_store(j.bh_out, 2 * y + 1, bb);
        mam_oracle_accumulate_ends(acc, ea, eb, ba, bb);
    }
    block_sum<6>(acc, lds);
    if (TICKET)
        grid_finish_n<6>(acc, lds, j.partials, j.counter, j.result, bid, nblocks);
    else
        grid_partials<6>(acc, j.partials, j.result);
}

// The job as its own launch when it cannot ride along a tree launch (k_basefold.hip).
int mam_open_fold_oracle_job(const OracleJob &j, hipStream_t s);
