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
    const u64 *chp;  // non-null: the challenge is read here instead of ch
    // 1: eq and bh in bit-reversed order (index k holds natural index brev(k)); output pair
    // k <- inputs k, k + 2Q and k + Q, k + 3Q (Q = half_out), written at k and k + Q
    int rev;
};

// `lds` holds 6 * (blockDim.x / 64) words; bid / nblocks number the job's workgroups.
// TICKET: finish the oracle with the one-launch ticketed reduction (the job rides along a
// tree launch); otherwise leave the workgroups' partials for k_sum_partials (a launch of its
// own, whose caller adds them in a second launch).
template <bool BH_BASE, bool TICKET>
__device__ __forceinline__ void open_fold_oracle_run(const OracleJob &j, unsigned bid, unsigned nblocks, u64 *lds) {
    u64 acc[6] = {0, 0, 0, 0, 0, 0};
    const ext2 ch = j.chp ? ext2{j.chp[0], j.chp[1]} : j.ch;
    const u64 ch7 = gl_mul7(ch.c1);
    const size_t Q = j.half_out;
    for (size_t y = bid * (size_t)blockDim.x + threadIdx.x; y < j.half_out; y += (size_t)nblocks * blockDim.x) {
        // inputs of output pair y: a0, a1 fold to its first element, a2, a3 to its second
        const size_t i0 = j.rev ? y : 4 * y, i1 = j.rev ? y + 2 * Q : 4 * y + 1;
        const size_t i2 = j.rev ? y + Q : 4 * y + 2, i3 = j.rev ? y + 3 * Q : 4 * y + 3;
        const size_t o0 = j.rev ? y : 2 * y, o1 = j.rev ? y + Q : 2 * y + 1;
        ext2 q0 = e2_load(j.eq, i0), q1 = e2_load(j.eq, i1), q2 = e2_load(j.eq, i2), q3 = e2_load(j.eq, i3);
        ext2 ea = e2_add(q0, e2_mul_fixed(ch, ch7, e2_sub(q1, q0)));
        ext2 eb = e2_add(q2, e2_mul_fixed(ch, ch7, e2_sub(q3, q2)));
        ext2 ba, bb;
        if (BH_BASE) {
            u64 a0 = gl_canon(j.bh[i0]), a1 = gl_canon(j.bh[i1]);
            u64 a2 = gl_canon(j.bh[i2]), a3 = gl_canon(j.bh[i3]);
            ba = e2_add(e2_base(a0), e2_scale(gl_sub(a1, a0), ch));
            bb = e2_add(e2_base(a2), e2_scale(gl_sub(a3, a2), ch));
        } else {
            ext2 b0 = e2_load(j.bh, i0), b1 = e2_load(j.bh, i1), b2 = e2_load(j.bh, i2), b3 = e2_load(j.bh, i3);
            ba = e2_add(b0, e2_mul_fixed(ch, ch7, e2_sub(b1, b0)));
            bb = e2_add(b2, e2_mul_fixed(ch, ch7, e2_sub(b3, b2)));
        }
        e2_store(j.eq_out, o0, ea);
        e2_store(j.eq_out, o1, eb);
        e2_store(j.bh_out, o0, ba);
        e2_store(j.bh_out, o1, bb);
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

// Product form of an opening round (single point): eq never stored. Folding eq at ch on its
// lowest variable gives the eq table of the remaining variables times a scalar, so a round's
// oracle needs only U0 = sum_y E(y) b(2y) and U1 = sum_y E(y) b(2y+1), E the eq table of the
// variables above the pair bit; the host applies the scalars. E(y) is the product of one
// entry per 6-bit group of y from 64-entry tables built in LDS by each workgroup.
struct UJob {
    const u64 *bh;
    u64 *bh_out;
    ext2 ch;
    size_t pairs;
    const u64 *r;  // Ext2 point words of E's variables
    int nb;        // E's variables (bits of y), <= 30
    u64 *partials, *result;
    u32 *counter;  // ticket of the one-launch reduction (a job riding along a tree launch)
    const u64 *chp;  // non-null: the challenge is read here instead of ch
    // 1: bh in bit-reversed order (index k holds natural index brev(k)) and r the point's
    // coordinates in reversed order: pair y reads k = y and y + Q (y + 2Q, y + Q, y + 3Q with
    // the fold, Q = pairs) and writes its folded pair at y and y + Q; E(y) is then the
    // natural E at brev(y), which is what the reversed coordinates give
    int rev;
};

template <bool BH_BASE, bool FOLD, bool TICKET>
__device__ __forceinline__ void open_u_run(const UJob &j, unsigned bid, unsigned nblocks) {
    __shared__ u64 tab[5][64][2];
    __shared__ u64 lds[4 * (MAM_BLOCK / WAVE)];
    for (int e = threadIdx.x; e < 5 * 64; e += blockDim.x) {
        const int g = e >> 6, m = e & 63;
        ext2 v = e2_one();
        for (int i = 0; i < 6 && 6 * g + i < j.nb; i++) {
            const ext2 ri = e2_load(j.r, 6 * g + i);
            v = e2_mul(v, ((m >> i) & 1) ? ri : e2_sub(e2_one(), ri));
        }
        tab[g][m][0] = v.c0;
        tab[g][m][1] = v.c1;
    }
    mam_barrier();
    const int ng = j.nb > 0 ? (j.nb + 5) / 6 : 1;
    const ext2 ch = j.chp ? ext2{j.chp[0], j.chp[1]} : j.ch;
    const u64 ch7 = gl_mul7(ch.c1);
    // A wave takes chunks of 64 * st consecutive y (st steps of 64 lanes, st a power of two <= 16
    // chosen so that every wave of the grid has a chunk). Within a chunk a lane's y keeps its
    // low group (the lane) and its groups above the second (the chunk), so the lane adds
    // T1[group 1] * b over the chunk and multiplies the sum by T0[lane] times the chunk's high
    // factor once: at st = 16, about one product per input pair instead of four.
    const int lane = threadIdx.x & 63;
    const ext2 t0{tab[0][lane][0], tab[0][lane][1]};
    const size_t nwaves = (size_t)nblocks * (blockDim.x >> 6);
    const size_t wave = (size_t)bid * (blockDim.x >> 6) + (threadIdx.x >> 6);
    int lst = 0;  // log2 st
    while (lst < 4 && ((j.pairs >> (6 + lst + 1)) >= nwaves)) lst++;
    const int lc = 6 + lst;  // log2 of the chunk length
    const size_t chunks = (j.pairs + ((size_t)1 << lc) - 1) >> lc;
    ext2 s0 = e2_zero(), s1 = e2_zero();
    for (size_t c = wave; c < chunks; c += nwaves) {
        ext2 c0 = e2_zero(), c1 = e2_zero();
        for (int i = 0; i < (1 << lst); i++) {
            const size_t y = (c << lc) + ((size_t)i << 6) + lane;
            if (y >= j.pairs) break;
            ext2 b0, b1;
            const size_t Q = j.pairs;
            if (FOLD) {
                const size_t i0 = j.rev ? y : 4 * y, i1 = j.rev ? y + 2 * Q : 4 * y + 1;
                const size_t i2 = j.rev ? y + Q : 4 * y + 2, i3 = j.rev ? y + 3 * Q : 4 * y + 3;
                if (BH_BASE) {
                    const u64 a0 = gl_canon(j.bh[i0]), a1 = gl_canon(j.bh[i1]);
                    const u64 a2 = gl_canon(j.bh[i2]), a3 = gl_canon(j.bh[i3]);
                    b0 = e2_add(e2_base(a0), e2_scale(gl_sub(a1, a0), ch));
                    b1 = e2_add(e2_base(a2), e2_scale(gl_sub(a3, a2), ch));
                } else {
                    const ext2 a0 = e2_load(j.bh, i0), a1 = e2_load(j.bh, i1);
                    const ext2 a2 = e2_load(j.bh, i2), a3 = e2_load(j.bh, i3);
                    b0 = e2_add(a0, e2_mul_fixed(ch, ch7, e2_sub(a1, a0)));
                    b1 = e2_add(a2, e2_mul_fixed(ch, ch7, e2_sub(a3, a2)));
                }
                e2_store(j.bh_out, j.rev ? y : 2 * y, b0);
                e2_store(j.bh_out, j.rev ? y + Q : 2 * y + 1, b1);
            } else {
                const size_t i0 = j.rev ? y : 2 * y, i1 = j.rev ? y + Q : 2 * y + 1;
                b0 = BH_BASE ? e2_base(gl_canon(j.bh[i0])) : e2_load(j.bh, i0);
                b1 = BH_BASE ? e2_base(gl_canon(j.bh[i1])) : e2_load(j.bh, i1);
            }
            const size_t m1 = (y >> 6) & 63;
            const ext2 t1{tab[1][m1][0], tab[1][m1][1]};
            if (BH_BASE && !FOLD) {
                c0 = e2_add(c0, e2_scale(b0.c0, t1));
                c1 = e2_add(c1, e2_scale(b1.c0, t1));
            } else {
                c0 = e2_add(c0, e2_mul(t1, b0));
                c1 = e2_add(c1, e2_mul(t1, b1));
            }
        }
        ext2 H = t0;
        for (int g = 2; g < ng; g++) {
            const size_t m = ((c << lc) >> (6 * g)) & 63;
            H = e2_mul(H, ext2{tab[g][m][0], tab[g][m][1]});
        }
        s0 = e2_add(s0, e2_mul(H, c0));
        s1 = e2_add(s1, e2_mul(H, c1));
    }
    u64 acc[4] = {s0.c0, s0.c1, s1.c0, s1.c1};
    block_sum<4>(acc, lds);
    if (TICKET)
        grid_finish_n<4>(acc, lds, j.partials, j.counter, j.result, bid, nblocks);
    else
        grid_partials<4>(acc, j.partials, j.result);
}

