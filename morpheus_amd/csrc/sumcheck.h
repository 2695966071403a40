// Internal interface of the fused sumcheck kernels (k_sumcheck.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "field.h"
#include "fieldq.h"

struct MailJob;

#define MAM_MAX_POLYS 8
// Upper bound on workgroups of a grid-wide reduction launch (sizes the partials buffer).
#define MAM_RED_MAX_BLOCKS 2048

struct ScScratch {
    u64 *partials;  // MAM_RED_MAX_BLOCKS * 2 * (MAM_MAX_POLYS + 1) words
    u32 *counter;   // one zero-initialized ticket counter
    u64 *msg;       // 2 * (MAM_MAX_POLYS + 1) words: the reduced message
};

int mam_sc_launch_msg(int np, const u64 *in, size_t stride, size_t half, bool base, const ScScratch &s,
                      hipStream_t st);
int mam_sc_launch_fold_msg(int np, const u64 *in, size_t in_stride, u64 *out, size_t out_stride, size_t half_out,
                           bool base_in, ext2 ch, const ScScratch &s, hipStream_t st);
int mam_sc_launch_fold_final(int np, const u64 *in, size_t in_stride, bool base_in, ext2 ch, u64 *finals,
                             hipStream_t st);
// The same three launches over the almost-Goldilocks field q and its Ext2 (X^2 - 3).
int mam_qsc_launch_msg(int np, const u64 *in, size_t stride, size_t half, bool base, const ScScratch &s,
                       hipStream_t st);
int mam_qsc_launch_fold_msg(int np, const u64 *in, size_t in_stride, u64 *out, size_t out_stride, size_t half_out,
                            bool base_in, qe2 ch, const ScScratch &s, hipStream_t st);
int mam_qsc_launch_fold_final(int np, const u64 *in, size_t in_stride, bool base_in, qe2 ch, u64 *finals,
                              hipStream_t st);
// Two rounds per pass (whole-proof prover, np <= MAM_LOOK_MAX_POLYS): see k_sumcheck.hip.
#define MAM_LOOK_MAX_POLYS 3
// mj (may be null): a post of the pass's words to that host mailbox, made by the launch that
// adds the workgroup partials when there is one; *posted tells whether it did. d_ch (may be
// null): device words ca.c0, ca.c1, cb.c0, cb.c1 that replace ca and cb (a pre-queued pass
// whose challenges arrive through a gate).
int mam_sc_launch_pass(int np, const u64 *in, size_t in_stride, bool base_in, int folds, bool look, u64 *out,
                       size_t out_stride, size_t items, ext2 ca, ext2 cb, const ScScratch &s, hipStream_t st,
                       const MailJob *mj = nullptr, bool *posted = nullptr, const u64 *d_ch = nullptr);
int mam_sc_launch_pass_final(int np, const u64 *in, size_t in_stride, bool base_in, int folds, ext2 ca, ext2 cb,
                             u64 *finals, hipStream_t st, const u64 *d_ch = nullptr);

// ---------------------------------------------------------------------------
// Batched product sumcheck with the transcript on the device (k_sumcheck.hip): every
// instance of a batch (two base polynomials f, g of its own size) advances one pass (two
// rounds, or the last single round) per launch pair, so a batch is a fixed launch sequence
// that can be captured in a graph.
// ---------------------------------------------------------------------------
#define MAM_BSC_WORDS 24  // reduced words per workgroup slot: msg (3 Ext2) + sym (3 x 3 Ext2)

// One instance's part of one pass.
struct BscJob {
    const u64 *in;          // arrays the pass reads (base f | g, or Ext2 after folds)
    u64 *out;               // arrays after the pending folds (FOLDS = 2), else null
    size_t in_stride, out_stride, items;
    int base_in, folds, look;
    int inst;               // instance index (transcript, challenge slot, outputs)
    int round;              // first round of the pass
    int last;               // the instance's last pass: the final values follow it
    unsigned blk0, nblk;    // the job's workgroups in the pass launch
    const u64 *fin_src;     // arrays the final fold reads
    size_t fin_stride;
    int fin_base, fin_folds;
};

// Device memory of one instance: transcript (mam_transcript, 18 words), the pending
// challenges (ca, cb: 4 words), outputs (round messages [nv][3][2], challenges [nv][2],
// final values [2][2]).
struct BscInst {
    u64 *tr;
    u64 *ch;
    u64 *msgs, *chs, *finals;
};

// pass k of a batch: jobs[0..njobs) sorted by blk0, nblk_total workgroups; partials holds
// MAM_BSC_WORDS words per workgroup.
int mam_bsc_launch_pass(const BscJob *d_jobs, int njobs, unsigned nblk_total, const BscInst *d_insts, u64 *partials,
                        hipStream_t st);
int mam_bsc_launch_round(const BscJob *d_jobs, int njobs, const BscInst *d_insts, const u64 *partials,
                         hipStream_t st);

