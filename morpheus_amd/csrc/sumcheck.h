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
// adds the workgroup partials when there is one; *posted tells whether it did.
int mam_sc_launch_pass(int np, const u64 *in, size_t in_stride, bool base_in, int folds, bool look, u64 *out,
                       size_t out_stride, size_t items, ext2 ca, ext2 cb, const ScScratch &s, hipStream_t st,
                       const MailJob *mj = nullptr, bool *posted = nullptr);
int mam_sc_launch_pass_final(int np, const u64 *in, size_t in_stride, bool base_in, int folds, ext2 ca, ext2 cb,
                             u64 *finals, hipStream_t st);
