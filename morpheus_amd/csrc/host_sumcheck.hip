// Linear sumcheck prover over Ext2 on the device, driven by the host transcript.
//
// Mirrors goldilocks-cuda-rs/src/sumcheck_prover.rs:205-423 (GpuSumcheckStateExt2:
// new / new_from_base / compute_round_message / fold / final_evaluations) and the proof
// loop of zk-torch-3/src/sumcheck/gpu_prover.rs:21-89 (labels "num_var", "num_poly",
// "round_message", "challenge"), which produces the same transcript as the CPU prover
// zk-torch-3/src/sumcheck/cpu_ext2_prover.rs:21-148.
//
// The state folds lazily: fold(challenge) records the challenge, and the next
// round-message request runs the fused fold + message kernel, so each round streams the
// polynomials once. The input buffer is never written.
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "common.h"
#include "sumcheck.h"
#include "transcript.h"

namespace {

struct ScState {
    int np = 0, nv = 0;
    size_t n = 0;              // elements per polynomial at round 0
    const u64 *input = nullptr;
    bool input_base = false;
    size_t in_stride = 0;      // words between polynomials of the input
    u64 *buf[2] = {nullptr, nullptr};
    size_t stride[2] = {0, 0};
    int cur = -1;              // -1: current arrays are the input; else buf[cur]
    size_t cur_size = 0;       // elements per polynomial now
    bool pending = false;      // a fold challenge is recorded but not applied
    ext2 pending_ch = {0, 0};
    bool finished = false;
    ScScratch s = {nullptr, nullptr, nullptr};
    u64 *finals = nullptr;     // device, np Ext2
    u64 *host = nullptr;       // pinned readback, 2 * (MAM_MAX_POLYS + 1) words
    hipStream_t st = nullptr;
    bool arena = false;        // buffers carved from the thread's scratch arena
};

static size_t align_words(size_t w) { return (w + 31) & ~(size_t)31; }

int state_alloc(ScState &S, bool arena) {
    hipStream_t st = S.st;
    size_t half = S.n / 2;
    // buf[0] holds rounds 1, 3, 5, ... (n/2 Ext2 max); buf[1] rounds 2, 4, ... (n/4 Ext2 max)
    S.stride[0] = 2 * half;
    S.stride[1] = half > 1 ? half : 2;
    if (arena) {
        const size_t w0 = align_words(S.np * S.stride[0]), w1 = align_words(S.np * S.stride[1]);
        const size_t wp = align_words(MAM_RED_MAX_BLOCKS * 2 * (MAM_MAX_POLYS + 1)), wc = 32,
                     wm = align_words(2 * (MAM_MAX_POLYS + 1)), wf = align_words(2 * MAM_MAX_POLYS);
        u64 *base = mam_dev_arena(w0 + w1 + wp + wc + wm + wf, 1);
        S.host = mam_host_arena(4 * (MAM_MAX_POLYS + 1));
        if (!base || !S.host) return -1;
        S.arena = true;
        S.buf[0] = base;
        S.buf[1] = base + w0;
        S.s.partials = base + w0 + w1;
        S.s.counter = (u32 *)(base + w0 + w1 + wp);
        S.s.msg = base + w0 + w1 + wp + wc;
        S.finals = S.s.msg + wm;
        return hipMemsetAsync(S.s.counter, 0, 64, st) == hipSuccess ? 0 : -1;
    }
    mam_stream_pool_setup();
    if (hipMallocAsync((void **)&S.buf[0], S.np * S.stride[0] * sizeof(u64), st) != hipSuccess) return -1;
    if (hipMallocAsync((void **)&S.buf[1], S.np * S.stride[1] * sizeof(u64), st) != hipSuccess) return -1;
    if (hipMallocAsync((void **)&S.s.partials, MAM_RED_MAX_BLOCKS * 2 * (MAM_MAX_POLYS + 1) * sizeof(u64), st) !=
        hipSuccess)
        return -1;
    if (hipMallocAsync((void **)&S.s.counter, 64, st) != hipSuccess) return -1;
    if (hipMallocAsync((void **)&S.s.msg, 2 * (MAM_MAX_POLYS + 1) * sizeof(u64), st) != hipSuccess) return -1;
    if (hipMallocAsync((void **)&S.finals, 2 * MAM_MAX_POLYS * sizeof(u64), st) != hipSuccess) return -1;
    if (hipMemsetAsync(S.s.counter, 0, 64, st) != hipSuccess) return -1;
    if (hipHostMalloc((void **)&S.host, 4 * (MAM_MAX_POLYS + 1) * sizeof(u64), hipHostMallocDefault) != hipSuccess)
        return -1;
    return 0;
}

void state_free(ScState &S) {
    hipStream_t st = S.st;
    if (S.arena) {
        S = ScState();
        return;
    }
    if (S.buf[0]) hipFreeAsync(S.buf[0], st);
    if (S.buf[1]) hipFreeAsync(S.buf[1], st);
    if (S.s.partials) hipFreeAsync(S.s.partials, st);
    if (S.s.counter) hipFreeAsync(S.s.counter, st);
    if (S.s.msg) hipFreeAsync(S.s.msg, st);
    if (S.finals) hipFreeAsync(S.finals, st);
    if (S.host) hipHostFree(S.host);
    S = ScState();
}

int state_init(ScState &S, const u64 *d_polys, int np, int nv, bool base, hipStream_t st, bool arena) {
    if (np < 1 || np > MAM_MAX_POLYS || nv < 1 || nv > 34 || !d_polys) return -2;
    S.np = np;
    S.nv = nv;
    S.n = (size_t)1 << nv;
    S.input = d_polys;
    S.input_base = base;
    S.in_stride = base ? S.n : 2 * S.n;
    S.cur = -1;
    S.cur_size = S.n;
    S.st = st;
    return state_alloc(S, arena);
}

const u64 *cur_ptr(const ScState &S, size_t &stride, bool &base) {
    if (S.cur < 0) {
        stride = S.in_stride;
        base = S.input_base;
        return S.input;
    }
    stride = S.stride[S.cur];
    base = false;
    return S.buf[S.cur];
}

// Compute the current round's message into S.host[0 .. 2(np+1)).
int state_round_message(ScState &S) {
    if (S.finished || S.cur_size < 2) return -2;
    size_t stride;
    bool base;
    const u64 *in = cur_ptr(S, stride, base);
    int rc;
    if (S.pending) {
        size_t half_out = S.cur_size / 4;
        if (half_out == 0) return -2;  // fold to size 1 is handled by final_evaluations
        int nxt = S.cur < 0 ? 0 : 1 - S.cur;
        rc = mam_sc_launch_fold_msg(S.np, in, stride, S.buf[nxt], S.stride[nxt], half_out, base, S.pending_ch, S.s,
                                    S.st);
        if (rc) return rc;
        S.cur = nxt;
        S.cur_size /= 2;
        S.pending = false;
    } else {
        rc = mam_sc_launch_msg(S.np, in, stride, S.cur_size / 2, base, S.s, S.st);
        if (rc) return rc;
    }
    return mam_publish_wait(S.s.msg, 2 * (S.np + 1), nullptr, 0, S.host, S.st);
}

int state_fold(ScState &S, ext2 ch) {
    if (S.finished || S.pending || S.cur_size < 2) return -2;
    S.pending = true;
    S.pending_ch = ext2{gl_canon(ch.c0), gl_canon(ch.c1)};
    return 0;
}

// Apply the last recorded fold (size 2 -> 1) and read the final values into S.host.
int state_final(ScState &S) {
    if (!S.finished) {
        if (!S.pending || S.cur_size != 2) return -2;
        size_t stride;
        bool base;
        const u64 *in = cur_ptr(S, stride, base);
        int rc = mam_sc_launch_fold_final(S.np, in, stride, base, S.pending_ch, S.finals, S.st);
        if (rc) return rc;
        S.pending = false;
        S.cur_size = 1;
        S.finished = true;
    }
    return mam_publish_wait(S.finals, 2 * S.np, nullptr, 0, S.host, S.st);
}

// Whole proof with two rounds per device pass (np <= MAM_LOOK_MAX_POLYS): each pass
// returns the round-r message and the round-(r+1) message as a polynomial in ch_r, so the
// host draws both challenges from one readback (same transcript as one round at a time).
int prove_lookahead(ScState &S, mam_transcript *t, const uint8_t *l_msg, long l_msg_len, const uint8_t *l_ch,
                    long l_ch_len, u64 *round_msgs, u64 *challenges, u64 *final_evals) {
    const int np = S.np, nv = S.nv, dp1 = np + 1;
    // The passes are fixed by nv: pass k covers rounds r_k, r_k + 1 (look) or r_k alone.
    struct Pass {
        int r;
        bool look;
        int folds;
        size_t items;
        const u64 *src;
        size_t src_stride;
        bool src_base;
        u64 *out;
        size_t out_stride;
    };
    std::vector<Pass> passes;
    const u64 *src = S.input;
    size_t src_stride = S.in_stride, size = S.n;
    bool src_base = S.input_base;
    int folds = 0, nxt = 0;
    for (int r = 0; r < nv;) {
        Pass p;
        p.r = r;
        p.look = nv - r >= 2;
        p.folds = folds;
        const size_t m = size >> folds;  // elements per polynomial after the pending folds
        p.items = m / (p.look ? 4 : 2);
        p.src = src;
        p.src_stride = src_stride;
        p.src_base = src_base;
        p.out = folds ? S.buf[nxt] : nullptr;
        p.out_stride = folds ? S.stride[nxt] : 0;
        passes.push_back(p);
        if (folds) {
            src = p.out;
            src_stride = p.out_stride;
            src_base = false;
            size = m;
            nxt ^= 1;
        }
        folds = p.look ? 2 : 1;
        r += p.look ? 2 : 1;
    }
    const int P = (int)passes.size();
    auto words = [&](const Pass &p) { return 2 * dp1 * (p.look ? 1 + dp1 : 1); };
    // labelled: the label of the round's first message is already absorbed (the host does it
    // while the device computes the message; the absorbed sequence is the same)
    auto draw = [&](const u64 *m, int r, bool labelled = false) {
        for (int c = 0; c < dp1; c++) {
            const bool skip = labelled && c == 0;
            mam::tr_append_ext2(t, skip ? nullptr : l_msg, skip ? 0 : l_msg_len, m + 2 * c);
        }
        u64 ch[2];
        mam::tr_challenge_ext2(t, l_ch, l_ch_len, ch);
        challenges[2 * r] = ch[0];
        challenges[2 * r + 1] = ch[1];
        return ext2{ch[0], ch[1]};
    };
    ext2 ca = e2_zero(), cb = e2_zero();
    int rc = 0;
    // Pass k + 1 (or the final values) is queued behind a gate before pass k's words are
    // read, and released with its challenges once they are drawn (common.h).
    const bool pre = mam_prequeue_enabled();
    const Pass &last = passes[P - 1];
    const u64 *fsrc = last.folds ? last.out : last.src;
    const size_t fstride = last.folds ? last.out_stride : last.src_stride;
    const bool fbase = last.folds ? false : last.src_base;
    struct Queued {
        u64 gate = 0, seq = 0;
        u64 *box = nullptr;
        bool posted = false;
    };
    // queue pass k (k == P: the final values) with its challenges from device memory
    auto queue = [&](int k, Queued &q) -> int {
        const u64 *d_ch = mam_gate_queue(S.st, 4, &q.gate);
        if (!d_ch) return -1;
        if (k == P)
            return mam_sc_launch_pass_final(np, fsrc, fstride, fbase, last.look ? 2 : 1, ca, cb, S.finals, S.st, d_ch);
        const Pass &p = passes[k];
        q.box = mam_mailbox_arm(&q.seq);
        if (!q.box) return -1;
        const MailJob mail = {nullptr, 0, q.box, q.seq};
        int r = mam_sc_launch_pass(np, p.src, p.src_stride, p.src_base, p.folds, p.look, p.out, p.out_stride, p.items,
                                   ca, cb, S.s, S.st, &mail, &q.posted, d_ch);
        if (!r && !q.posted) r = mam_mailbox_post(S.s.msg, words(p), nullptr, 0, q.box, q.seq, S.st);
        return r;
    };
    Queued next;
    bool have_next = false;
    for (int k = 0; k < P && !rc; k++) {
        const Pass &p = passes[k];
        u64 seq = 0;
        u64 *box = nullptr;
        if (have_next) {
            box = next.box;
            seq = next.seq;
            have_next = false;
        } else {
            box = mam_mailbox_arm(&seq);
            if (!box) rc = -1;
            const MailJob mail = {nullptr, 0, box, seq};
            bool posted = false;
            if (!rc)
                rc = mam_sc_launch_pass(np, p.src, p.src_stride, p.src_base, p.folds, p.look, p.out, p.out_stride,
                                        p.items, ca, cb, S.s, S.st, &mail, &posted);
            if (!rc && !posted) rc = mam_mailbox_post(S.s.msg, words(p), nullptr, 0, box, seq, S.st);
        }
        Queued q;
        if (!rc && pre) {
            if (queue(k + 1, q)) {
                mam_gate_open_all();
                rc = -1;
            }
        }
        // the first message's label does not depend on the message: absorb it (its full
        // blocks run the permutation) before waiting for the pass
        if (!rc && l_msg_len > 0) mam::tr_append_scalars(t, l_msg, l_msg_len, nullptr, 0);
        if (!rc) rc = mam_mailbox_wait(box, seq, words(p), S.host, S.st);
        if (rc) {
            mam_gate_open_all();
            break;
        }
        u64 *m0 = round_msgs + (size_t)p.r * dp1 * 2;
        memcpy(m0, S.host, dp1 * 2 * sizeof(u64));
        ca = draw(m0, p.r, l_msg_len > 0);
        if (p.look) {
            // round r + 1: msg(c) = sum_j ch_r^j sym[c][j] (Horner)
            u64 *m1 = round_msgs + (size_t)(p.r + 1) * dp1 * 2;
            const u64 *sym = S.host + 2 * dp1;
            for (int c = 0; c < dp1; c++) {
                ext2 v{gl_canon(sym[2 * (c * dp1 + np)]), gl_canon(sym[2 * (c * dp1 + np) + 1])};
                for (int j = np - 1; j >= 0; j--)
                    v = e2_add(e2_mul(v, ca), ext2{gl_canon(sym[2 * (c * dp1 + j)]), gl_canon(sym[2 * (c * dp1 + j) + 1])});
                m1[2 * c] = v.c0;
                m1[2 * c + 1] = v.c1;
            }
            cb = draw(m1, p.r + 1);
        }
        if (pre) {
            const u64 w[4] = {ca.c0, ca.c1, cb.c0, cb.c1};
            mam_gate_open(q.gate, w, 4);
            if (k + 1 < P) {
                next = q;
                have_next = true;
            }
        }
    }
    if (rc) {
        S.finished = true;
        return rc;
    }
    // the final values (already queued behind the last gate when pre-queuing)
    if (!pre) rc = mam_sc_launch_pass_final(np, fsrc, fstride, fbase, last.look ? 2 : 1, ca, cb, S.finals, S.st);
    if (!rc) rc = mam_publish_wait(S.finals, 2 * np, nullptr, 0, S.host, S.st);
    if (!rc) memcpy(final_evals, S.host, 2 * np * sizeof(u64));
    S.finished = true;
    return rc;
}

const uint8_t L_NUM_VAR[] = "num_var";
const uint8_t L_NUM_POLY[] = "num_poly";
const uint8_t L_ROUND_MSG[] = "round_message";
const uint8_t L_CHALLENGE[] = "challenge";

}  // namespace

// ---------------------------------------------------------------------------
// Whole proof: GpuLinearSumcheckProver::new + prove (gpu_prover.rs:21-89)
// d_polys: num_polys polynomials of 2^num_vars elements, packed back to back, base-field
// words (input_is_base = 1) or interleaved Ext2. Outputs (host): round_msgs
// [num_vars][num_polys + 1][2], challenges [num_vars][2], final_evals [num_polys][2],
// final_eval [2] (product of the final evaluations).
// ---------------------------------------------------------------------------
extern "C" __attribute__((visibility("default"))) int mam_sumcheck_prove(
    mam_transcript *t, const u64 *d_polys, int num_polys, int num_vars, int input_is_base, u64 *round_msgs,
    u64 *challenges, u64 *final_evals, u64 *final_eval) {
    ScState S;
    int rc = state_init(S, d_polys, num_polys, num_vars, input_is_base != 0, mam_stream(), true);
    if (rc) {
        state_free(S);
        return rc;
    }
    mam::tr_append_u64(t, L_NUM_VAR, 7, (u64)num_vars);
    mam::tr_append_u64(t, L_NUM_POLY, 8, (u64)num_polys);
    const int dp1 = num_polys + 1;
    if (num_polys <= MAM_LOOK_MAX_POLYS) {
        rc = prove_lookahead(S, t, L_ROUND_MSG, 13, L_CHALLENGE, 9, round_msgs, challenges, final_evals);
        if (!rc) {
            ext2 fe = e2_one();
            for (int i = 0; i < num_polys; i++) fe = e2_mul(fe, ext2{final_evals[2 * i], final_evals[2 * i + 1]});
            final_eval[0] = fe.c0;
            final_eval[1] = fe.c1;
        }
        state_free(S);
        return rc;
    }
    for (int r = 0; r < num_vars && !rc; r++) {
        rc = state_round_message(S);
        if (rc) break;
        u64 *m = round_msgs + (size_t)r * dp1 * 2;
        memcpy(m, S.host, dp1 * 2 * sizeof(u64));
        for (int c = 0; c < dp1; c++) mam::tr_append_ext2(t, L_ROUND_MSG, 13, m + 2 * c);
        u64 ch[2];
        mam::tr_challenge_ext2(t, L_CHALLENGE, 9, ch);
        challenges[2 * r] = ch[0];
        challenges[2 * r + 1] = ch[1];
        rc = state_fold(S, ext2{ch[0], ch[1]});
    }
    if (!rc) rc = state_final(S);
    if (!rc) {
        ext2 fe = e2_one();
        for (int i = 0; i < num_polys; i++) {
            final_evals[2 * i] = S.host[2 * i];
            final_evals[2 * i + 1] = S.host[2 * i + 1];
            fe = e2_mul(fe, ext2{S.host[2 * i], S.host[2 * i + 1]});
        }
        final_eval[0] = fe.c0;
        final_eval[1] = fe.c1;
    }
    state_free(S);
    return rc;
}

// ---------------------------------------------------------------------------
// Stateful interface (GpuSumcheckStateExt2). The state reads d_polys (device, not copied)
// until the first fold has been applied; the caller keeps it alive until then.
// ---------------------------------------------------------------------------
extern "C" __attribute__((visibility("default"))) int mam_sc_state_new(const u64 *d_polys, int num_polys, int num_vars,
                                                                      int input_is_base, void **handle) {
    if (!handle) return -2;
    ScState *S = new ScState();
    int rc = state_init(*S, d_polys, num_polys, num_vars, input_is_base != 0, mam_stream(), false);
    if (rc) {
        state_free(*S);
        delete S;
        *handle = nullptr;
        return rc;
    }
    *handle = S;
    return 0;
}

extern "C" __attribute__((visibility("default"))) int mam_sc_state_round_message(void *handle, u64 *msg_out) {
    ScState *S = (ScState *)handle;
    if (!S) return -2;
    int rc = state_round_message(*S);
    if (!rc) memcpy(msg_out, S->host, 2 * (S->np + 1) * sizeof(u64));
    return rc;
}

extern "C" __attribute__((visibility("default"))) int mam_sc_state_fold(void *handle, u64 c0, u64 c1) {
    ScState *S = (ScState *)handle;
    if (!S) return -2;
    return state_fold(*S, ext2{c0, c1});
}

extern "C" __attribute__((visibility("default"))) int mam_sc_state_final_evaluations(void *handle, u64 *out) {
    ScState *S = (ScState *)handle;
    if (!S) return -2;
    int rc = state_final(*S);
    if (!rc) memcpy(out, S->host, 2 * S->np * sizeof(u64));
    return rc;
}

extern "C" __attribute__((visibility("default"))) int mam_sc_state_free(void *handle) {
    ScState *S = (ScState *)handle;
    if (!S) return 0;
    state_free(*S);
    delete S;
    return hipStreamSynchronize(mam_stream()) == hipSuccess ? 0 : -1;
}
