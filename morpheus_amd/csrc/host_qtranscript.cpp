// Host zk-torch-4 transcript (zk-torch-4/src/transcript.rs:21-240) and its C-ABI.
#include <string.h>

#include "monolith_q.h"
#include "qtranscript.h"

#define MAM_API extern "C" __attribute__((visibility("default")))

namespace mam {

static constexpr int QRATE = 8;

static void qduplex(mam_qtranscript *t) {
    for (int64_t i = 0; i < t->in_count; i++) t->state[i] = t->in_buf[i];
    for (int64_t i = t->in_count; i < QRATE; i++) t->state[i] = 0;
    t->in_count = 0;
    monoq::permute(t->state);
    for (int i = 0; i < QRATE; i++) t->out_buf[QRATE - 1 - i] = t->state[i];
    t->out_count = QRATE;
}

static void qabsorb(mam_qtranscript *t, uint64_t v) {
    t->out_count = 0;
    t->in_buf[t->in_count++] = fq::canon(v);
    if (t->in_count == QRATE) qduplex(t);
}

static uint64_t qsqueeze(mam_qtranscript *t) {
    if (t->out_count == 0) qduplex(t);
    t->out_count--;
    return t->out_buf[t->out_count];
}

static void qabsorb_label(mam_qtranscript *t, const uint8_t *label, long len) {
    if (len <= 0) return;
    qabsorb(t, (uint64_t)len);
    for (long i = 0; i < len; i++) qabsorb(t, label[i]);
}

void qtr_init(mam_qtranscript *t, const uint8_t *label, long len) {
    memset(t, 0, sizeof(*t));
    for (long i = 0; i < len; i++) qabsorb(t, label[i]);
}

void qtr_append_u64s(mam_qtranscript *t, const uint8_t *label, long len, const uint64_t *v, long n) {
    qabsorb_label(t, label, len);
    for (long i = 0; i < n; i++) qabsorb(t, v[i]);
}

void qtr_append_ext2(mam_qtranscript *t, const uint8_t *label, long len, const uint64_t *v2) {
    qtr_append_u64s(t, label, len, v2, 2);
}

void qtr_challenge_scalars(mam_qtranscript *t, const uint8_t *label, long len, uint64_t *out, long n) {
    qabsorb_label(t, label, len);
    for (long i = 0; i < n; i++) out[i] = qsqueeze(t);
}

}  // namespace mam

// C-ABI: the methods of zk-torch-4's Transcript (transcript.rs:36-240), labels as byte strings
// with explicit lengths. append_scalar(s) / append_u64(_slice) / append_digest all absorb the
// label and then the values reduced mod q, so one entry point serves them.
MAM_API void mam_qtranscript_init(mam_qtranscript *t, const uint8_t *label, long len) { mam::qtr_init(t, label, len); }
MAM_API void mam_qtranscript_append_u64s(mam_qtranscript *t, const uint8_t *label, long len, const uint64_t *v,
                                         long n) {
    mam::qtr_append_u64s(t, label, len, v, n);
}
MAM_API void mam_qtranscript_append_ext2(mam_qtranscript *t, const uint8_t *label, long len, const uint64_t *v2) {
    mam::qtr_append_ext2(t, label, len, v2);
}
// challenge_scalar (n = 1), challenge_ext2 (n = 2, c0 then c1), challenge_vector (n = len)
MAM_API void mam_qtranscript_challenge(mam_qtranscript *t, const uint8_t *label, long len, uint64_t *out, long n) {
    mam::qtr_challenge_scalars(t, label, len, out, n);
}
// fork(family, branch_id) (transcript.rs:166-171): *child = a copy of *t that absorbed the
// family label and branch_id
MAM_API void mam_qtranscript_fork(const mam_qtranscript *t, const uint8_t *family, long len, uint64_t branch_id,
                                  mam_qtranscript *child) {
    mam_qtranscript c = *t;
    mam::qtr_append_u64s(&c, family, len, &branch_id, 1);
    *child = c;
}
// fingerprint (transcript.rs:133-139): debug hash of the counters and the sponge state
MAM_API uint64_t mam_qtranscript_fingerprint(const mam_qtranscript *t) {
    uint64_t h = (uint64_t)(t->in_count * 100 + t->out_count);
    for (int i = 0; i < 12; i++) h = h * 0x9e3779b97f4a7c15ULL + t->state[i];
    return h;
}
// The Monolith-q permutation on its own, in place on 12 words.
MAM_API void mam_monolith_q_permute(uint64_t *state12) { monoq::permute(state12); }
