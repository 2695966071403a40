// Host Fiat-Shamir transcript (zk-torch-3/src/transcript.rs:167-330) and its C-ABI.
// Absorbed values are reduced to residues: the sponge output depends only on residues
// (the first permutation step adds a round constant and the S-box reduces), so this is
// the same transcript for any representation the caller passes.
#include <string.h>

#include "poseidon2_host.h"
#include "transcript.h"

#define MAM_API extern "C" __attribute__((visibility("default")))

namespace mam {

static void duplex(mam_transcript *t) {
    for (int64_t i = 0; i < t->in_count; i++) t->state[i] = t->in_buf[i];
    for (int64_t i = t->in_count; i < 4; i++) t->state[i] = 0;
    t->in_count = 0;
    p2h::permute(t->state);
    for (int i = 0; i < 4; i++) t->out_buf[3 - i] = t->state[i];
    t->out_count = 4;
}

static void absorb(mam_transcript *t, uint64_t v) {
    t->out_count = 0;
    t->in_buf[t->in_count++] = gl_canon(v);
    if (t->in_count == 4) duplex(t);
}

static uint64_t squeeze(mam_transcript *t) {
    if (t->out_count == 0) duplex(t);
    t->out_count--;
    return t->out_buf[t->out_count];
}

static void absorb_label(mam_transcript *t, const uint8_t *label, long len) {
    if (len <= 0) return;
    absorb(t, (uint64_t)len);
    for (long i = 0; i < len; i++) absorb(t, label[i]);
}

void tr_init(mam_transcript *t, const uint8_t *label, long len) {
    memset(t, 0, sizeof(*t));
    for (long i = 0; i < len; i++) absorb(t, label[i]);
}

void tr_append_u64(mam_transcript *t, const uint8_t *label, long len, uint64_t v) {
    absorb_label(t, label, len);
    absorb(t, v);
}

void tr_append_scalars(mam_transcript *t, const uint8_t *label, long len, const uint64_t *v, long n) {
    absorb_label(t, label, len);
    for (long i = 0; i < n; i++) absorb(t, v[i]);
}

void tr_append_ext2(mam_transcript *t, const uint8_t *label, long len, const uint64_t *v2) {
    absorb_label(t, label, len);
    absorb(t, v2[0]);
    absorb(t, v2[1]);
}

uint64_t tr_challenge_scalar(mam_transcript *t, const uint8_t *label, long len) {
    absorb_label(t, label, len);
    return squeeze(t);
}

void tr_challenge_ext2(mam_transcript *t, const uint8_t *label, long len, uint64_t *out2) {
    absorb_label(t, label, len);
    out2[0] = squeeze(t);
    out2[1] = squeeze(t);
}

}  // namespace mam

// C-ABI (labels as byte strings with explicit lengths)
MAM_API void mam_transcript_init(mam_transcript *t, const uint8_t *label, long len) { mam::tr_init(t, label, len); }
MAM_API void mam_transcript_append_u64(mam_transcript *t, const uint8_t *label, long len, uint64_t v) {
    mam::tr_append_u64(t, label, len, v);
}
MAM_API void mam_transcript_append_scalars(mam_transcript *t, const uint8_t *label, long len, const uint64_t *v,
                                           long n) {
    mam::tr_append_scalars(t, label, len, v, n);
}
MAM_API void mam_transcript_append_ext2(mam_transcript *t, const uint8_t *label, long len, const uint64_t *v2) {
    mam::tr_append_ext2(t, label, len, v2);
}
MAM_API uint64_t mam_transcript_challenge_scalar(mam_transcript *t, const uint8_t *label, long len) {
    return mam::tr_challenge_scalar(t, label, len);
}
MAM_API void mam_transcript_challenge_ext2(mam_transcript *t, const uint8_t *label, long len, uint64_t *out2) {
    mam::tr_challenge_ext2(t, label, len, out2);
}
// Host Poseidon2-w8 permutation (for checking the device one).
MAM_API void mam_poseidon2_permute_host(uint64_t *state8) { p2h::permute(state8); }
