// Fiat-Shamir transcript of the zk-torch-3 prover: a Poseidon2-w8 duplex sponge in
// overwrite mode, rate 4, capacity 4 (zk-torch-3/src/transcript.rs:167-330).
//
//   new(label):        absorb every label byte as a field element, no length prefix
//   append_*(label,.): absorb len(label) and the label bytes (nothing for an empty label),
//                      then the values
//   challenge_*(label):the same label absorption, then squeeze
//   absorb:            clears pending outputs; a full rate buffer triggers a duplex
//   duplex:            overwrite state[0..k) with the k buffered inputs, zero state[k..4),
//                      permute, and queue state[0..4) so squeezes pop state[0], state[1], ...
//   The Basefold opening uses empty labels (transcript.rs:332-350).
// The struct is plain data so callers in any language can hold it by value.
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t state[8];
    uint64_t in_buf[4];
    uint64_t out_buf[4];
    int64_t in_count;
    int64_t out_count;
} mam_transcript;

#ifdef __cplusplus
}

namespace mam {

void tr_init(mam_transcript *t, const uint8_t *label, long len);
void tr_append_u64(mam_transcript *t, const uint8_t *label, long len, uint64_t v);
void tr_append_scalars(mam_transcript *t, const uint8_t *label, long len, const uint64_t *v, long n);
void tr_append_ext2(mam_transcript *t, const uint8_t *label, long len, const uint64_t *v2);
uint64_t tr_challenge_scalar(mam_transcript *t, const uint8_t *label, long len);
void tr_challenge_ext2(mam_transcript *t, const uint8_t *label, long len, uint64_t *out2);

}  // namespace mam
#endif
