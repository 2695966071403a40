// Fiat-Shamir transcript of the zk-torch-4 prover (zk-torch-4/src/transcript.rs:21-240): a
// Monolith-w12 duplex sponge over q in overwrite mode, rate 8, capacity 4.
//
//   new(label):         absorb every label byte as a field element, no length prefix
//   append_*(label, .): absorb len(label) and the label bytes (nothing for an empty label),
//                       then the values reduced mod q
//   challenge_*(label): the same label absorption, then squeeze
//   fork(family, id):   a copy that absorbs the family label and then id
//   absorb:             clears pending outputs; a full rate buffer triggers a duplex
//   duplex:             overwrite state[0..k) with the k buffered inputs, zero state[k..8),
//                       permute, and queue state[0..8) so squeezes pop state[0], state[1], ...
// The struct is plain data: a copy is a clone.
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t state[12];
    uint64_t in_buf[8];
    uint64_t out_buf[8];
    int64_t in_count;
    int64_t out_count;
} mam_qtranscript;

#ifdef __cplusplus
}

namespace mam {

void qtr_init(mam_qtranscript *t, const uint8_t *label, long len);
void qtr_append_u64s(mam_qtranscript *t, const uint8_t *label, long len, const uint64_t *v, long n);
void qtr_append_ext2(mam_qtranscript *t, const uint8_t *label, long len, const uint64_t *v2);
void qtr_challenge_scalars(mam_qtranscript *t, const uint8_t *label, long len, uint64_t *out, long n);

}  // namespace mam
#endif
