// Proof wire format: bincode 1.x bytes of zk-torch-3's SumcheckProof
// (zk-torch-3/src/sumcheck/mod.rs:16-23, #[derive(Serialize)]), as bincode::serialize
// writes them (zk-torch-4/src/lib.rs:147-161 measures and fingerprints proofs this way):
// fixed-width little-endian integers, u64 length before every Vec, structs as their fields
// in declaration order, newtypes as their content.
//
//   SumcheckProof { final_eval: GoldilocksExt2, round_messages: Vec<Vec<GoldilocksExt2>> }
//   GoldilocksExt2 { c0: GoldilocksField, c1: GoldilocksField }, GoldilocksField(u64)
//
// bytes = c0 | c1 of final_eval (8 + 8) | u64 rounds | per round: u64 entries | entries x 16
// Field words are written as canonical residues (the values the Rust prover holds).
#include <stdint.h>
#include <string.h>

#include "field.h"

#define MAM_API extern "C" __attribute__((visibility("default")))

namespace {
inline uint8_t *put_u64(uint8_t *p, uint64_t v) {
    for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
    return p + 8;
}
}  // namespace

// Bytes of the serialized proof for `rounds` round messages of `entries` Ext2 values each.
MAM_API long mam_sumcheck_proof_bincode_len(int rounds, int entries) {
    if (rounds < 0 || entries < 0) return -2;
    return 16 + 8 + (long)rounds * (8 + 16L * entries);
}

// round_msgs: [rounds][entries][2] words, final_eval: [2] words (the layout of
// mam_sumcheck_prove's outputs). Writes mam_sumcheck_proof_bincode_len bytes to out (cap
// bytes available); returns the length, or -2 when the arguments or the capacity do not fit.
MAM_API long mam_sumcheck_proof_bincode(const uint64_t *round_msgs, int rounds, int entries,
                                        const uint64_t *final_eval, uint8_t *out, long cap) {
    const long len = mam_sumcheck_proof_bincode_len(rounds, entries);
    if (len < 0 || !final_eval || !out || (rounds > 0 && entries > 0 && !round_msgs) || cap < len) return -2;
    uint8_t *p = out;
    p = put_u64(p, gl_canon(final_eval[0]));
    p = put_u64(p, gl_canon(final_eval[1]));
    p = put_u64(p, (uint64_t)rounds);
    for (int r = 0; r < rounds; r++) {
        p = put_u64(p, (uint64_t)entries);
        for (int e = 0; e < entries; e++) {
            const uint64_t *v = round_msgs + 2 * ((size_t)r * entries + e);
            p = put_u64(p, gl_canon(v[0]));
            p = put_u64(p, gl_canon(v[1]));
        }
    }
    return (long)(p - out);
}
