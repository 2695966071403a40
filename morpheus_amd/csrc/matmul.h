// The matmul proof statement shared by the prover (host_matmul.hip) and the proving service
// (host_serve.hip); the C-ABI declares the same layout (include/morpheus_prover.h).
#pragma once

#include <stdint.h>

#include "field.h"

struct mam_matmul_instance {
    const u64 *d_a, *d_b, *d_c;  // device, canonical words, row-major
    int lm, lk, ln;
    const uint8_t *label;
    long label_len;
    u64 *roots;        // [3][4] A, B, C
    u64 *ajtai;        // [ajtai_rows], or null to skip the Ajtai commitment
    u64 *point;        // [ln + lm][2]: r_j then r_i
    u64 *claim;        // [2]
    u64 *round_msgs;   // [lk][4][2]
    u64 *challenges;   // [lk][2]
    u64 *final_evals;  // [3][2]
    u64 *proof_a, *proof_b, *proof_c;  // Basefold proofs (mam_basefold_proof_len words each)
    double start_ms, end_ms;
    int status;
};

extern "C" int mam_matmul_prove(mam_matmul_instance *inst, const u64 *d_table, int table_nv, int log_rate, int hash,
                                int num_queries, void *ajtai_handle, int ajtai_rows);
