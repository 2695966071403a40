/*
 * morpheus_prover.h -- prover-level C-ABI of libmorpheus_amd.so.
 *
 * In the reference these operations are Rust methods of goldilocks-cuda-rs and
 * zk-torch-3 layered over the raw FFI; this library implements them natively (C++ host
 * code over HIP kernels) and exports them as C functions, so the Rust types become thin
 * shells over these calls. Each declaration cites the Rust interface it replaces.
 *
 * Conventions: d_* device pointers, h_* / unprefixed pointers host memory; Ext2 values
 * are [c0, c1] word pairs; returns 0 / -1 (runtime error) / -2 (rejected argument) unless
 * stated; every function completes before returning.
 */
#ifndef MORPHEUS_PROVER_H
#define MORPHEUS_PROVER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Fiat-Shamir transcript: zk-torch-3/src/transcript.rs:167-330 (Transcript) ---- */
typedef struct {
    uint64_t state[8];
    uint64_t in_buf[4];
    uint64_t out_buf[4];
    int64_t in_count;
    int64_t out_count;
} mam_transcript;

void mam_transcript_init(mam_transcript *t, const uint8_t *label, long len);        /* Transcript::new, :182 */
void mam_transcript_append_u64(mam_transcript *t, const uint8_t *label, long len,
                               uint64_t v);                                         /* append_u64, :225 */
void mam_transcript_append_scalars(mam_transcript *t, const uint8_t *label, long len, const uint64_t *v,
                                   long n);                                         /* append_scalars, :217 */
void mam_transcript_append_ext2(mam_transcript *t, const uint8_t *label, long len,
                                const uint64_t *v2);                                /* append_ext2, :255 */
uint64_t mam_transcript_challenge_scalar(mam_transcript *t, const uint8_t *label,
                                         long len);                                 /* challenge_scalar, :231 */
void mam_transcript_challenge_ext2(mam_transcript *t, const uint8_t *label, long len,
                                   uint64_t *out2);                                 /* challenge_ext2, :247 */
void mam_poseidon2_permute_host(uint64_t *state8);                                  /* poseidon2_permute, :136 */
void mam_monolith_permute_host(uint64_t *state12);       /* goldilocks-cuda-rs/src/cpu_monolith.rs:141 */

/* ---- the zk-torch-3 transcript on the device (no reference counterpart: the reference's
 * device challengers, challenger_*, take no labels) ----
 * d_state: a mam_transcript (18 words) in device memory. The program (u64 words) is a list of
 * [kind, label_len, label bytes packed little-endian into ceil(label_len / 8) words, n]:
 * kind 0 appends the next n words of d_src (append_u64 / append_scalars / append_ext2 of
 * transcript.rs:217-260), kind 1 draws n challenges into the next n words of d_dst
 * (challenge_scalar / challenge_ext2, :231-253). One wavefront runs it, queued on the
 * calling thread's stream; *d_err (optional) becomes 1 for a malformed program or an
 * overrun of n_src / n_dst. */
int mam_dtranscript_run(uint64_t *d_state, const uint64_t *d_prog, long prog_words, const uint64_t *d_src,
                        long n_src, uint64_t *d_dst, long n_dst, int *d_err);
/* n Poseidon2-w8 permutations of 8-word states in place, 8 lanes per state (synchronizes). */
int mam_poseidon2_wide_permute(uint64_t *d_states, long n);

/* ---- zk-torch-4 transcript: Monolith-w12 over q = 2^64 - 2^32 - 31, rate 8, capacity 4
 * (zk-torch-4/src/transcript.rs:21-240, crypto/monolith.rs:33-213). Plain data: a copy is a
 * clone. append_u64s serves append_scalar(s) / append_u64(_slice) / append_digest (label, then
 * the values reduced mod q); challenge serves challenge_scalar (n = 1), challenge_ext2 (n = 2)
 * and challenge_vector. */
typedef struct {
    uint64_t state[12];
    uint64_t in_buf[8];
    uint64_t out_buf[8];
    int64_t in_count;
    int64_t out_count;
} mam_qtranscript;

void mam_qtranscript_init(mam_qtranscript *t, const uint8_t *label, long len);      /* Transcript::new, :36 */
void mam_qtranscript_append_u64s(mam_qtranscript *t, const uint8_t *label, long len, const uint64_t *v,
                                 long n);                                           /* append_*, :67-110 */
void mam_qtranscript_append_ext2(mam_qtranscript *t, const uint8_t *label, long len,
                                 const uint64_t *v2);                               /* append_ext2, :99 */
void mam_qtranscript_challenge(mam_qtranscript *t, const uint8_t *label, long len, uint64_t *out,
                               long n);                                             /* challenge_*, :112-130 */
void mam_qtranscript_fork(const mam_qtranscript *t, const uint8_t *family, long len, uint64_t branch_id,
                          mam_qtranscript *child);                                  /* fork, :166 */
uint64_t mam_qtranscript_fingerprint(const mam_qtranscript *t);                     /* fingerprint, :133 */
void mam_monolith_q_permute(uint64_t *state12);                                     /* monolith.rs:194 */
/* Whole linear sumcheck over the q Ext2 (X^2 - 3) with that transcript:
 * zk-torch-4/src/sumcheck/gpu_prover.rs:26-112 (GpuLinearSumcheckProver::new + prove /
 * prove_ext2), the proofs of cpu_ext2_prover.rs:23-100. Arguments and outputs as
 * mam_sumcheck_prove below; num_polys <= 8. */
int mam_q_sumcheck_prove(mam_qtranscript *t, const uint64_t *d_polys, int num_polys, int num_vars,
                         int input_is_base, uint64_t *round_msgs, uint64_t *challenges, uint64_t *final_evals,
                         uint64_t *final_eval);

/* ---- linear sumcheck over Ext2 ----
 * Whole proof: zk-torch-3/src/sumcheck/gpu_prover.rs:21-89 (GpuLinearSumcheckProver::new
 * + prove / prove_ext2). d_polys holds num_polys polynomials of 2^num_vars elements back
 * to back, base words (input_is_base = 1) or interleaved Ext2; it is not modified.
 * Outputs: round_msgs[num_vars][num_polys + 1][2], challenges[num_vars][2],
 * final_evals[num_polys][2], final_eval[2]. num_polys <= 8. */
int mam_sumcheck_prove(mam_transcript *t, const uint64_t *d_polys, int num_polys, int num_vars, int input_is_base,
                       uint64_t *round_msgs, uint64_t *challenges, uint64_t *final_evals, uint64_t *final_eval);
/* GpuSumcheckStateExt2 (goldilocks-cuda-rs/src/sumcheck_prover.rs:205-423):
 * new / new_from_base (:218, :262), compute_round_message (:325), fold (:373),
 * final_evaluations (:401). The state reads d_polys until its first fold is applied. */
int mam_sc_state_new(const uint64_t *d_polys, int num_polys, int num_vars, int input_is_base, void **handle);
int mam_sc_state_round_message(void *handle, uint64_t *msg_out /* (num_polys + 1) * 2 */);
int mam_sc_state_fold(void *handle, uint64_t challenge_c0, uint64_t challenge_c1);
int mam_sc_state_final_evaluations(void *handle, uint64_t *out /* num_polys * 2 */);
int mam_sc_state_free(void *handle);

/* ---- Basefold PCS: goldilocks-cuda-rs/src/basefold.rs ---- */
#define MAM_HASH_MONOLITH 0  /* default Merkle hash (goldilocks-cuda-rs/Cargo.toml:22-24) */
#define MAM_HASH_POSEIDON2 1
/* BasefoldTable::generate (:281-321): entries as (point, weight) pairs, host memory */
long mam_basefold_table_len(int num_vars, int log_rate, int num_rounds);
int mam_basefold_table_generate(int num_vars, int log_rate, int num_rounds, uint64_t seed, uint64_t *entries);
/* BasefoldCommitment::commit_device (:382-414); handle owns its device buffers */
int mam_basefold_commit(const uint64_t *d_evals, int num_vars, int log_rate, int hash, void **handle,
                        uint64_t *root_out /* 4 */);
/* Same into caller buffers: codeword 2^(nv+lr), bh_evals 2^nv, tree (2^(nv+lr) - 1) * 4 words */
int mam_basefold_commit_into(const uint64_t *d_evals, int num_vars, int log_rate, int hash, uint64_t *d_codeword,
                             uint64_t *d_bh_evals, uint64_t *d_tree, uint64_t *root_out);
/* HostCommitmentCache re-upload (:873-900): wrap caller-owned buffers as a commitment */
int mam_basefold_commitment_wrap(uint64_t *d_codeword, uint64_t *d_bh_evals, uint64_t *d_tree,
                                 const uint64_t *root4, int num_vars, int log_rate, int hash, void **handle);
int mam_basefold_commitment_buffers(void *handle, uint64_t **d_codeword, uint64_t **d_bh_evals, uint64_t **d_tree,
                                    uint64_t *root_out);
int mam_basefold_commitment_free(void *handle);
/* Free the calling thread's cached scratch (re-created on demand). */
int mam_release_scratch(void);
/* No reference counterpart: page-locked host memory for outputs written again and again (a
 * proof buffer). Device-to-host copies into it are direct; into pageable memory they pass
 * through a staging buffer and a host copy (about 40 us per 2^22 proof's 0.9 MB of queries).
 * For callers that reuse their output buffers. */
int mam_host_alloc(size_t bytes, void **p);
int mam_host_free(void *p);
/* BasefoldCommitment::open_ext2 (:578-863) with extract_ext2_queries (:2310-2402).
 * point2: num_vars Ext2 (host); d_table: device table generated for table_nv >= num_vars;
 * proof: mam_basefold_proof_len() words in the flat layout
 *   eval[2] | oracles[nv][3][2] | folded_roots[nv-1][4] | final_cw[2^lr][2] |
 *   queries x { index | values[nv][2][2] | initial path[nv+lr-1][4] | folded paths } */
long mam_basefold_proof_len(int num_vars, int log_rate, int num_queries);
int mam_basefold_open_ext2(void *handle, const uint64_t *point2, const uint64_t *d_table, int table_nv,
                           mam_transcript *t, int num_queries, uint64_t *proof);
/* BasefoldVerifier::verify_ext2 (:1759-1925); host only; returns 1 accept, 0 reject */
int mam_basefold_verify_ext2(const uint64_t *root4, const uint64_t *point2, int num_vars, int log_rate, int hash,
                             const uint64_t *proof, int num_queries, const uint64_t *table, int table_nv,
                             mam_transcript *t);

/* ---- batched proving of uneven instances (benchmark configuration 4) ----
 * No single reference function: each instance runs the reference's product sumcheck
 * (zk-torch-3/src/sumcheck/gpu_prover.rs:21-89) on T = Transcript::new(label), then
 * commits f (basefold.rs:382-414) and opens it at the sumcheck challenges on the same T
 * (basefold.rs:578-863). `workers` persistent host threads, each on its own stream, take
 * instances largest first; workers = 1 is one-at-a-time proving. Inputs on the device,
 * outputs on the host; start/end times in ms from the batch start. num_vars >= 2. */
typedef struct {
    const uint64_t *d_fg;  /* f then g, 2^num_vars base words each (device) */
    int num_vars;
    const uint8_t *label;
    long label_len;
    uint64_t *round_msgs;  /* [num_vars][3][2] */
    uint64_t *challenges;  /* [num_vars][2] */
    uint64_t *final_evals; /* [2][2] */
    uint64_t *root;        /* [4] */
    uint64_t *proof;       /* mam_basefold_proof_len(num_vars, log_rate, num_queries) words */
    double start_ms, end_ms;
    int status;
} mam_batch_instance;
int mam_prove_batch(mam_batch_instance *inst, int n, const uint64_t *d_table, int table_nv, int log_rate, int hash,
                    int num_queries, int workers);
/* Graph form, the same proofs word for word: one hipGraph launch runs every instance's
 * sumcheck with its transcript on the device (one launch pair per two rounds for the whole
 * batch) and every commitment (on `branches` side streams, 0..4; 0: after the sumchecks),
 * then `workers` threads open the commitments, each transcript continuing from the device.
 * max_wg (1..4096) caps the workgroups of one instance in one pass. The graph and its
 * buffers are kept for the next call with the same instance buffers and settings;
 * mam_batch_graph_release frees them. num_vars + log_rate <= 30. */
int mam_prove_batch_graph(mam_batch_instance *inst, int n, const uint64_t *d_table, int table_nv, int log_rate,
                          int hash, int num_queries, int workers, int branches, int max_wg);
int mam_batch_graph_release(void);
/* Wall time (ms) of the last graph-form call: out2[0] the graph phase, out2[1] the openings. */
int mam_batch_graph_phase_ms(double *out2);

/* ---- proof wire format: bincode 1.x bytes of zk-torch-3's SumcheckProof
 * (zk-torch-3/src/sumcheck/mod.rs:16-23): final_eval (c0, c1 as u64 LE) | u64 rounds |
 * per round: u64 entries | entries x (c0, c1). Inputs in mam_sumcheck_prove's output layout.
 * Returns the byte length, -2 on bad arguments or too small a buffer. */
long mam_sumcheck_proof_bincode_len(int rounds, int entries);
long mam_sumcheck_proof_bincode(const uint64_t *round_msgs, int rounds, int entries, const uint64_t *final_eval,
                                uint8_t *out, long cap);

/* ---- batch opening of several commitments at several Ext2 points ----
 * batch_open_ext2 / BasefoldVerifier::batch_verify_ext2 (goldilocks-cuda-rs/src/basefold.rs:
 * 1202-1525, 1931-2221): commitments of equal num_vars, log_rate and hash; points (host)
 * [num_points][nv][2]; claims (host) name a commitment and a point. Flat proof (u64):
 *   outer_oracles[nv][3][2] | combined_root[4] | inner_eval[2] | inner_oracles[nv][3][2] |
 *   folded_roots[nv-1][4] | final_cw[2^lr][2] |
 *   nq x { index | values[nv][2][2] (Ext2 at level 0) | paths as in the single opening } |
 *   nq x num_evals x { values[2] | path[nv+lr-1][4] }
 * The verifier takes the host table and returns 1 on accept. */
typedef struct {
    int poly, point;
    uint64_t value[2];
} mam_eval_claim;
long mam_basefold_batch_proof_len(int num_vars, int log_rate, int num_queries, int num_evals);
int mam_basefold_batch_open_ext2(void *const *handles, int num_comms, const uint64_t *points, int num_points,
                                 const mam_eval_claim *evals, int num_evals, const uint64_t *d_table, int table_nv,
                                 mam_transcript *t, int num_queries, uint64_t *proof);
int mam_basefold_batch_verify_ext2(const uint64_t *roots, int num_comms, int num_vars, int log_rate, int hash,
                                   const uint64_t *points, int num_points, const mam_eval_claim *evals,
                                   int num_evals, const uint64_t *proof, int num_queries, const uint64_t *table,
                                   int table_nv, mam_transcript *t);

/* ---- matmul proof C = A B (benchmark configuration 5) ----
 * Composed from the einsum proof of zk-torch-3/src/basicblock/einsum.rs:807-1098 for
 * "ij,jk->ik": Basefold commitments of A (m x k), B (k x n), C (m x n) (row-major, m = 2^lm,
 * ...), an optional int8 Ajtai commitment of C, the claim C(r) at a transcript point, one
 * sumcheck over the inner index of [A(r_i, .), B(., r_j), 1], and Basefold openings of A, B, C
 * at their points, all on T = Transcript::new(label). The step list is in host_matmul.hip.
 * The Ajtai commitment needs |C| = 2^(lm+ln) >= 64 (a multiple of 64 matrix columns);
 * smaller C with a non-null ajtai returns -2. */
typedef struct {
    const uint64_t *d_a, *d_b, *d_c; /* device, canonical words, row-major */
    int lm, lk, ln;
    const uint8_t *label;
    long label_len;
    uint64_t *roots;       /* [3][4] */
    uint64_t *ajtai;       /* [ajtai_rows] or NULL */
    uint64_t *point;       /* [ln + lm][2]: r_j then r_i */
    uint64_t *claim;       /* [2] */
    uint64_t *round_msgs;  /* [lk][4][2] */
    uint64_t *challenges;  /* [lk][2] */
    uint64_t *final_evals; /* [3][2] */
    uint64_t *proof_a, *proof_b, *proof_c;
    double start_ms, end_ms;
    int status;
} mam_matmul_instance;
int mam_matmul_prove(mam_matmul_instance *inst, const uint64_t *d_table, int table_nv, int log_rate, int hash,
                     int num_queries, void *ajtai_handle, int ajtai_rows);

/* ---- proving service for matmul proofs (benchmark configuration 5) ----
 * Requests are released at their arrival offsets (ms from the call), admitted in arrival
 * order while the sum of admitted requests' device footprints fits hbm_budget (a request
 * larger than the budget runs alone), and proved by `workers` persistent threads on their
 * own streams. admit_ms / done_ms are filled in; latency = done_ms - arrival_ms. No
 * reference counterpart (the reference proves layers one at a time). */
typedef struct {
    mam_matmul_instance *inst;
    double arrival_ms;
    double admit_ms, done_ms;
    int status;
} mam_serve_request;
int mam_serve_matmul(mam_serve_request *req, int n, const uint64_t *d_table, int table_nv, int log_rate, int hash,
                     int num_queries, void *ajtai_handle, int ajtai_rows, int workers, uint64_t hbm_budget);

/* ---- int8-matrix Ajtai commitment on INT8 MFMA (benchmark configuration 3) ----
 * No reference function has these semantics (the reference's Ajtai is the ring-SIS commit
 * of almost-goldilocks-cuda-rs/src/ajtai.rs:138-386; its INT8 paths are a multifold GEMM,
 * cuda_almost_goldilocks/ajtai_tc.cuh:166-225). Defined here:
 *   M_ij = (int8) byte (j mod 64) of ChaCha8(key, counter = j / 64, nonce = (i, 0, 0)),
 *   out[w * rows + i] = sum_j M_ij z_w[j]  (mod p).
 * setup expands M once (rows * n bytes of HBM); rows a multiple of 16 up to 256, n a
 * multiple of 64. commit: batch <= 32 witnesses of n words each (device), out device. */
int mam_ajtai_i8_setup(const uint8_t *key32, int rows, uint64_t n, void **handle);
int mam_ajtai_i8_commit(void *handle, const uint64_t *d_z, int batch, uint64_t *d_out);
/* The same against the first n_used columns (witnesses of n_used words; n_used a multiple of
 * 64, <= n; batch <= 16 when n_used < n): equal to a commitment with an n_used-column set-up. */
int mam_ajtai_i8_commit_prefix(void *handle, const uint64_t *d_z, int batch, uint64_t n_used, uint64_t *d_out);
/* Allocation-free twin of mam_ajtai_i8_commit_prefix. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The chunk partials live in a file-scope device array
 * of 16 MiB per GPU (held until the process ends; its address is resolved at set-up), so the
 * call makes no hipMalloc / hipFree and no allocating library call. Returns once the work is
 * queued on the calling thread's stream: d_out is complete when that stream is synchronized.
 * Returns -2 for the arguments mam_ajtai_i8_commit_prefix rejects and when the partials would
 * exceed 16 MiB: (n_used / 2^13) x rows x 16 ceil((8 batch + [n_used < n]) / 16) x 4 bytes for
 * up to 32 planes (2^14 columns per chunk above); at rows = 256 and n_used = 2^22 that admits
 * batch <= 8 (a prefix commit, with its all-ones plane, batch <= 7). Every call on one GPU shares that GPU's array: calls that overlap on one GPU race
 * (different threads or streams); calls queued on one stream do not. */
int mam_ajtai_i8_commit_noalloc(void *handle, const uint64_t *d_z, int batch, uint64_t n_used, uint64_t *d_out);
int mam_ajtai_i8_matrix_row(void *handle, int row, uint64_t col0, uint64_t count, int8_t *h_out);
int mam_ajtai_i8_free(void *handle);

/* ---- device primitives without a reference symbol of their own ---- */
/* Monolith permutation of n 12-word states in place (the kernel behind the reference's
 * monolith_merkle_tree_* builders, cuda/monolith.cuh:94-216) */
int mam_monolith_permute_batch(uint64_t *d_states, int n);
/* Merkle build on the calling thread's stream without synchronizing (ffi.rs:549-576 semantics) */
int mam_merkle_tree_async(const uint64_t *d_codeword, uint64_t *d_tree, size_t num_leaves, int hash,
                          int ext2_leaves);

/* Opt-in timing of kernel groups with HIP events on their launch stream (no reference
 * counterpart; used by bench.py). Categories: 0 Merkle builds (work = permutations),
 * 1 opening fold+oracle (bytes), 2 commit transform (bytes), 3 sumcheck fold+message
 * (bytes), 4 single leaf-hashing launches over base-field codewords (permutations),
 * 5 int8 Ajtai GEMM launches (bytes), 6 dense ring-Ajtai launches (ring columns x witnesses).
 * collect() fills 7 entries of each array. enable(1) times every category, enable_mask
 * only the categories whose bit is set (each timed scope records two events). */
int mam_profile_enable(int on);
int mam_profile_enable_mask(unsigned mask);
/* VALU issue-rate microbenchmark (no reference counterpart; used by bench.py for the
 * measured VALU peak): kind 0 = v_xad_u32 chains, 1 = v_mad_u64_u32, 2 = v_lshl_add_u64,
 * 3 = v_add_u32; each thread of `blocks` x 256 runs iters x 8 single-instruction steps.
 * *ms = the timed launch. */
int mam_valu_probe(int kind, int blocks, int iters, double *ms);
/* No reference counterpart: median launch time (ms) of a plain device-to-device copy of
 * `bytes` (a multiple of 4096), the best of three grid shapes: the bandwidth a streaming
 * kernel of that size can reach, launch and ramp included. */
int mam_copy_probe(size_t bytes, int reps, double *ms);
int mam_profile_collect(double *ms, long *count, double *work);

#ifdef __cplusplus
}
#endif
#endif /* MORPHEUS_PROVER_H */
