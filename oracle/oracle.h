/*
 * oracle/oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * The oracle: a scalar CPU restatement of the reference prover path
 * (Vimalakirti/morpheus: zk-torch-3 transcript + CPU provers, goldilocks-cuda-rs
 * Basefold / Merkle semantics).  It is the checker for the HIP library in
 * morpheus_amd/ and the CPU baseline timed by bench.py.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links or calls it.  All outputs are canonical residues.
 *
 * Pinning status of each function is recorded next to its declaration.
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_HASH_MONOLITH 0
#define OR_HASH_POSEIDON2 1

/* ---- field (pinned: brute force vs Python big ints, algebraic identities) ---- */
uint64_t or_gl_add(uint64_t a, uint64_t b);
uint64_t or_gl_sub(uint64_t a, uint64_t b);
uint64_t or_gl_mul(uint64_t a, uint64_t b);
uint64_t or_gl_inv(uint64_t a);
void or_ext2_mul(const uint64_t *a, const uint64_t *b, uint64_t *out);
void or_ext2_inv(const uint64_t *a, uint64_t *out);

/* ---- hashes ---- */
/* Poseidon2 width 8.  initial_linear_layer=1 is the Plonky3 schedule (pinned by
 * the Plonky3 known-answer vector, cuda/poseidon2_kernels.cu:546-555);
 * initial_linear_layer=0 is the reference code's schedule
 * (zk-torch-3/src/transcript.rs:136-161), which shares every constant. */
void or_poseidon2_permute(uint64_t *state8, int initial_linear_layer);
/* Monolith width 12 (pinned: goldilocks-cuda-rs/src/cpu_monolith.rs:236-256) */
void or_monolith_permute(uint64_t *state12);
void or_monolith_hash_gl_leaf(uint64_t a, uint64_t b, uint64_t *out4);
void or_monolith_hash_ext2_leaf(const uint64_t *a2, const uint64_t *b2, uint64_t *out4);
void or_monolith_compress(const uint64_t *l4, const uint64_t *r4, uint64_t *out4);
void or_poseidon2_hash_gl_leaf(uint64_t a, uint64_t b, uint64_t *out4);
void or_poseidon2_hash_ext2_leaf(const uint64_t *a2, const uint64_t *b2, uint64_t *out4);
void or_poseidon2_compress(const uint64_t *l4, const uint64_t *r4, uint64_t *out4);

/* Merkle tree, layout [N leaves][N/2]...[root], 4 u64 per digest.
 * tree must hold (2N-1)*4 u64. */
void or_merkle_tree_gl(const uint64_t *cw, long num_leaves, int hash, uint64_t *tree);
void or_merkle_tree_ext2(const uint64_t *cw, long num_leaves, int hash, uint64_t *tree);
long or_merkle_root_offset(long num_leaves);
int or_merkle_verify_path(const uint64_t *leaf4, const uint64_t *path, int path_len, long index,
                          const uint64_t *root4, int hash);

/* ---- Fiat-Shamir transcript (zk-torch-3/src/transcript.rs:167-330) ---- */
typedef struct {
    uint64_t state[8];
    uint64_t in_buf[4];
    uint64_t out_buf[4];
    int64_t in_count;
    int64_t out_count;
} or_transcript;

void or_tr_new(or_transcript *t, const uint8_t *label, long len);
void or_tr_append_u64(or_transcript *t, const uint8_t *label, long len, uint64_t v);
void or_tr_append_scalars(or_transcript *t, const uint8_t *label, long len, const uint64_t *v, long n);
void or_tr_append_ext2(or_transcript *t, const uint8_t *label, long len, const uint64_t *v2);
uint64_t or_tr_challenge_scalar(or_transcript *t, const uint8_t *label, long len);
void or_tr_challenge_ext2(or_transcript *t, const uint8_t *label, long len, uint64_t *out2);

/* ---- eq / Lagrange tables, LSB-first (zk-torch-3/src/poly/mod.rs:106-163) ---- */
void or_eq_table(const uint64_t *r, int nv, uint64_t *out);
void or_eq_table_ext2(const uint64_t *r2, int nv, uint64_t *out2);
/* MLE evaluation by repeated LSB-first folding (closed-form check for partial eval) */
void or_partial_eval_ext2(const uint64_t *in2, int nv, const uint64_t *r2, int m, uint64_t *out2);

/* ---- linear sumcheck over Ext2 (zk-torch-3/src/sumcheck/cpu_ext2_prover.rs:21-148) ----
 * polys: np polynomials, each 2^nv Ext2 values interleaved [c0,c1]; consumed.
 * Writes round messages [nv][np+1][2], challenges [nv][2], final evals [np][2],
 * final_eval [2] (product of final evals). */
void or_sumcheck_prove_ext2(or_transcript *t, uint64_t *polys, int np, int nv,
                            uint64_t *round_msgs, uint64_t *challenges,
                            uint64_t *final_evals, uint64_t *final_eval);
/* zk-torch-3/src/sumcheck/verifier.rs:17-65 */
int or_sumcheck_verify_ext2(or_transcript *t, const uint64_t *claim2, int np, int nv,
                            const uint64_t *round_msgs, const uint64_t *final_eval);

/* ---- Basefold (goldilocks-cuda-rs/src/basefold.rs, zk-torch-3/src/commit/cpu_basefold.rs) ---- */
/* number of table entries = sum over levels of 2^(nv+lr-i-1) */
long or_basefold_table_len(int nv, int lr, int nrounds);
/* BasefoldTable::generate (basefold.rs:281-321); entries as [x0, w] pairs */
void or_basefold_table_generate(int nv, int lr, int nrounds, uint64_t seed, uint64_t *entries);
/* commit (wrapper.cu:1242-1312 + merkle.rs:34-59): codeword 2^(nv+lr) (Type1),
 * bh_evals 2^nv (Type1), tree (2N-1)*4 with N = 2^(nv+lr-1) leaves */
void or_basefold_commit(const uint64_t *evals, int nv, int lr, int hash,
                        uint64_t *codeword, uint64_t *bh_evals, uint64_t *tree);
/* flat proof length in u64 for (nv, lr, nq) */
long or_basefold_proof_len(int nv, int lr, int nq);
/* cpu_full_open_ext2 (cpu_basefold.rs:235-425); table entries as generated for
 * (table_nv, lr, table_nv).  Writes the flat proof layout. */
void or_basefold_open_ext2(const uint64_t *codeword, const uint64_t *bh_evals, const uint64_t *root4,
                           int nv, int lr, int hash, const uint64_t *point2,
                           const uint64_t *table, int table_nv, or_transcript *t, int nq,
                           uint64_t *proof);
void or_basefold_open_ext2_tree(const uint64_t *codeword, const uint64_t *bh_evals, const uint64_t *root4,
                                const uint64_t *tree0, int nv, int lr, int hash, const uint64_t *point2,
                                const uint64_t *table, int table_nv, or_transcript *t, int nq, uint64_t *proof);
void or_set_threads(int n);
int or_get_threads(void);
/* batch_open_ext2 / batch_verify_ext2 (basefold.rs:1202-1525, 1931-2221): num_comms
 * commitments (codewords, bh_evals, roots[4]) of the same nv, lr opened at num_points Ext2
 * points (points[num_points][nv][2]) for num_evals claims (poly, point, value[2]). */
long or_basefold_batch_proof_len(int nv, int lr, int nq, int num_evals);
void or_basefold_batch_open_ext2(int num_comms, const uint64_t *const *codewords, const uint64_t *const *bh_list,
                                 const uint64_t *roots, int nv, int lr, int hash, int num_points,
                                 const uint64_t *points, int num_evals, const int *eval_poly, const int *eval_point,
                                 const uint64_t *eval_values, const uint64_t *table, int table_nv, or_transcript *t,
                                 int nq, uint64_t *proof);
int or_basefold_batch_verify_ext2(int num_comms, const uint64_t *roots, int nv, int lr, int hash, int num_points,
                                  const uint64_t *points, int num_evals, const int *eval_poly, const int *eval_point,
                                  const uint64_t *eval_values, const uint64_t *proof, int nq, const uint64_t *table,
                                  int table_nv, or_transcript *t);
/* BasefoldVerifier::verify_ext2 (basefold.rs:1759-1925); returns 1 on accept */
int or_basefold_verify_ext2(const uint64_t *root4, const uint64_t *point2, int nv, int lr, int hash,
                            const uint64_t *proof, int nq, const uint64_t *table, int table_nv,
                            or_transcript *t);

/* ---- ChaCha and the int8-matrix Ajtai commitment (oracle/ajtai.c) ----
 * ChaCha block: pinned by the RFC 8439 section 2.3.2 ChaCha20 vector.
 * Commitment: exact integer matrix-vector product (checked against Python big ints). */
void or_chacha_block(int rounds, const uint32_t *key8, uint32_t counter, const uint32_t *nonce3, uint32_t *out16);
void or_ajtai_i8_row(const uint8_t *key32, uint32_t row, uint64_t n, int8_t *out);
void or_ajtai_i8_commit(const uint8_t *key32, int rows, uint64_t n, const uint64_t *z, int batch, uint64_t *out);

/* ---- almost-Goldilocks q = 2^64 - 2^32 - 31 (oracle/q.c) ----
 * field/Ext2: pinned by the reference's golden vectors (almost_field_test.cu:143-180,
 * almost_extension_test.cu:107-126) and algebraic identities checked against Python ints.
 * eq / partial eval / sumcheck: closed forms (sum of eq = 1, MLE evaluation, g(0)+g(1)). */
uint64_t or_q_add(uint64_t a, uint64_t b);
uint64_t or_q_sub(uint64_t a, uint64_t b);
uint64_t or_q_mul(uint64_t a, uint64_t b);
uint64_t or_q_inv(uint64_t a);
void or_q2_mul(const uint64_t *a, const uint64_t *b, uint64_t *out);
void or_q2_inv(const uint64_t *a, uint64_t *out);
void or_q_eq_table(int ext, const uint64_t *r, int nv, uint64_t *out);
void or_q_partial_eval(const uint64_t *in, int in_ext, int nv, const uint64_t *r, int r_ext, int m, uint64_t *out);
/* zk-torch-4 transcript (Monolith-w12 over q): the round constants of rounds 1..5 (60 words)
 * are carried in the struct */
typedef struct {
    uint64_t st[12], in[8], out[8];
    int64_t nin, nout;
    uint64_t rc[60];
} or_qtr;
void or_monoq_permute(uint64_t *x, const uint64_t *rc60);
void or_qtr_init(or_qtr *t, const uint8_t *label, long len, const uint64_t *rc60);
void or_qtr_absorb(or_qtr *t, uint64_t v);
void or_qtr_append(or_qtr *t, const uint8_t *label, long len, const uint64_t *v, long n);
void or_qtr_challenge(or_qtr *t, const uint8_t *label, long len, uint64_t *out, long n);
void or_q_sumcheck_prove(or_qtr *t, int ext, const uint64_t *polys, int np, int nv, uint64_t *msgs, uint64_t *chs,
                         uint64_t *finals, uint64_t *fe);
void or_q_sumcheck_msg(int ext, const uint64_t *polys, int d, uint64_t original_size, uint64_t half, uint64_t *out);
void or_q_sumcheck_fold(int ext, const uint64_t *in, uint64_t *out, const uint64_t *ch2, int d,
                        uint64_t original_size, uint64_t half);
/* ring Ajtai over F_q[X]/(X^64+1), KAPPA = 42 (ajtai_cpu_reference.cuh:58-242).
 * Pinned by the ring identities (negacyclic X^64 = -1, linearity, binary = sparse = wide
 * on the same witness) and the RFC-pinned ChaCha block; no reference vector exists. */
void or_ring_matrix_entry(const uint32_t *key8, uint32_t row, uint64_t j, uint64_t *out64);
void or_ring_commit_binary(const uint32_t *key8, const uint64_t *z, uint64_t N, int B, uint64_t col_offset,
                           uint64_t *out);
void or_ring_commit_wide(const uint32_t *key8, const uint64_t *z_wide, uint64_t N, uint64_t col_offset,
                         uint64_t *out);
void or_ring_commit_sparse(const uint32_t *key8, const uint64_t *pos, uint64_t K, uint64_t *out);
void or_ring_fold_witness(const uint64_t *z1, const uint64_t *z2, const int8_t *r, uint64_t N, uint64_t *out);
void or_ring_fold_commitment(const uint64_t *c1, const uint64_t *c2, const int8_t *r, uint64_t *out);
void or_ring_multifold_mixed_witness(const uint64_t *bin, const uint64_t *pos, const uint64_t *neg, const int8_t *r_all,
                                     int K, int T, uint64_t N, int16_t *out);
void or_ring_multifold_commitment(const uint64_t *c_all, const int8_t *r_all, int M, uint64_t *out);
void or_ring_split_witness(const int16_t *z_wide, uint64_t N, uint64_t *pos, uint64_t *neg);
void or_ring_commit_ternary(const uint32_t *key8, const uint64_t *pos, const uint64_t *neg, uint64_t N, uint64_t *out);
void or_ring_materialize(const uint32_t *key8, uint64_t N, uint64_t *m);
void or_q_selective_add_planes(const uint64_t *eq2, const uint64_t *planes, uint64_t total, int n_planes,
                               uint64_t packed, uint64_t *out2);
/* fused permute + partial eval over modulus p or q (pinned by: identity LUTs reduce it to
 * the partial evaluation, which is pinned by direct MLE sums) */
void or_fused_permute_peval(uint64_t mod, const uint64_t *evals, const uint64_t *eq, const uint32_t *lo,
                            const uint32_t *hi, int n, int m, int half, long out_size, uint64_t *out);

#ifdef __cplusplus
}
#endif
#endif
