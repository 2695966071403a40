/*
 * morpheus_agl.h -- the almost-Goldilocks (q = 2^64 - 2^32 - 31) hot-path subset of
 * libmorpheus_amd.so, under the names and signatures of the reference's
 * almost-goldilocks-cuda-rs/src/ffi.rs (static library `almost_goldilocks_cuda_wrapper`),
 * so that crate's Rust wrappers link against this library unchanged.
 *
 * Conventions as in the reference: device pointers to u64 field elements (canonical on
 * output), Ext2 = F_q[X]/(X^2 - 3) interleaved [c0, c1]; return 0 on success, -1 on a
 * runtime error or rejected argument, -2 unsupported batch size B, -3 unsupported chunk
 * size (wrapper.cu:1132-1205); calls complete before returning unless named `_stream`.
 * The memory/device entry points (cuda_malloc, cuda_memcpy_*, ...) are shared with the
 * Goldilocks crate and declared in morpheus_gl.h.
 */
#ifndef MORPHEUS_AGL_H
#define MORPHEUS_AGL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int almost_goldilocks_cuda_init(void); /* ffi.rs:13 */

/* memory and streams of the q crate beyond the shared ones in morpheus_gl.h (ffi.rs:61-100,
 * 342-350, 687). cuda_malloc / cuda_free cache freed blocks per device (best fit within
 * 1.25 x + 2 MiB, event-guarded reuse; AGL_ALLOC_CACHE=0 turns the cache off), as the
 * q crate's allocator does (wrapper.cu:118-248). */
size_t agl_cached_bytes(void);                                           /* ffi.rs:61 */
int agl_d2h_staged(void *h_dst, const void *d_src, size_t bytes);        /* ffi.rs:63 */
int agl_h2d_staged(void *d_dst, const void *h_src, size_t bytes);        /* ffi.rs:64 */
const char *agl_cuda_error_string(int code);                            /* ffi.rs:80 */
int cuda_memset(void *dst, int value, size_t size);                     /* ffi.rs:84 */
int cuda_pool_trim(size_t min_bytes_to_keep);                           /* ffi.rs:85 */
int cuda_sync_mine(void);                                               /* ffi.rs:87; AGL_DEVICE_SYNC=1: device-wide */
int cuda_stream_create(void **stream);                                  /* ffi.rs:96 */
int cuda_stream_destroy(void *stream);                                  /* ffi.rs:97 */
int cuda_stream_synchronize(void *stream);                              /* ffi.rs:98 */
int cuda_memcpy_dtod_async(void *dst, const void *src, size_t size, void *stream); /* ffi.rs:99 */
int cuda_memcpy_htod_async(void *dst, const void *src, size_t size, void *stream); /* ffi.rs:100 */
int cuda_memcpy_peer(void *dst, int dst_dev, const void *src, int src_dev, size_t size); /* ffi.rs:342-348 */
int cuda_enable_peer_access(int peer);                                  /* ffi.rs:350 */
int agl_zero_buffer_ffi(uint64_t *d_buf, size_t n);                     /* ffi.rs:687; returns the runtime error code */

/* base field batch ops, ffi.rs:103-112 */
int agl_batch_add_ffi(const uint64_t *a, const uint64_t *b, uint64_t *r, int n);
int agl_batch_sub_ffi(const uint64_t *a, const uint64_t *b, uint64_t *r, int n);
int agl_batch_mul_ffi(const uint64_t *a, const uint64_t *b, uint64_t *r, int n);
int agl_batch_neg_ffi(const uint64_t *a, uint64_t *r, int n);
int agl_batch_square_ffi(const uint64_t *a, uint64_t *r, int n);
int agl_batch_double_ffi(const uint64_t *a, uint64_t *r, int n);
int agl_batch_inverse_ffi(const uint64_t *a, uint64_t *r, int n);
int agl_batch_exp_ffi(const uint64_t *a, uint64_t exp, uint64_t *r, int n);
int agl_batch_mul_scalar_ffi(uint64_t scalar, const uint64_t *a, uint64_t *r, int n);
int agl_batch_div_ffi(const uint64_t *a, const uint64_t *b, uint64_t *r, int n);
/* out[x] = in[y], bit b of x = bit inv_map[b] of y (ffi.rs:113-119), total = 2^n_bits */
int agl_bit_permute_ffi(const uint64_t *d_input, uint64_t *d_output, const int *d_perm_map, int n_bits, int total);

/* Ext2 batch ops, ffi.rs:122-138 */
int aext2_batch_add_ffi(const uint64_t *a, const uint64_t *b, uint64_t *r, int n);
int aext2_batch_sub_ffi(const uint64_t *a, const uint64_t *b, uint64_t *r, int n);
int aext2_batch_mul_ffi(const uint64_t *a, const uint64_t *b, uint64_t *r, int n);
int aext2_batch_inverse_ffi(const uint64_t *a, uint64_t *r, int n);
int aext2_batch_neg_ffi(const uint64_t *a, uint64_t *r, int n);
int aext2_batch_square_ffi(const uint64_t *a, uint64_t *r, int n);
int aext2_batch_frobenius_ffi(const uint64_t *a, uint64_t *r, int n);
int aext2_batch_conjugate_ffi(const uint64_t *a, uint64_t *r, int n);
int aext2_batch_exp_ffi(const uint64_t *a, uint64_t exp, uint64_t *r, int n);
int aext2_batch_mul_scalar_ffi(uint64_t scalar, const uint64_t *a, uint64_t *r, int n);
int aext2_scale_accumulate_ffi(uint64_t scalar_c0, uint64_t scalar_c1, const uint64_t *src, uint64_t *acc, int n);
int agl_to_aext2_batch_ffi(const uint64_t *input, uint64_t *output, int n);
int aext2_to_agl_batch_ffi(const uint64_t *input, uint64_t *output, int n);

/* eq tables, ffi.rs:161-196: the table lands in d_buf_a for even log_n, d_buf_b for odd
 * (where the reference's ping-pong ends) and *d_result points at it. Batched: point l at
 * d_r_all + l * log_n Ext2, table l at buffer + l * leaf_stride Ext2 elements. */
int agl_eq_dp_all_ffi(const uint64_t *d_r, uint64_t *d_buf_a, uint64_t *d_buf_b, int log_n, uint64_t **d_result);
int aext2_eq_dp_all_ffi(const uint64_t *d_r, uint64_t *d_buf_a, uint64_t *d_buf_b, int log_n, uint64_t **d_result);
int aext2_eq_dp_all_stream_ffi(const uint64_t *d_r, uint64_t *d_buf_a, uint64_t *d_buf_b, int log_n,
                               uint64_t **d_result, void *stream);
int aext2_eq_dp_all_batched_ffi(const uint64_t *d_r_all, uint64_t *d_buf_a_all, uint64_t *d_buf_b_all, int log_n,
                                int num_leaves, size_t leaf_stride, uint64_t **d_result, void *stream);

/* partial evaluation, ffi.rs:198-212 (first m variables, LSB first) */
int agl_partial_eval_ffi(uint64_t *d_data, const uint64_t *d_r, int log_n, int m);
int agl_partial_eval_ext2_from_base_ffi(const uint64_t *d_input, uint64_t *d_output, const uint64_t *d_r, int log_n,
                                        int m);

/* fused permute + partial evaluation, ffi.rs:214-226 (semantics as the Goldilocks twin) */
int agl_fused_permute_partial_eval_ffi(const uint64_t *d_evals, uint64_t *d_output, const uint64_t *d_eq_table,
                                       const uint32_t *d_lo_lut, const uint32_t *d_hi_lut, int n, int m, int half,
                                       int output_size, int smem_bytes);

/* sumcheck with per-workgroup partials, ffi.rs:228-263 */
int agl_sumcheck_round_message_ffi(const uint64_t *d_polys, uint64_t *d_partial, int d, size_t original_size,
                                   size_t half, int num_blocks);
int agl_sumcheck_fold_ffi(const uint64_t *d_input, uint64_t *d_output, uint64_t challenge, int d,
                          size_t original_size, size_t half);
int aext2_sumcheck_round_message_ffi(const uint64_t *d_polys, uint64_t *d_partial, int d, size_t original_size,
                                     size_t half, int num_blocks);
int aext2_sumcheck_fold_ffi(const uint64_t *d_input, uint64_t *d_output, uint64_t challenge_c0,
                            uint64_t challenge_c1, int d, size_t original_size, size_t half);
/* batched same-size degree-2 sumcheck, ffi.rs:265-272 and :320-328 (GpuBatchedSamePointState,
 * sumcheck_prover.rs:468-1027): partial[(block_x * num_leaves + leaf) * 3 + c] Ext2 */
int aext2_sumcheck_batched_round_message_ffi(const uint64_t *d_polys, uint64_t *d_partial, size_t original_size,
                                             size_t half, int num_leaves, int num_blocks_x);
int aext2_sumcheck_batched_fold_ffi(const uint64_t *d_input, uint64_t *d_output, uint64_t challenge_c0,
                                    uint64_t challenge_c1, size_t original_size, size_t half, int num_leaves);
/* round 0 of the same state with a binary f given as bits (packed_size_u64 words per leaf,
 * bit i of word i / 64 = f(i)): message from the eq slot and the bits, ffi.rs:402-411; fold
 * writing eq' and the Ext2 f' = f0 + ch (f1 - f0) at half size, ffi.rs:413-423 */
int aext2_sumcheck_batched_round0_binary_msg_ffi(const uint64_t *d_polys, const uint64_t *d_packed,
                                                 uint64_t *d_partial, size_t original_size, size_t half,
                                                 int num_leaves, int num_blocks_x, size_t packed_size_u64);
int aext2_sumcheck_batched_round0_binary_fold_ffi(const uint64_t *d_polys, const uint64_t *d_packed,
                                                  uint64_t *d_output, uint64_t challenge_c0, uint64_t challenge_c1,
                                                  size_t original_size, size_t half, int num_leaves,
                                                  size_t packed_size_u64);
/* lifts into the f slot of that layout: bits (ffi.rs:293-299), or pos - neg in {0, 1, q - 1}
 * (ffi.rs:274-281) */
int aext2_batched_lift_binary_ffi(const uint64_t *d_packed, uint64_t *d_polys, size_t original_size,
                                  int num_leaves, size_t packed_size_u64);
int aext2_batched_lift_ternary_single_ffi(const uint64_t *d_pos, const uint64_t *d_neg, uint64_t *d_polys,
                                          size_t original_size, int num_leaves, size_t packed_size_u64);
/* Selective add over a shared eq table (ffi.rs:283-291): for each of n_planes packed binary
 * planes, d_partial[(bx * n_planes + p) * 2 ..] = the Ext2 sum of eq[i] over the set bits i that
 * workgroup bx of num_blocks_x (256 threads, grid-stride) visits; callers add the partials. */
int aext2_selective_add_batched_planes_ffi(const uint64_t *d_eq, const uint64_t *d_packed_planes,
                                           uint64_t *d_partial, size_t total, int n_planes, size_t packed_size_u64,
                                           int num_blocks_x);

/* ring Ajtai over F_q[X]/(X^64 + 1), KAPPA = 42 (ajtai.rs:138-386); d_key: 8 u32 on the
 * device; outputs [B][42][64] canonical coefficients. */
int ajtai_commit_dense_batched_ffi(const uint32_t *d_key, const uint64_t *d_z, uint64_t N, int B, int chunk,
                                   uint64_t *d_out);                                   /* ffi.rs:426-433 */
int ajtai_commit_dense_batched_at_ffi(const uint32_t *d_key, const uint64_t *d_z, uint64_t N, int B, int chunk,
                                      uint64_t col_offset, uint64_t *d_out);           /* ffi.rs:500-508 */
int ajtai_commit_wide_ffi(const uint32_t *d_key, const uint64_t *d_z_wide, uint64_t N, uint64_t col_offset,
                          int chunk, uint64_t *d_out);                                 /* ffi.rs:595-602 */
int ajtai_commit_sparse_ffi(const uint32_t *d_key, const uint64_t *d_positions, uint64_t K, int chunk,
                            uint64_t *d_out);                                          /* ffi.rs:649-655 */
/* Ternary commits of the 13 split chunks z_b = pos_b - neg_b (pos, neg: [13][N_ring] bit words;
 * out [13][42][64]); from ChaCha8, or from a matrix materialized as M[j][row][64]
 * (ajtai_materialize_m_ffi, N * 42 * 64 words). Queued on the calling thread's stream; chunk as
 * above (else -3). */
int ajtai_commit_ternary_ffi(const uint32_t *d_key, const uint64_t *d_pos, const uint64_t *d_neg, uint64_t N_ring,
                             int chunk, uint64_t *d_out);                              /* ffi.rs:603-610 */
int ajtai_materialize_m_ffi(const uint32_t *d_chacha_key, uint64_t *d_M, uint64_t N);  /* ffi.rs:612-616 */
int ajtai_commit_ternary_premat_ffi(const uint64_t *d_M, const uint64_t *d_pos, const uint64_t *d_neg,
                                    uint64_t N_ring, int chunk, uint64_t *d_out);      /* ffi.rs:618-625 */
/* Folds with challenges of 64 coefficients in {-1, 0, 1, 2}: out_j = z1_j + z2_j r (binary
 * witnesses [N_ring], out [N_ring][64] canonical; r_coeffs host memory) and out_i = c1_i + r c2_i
 * (commitments [42][64]). */
int ajtai_fold_witness_ffi(const uint64_t *d_z1, const uint64_t *d_z2, const int8_t *r_coeffs, uint64_t N_ring,
                           int chunk, uint64_t *d_out);                                /* ffi.rs:435-442 */
int ajtai_fold_commitment_ffi(const uint64_t *d_c1, const uint64_t *d_c2, const int8_t *r_coeffs,
                              uint64_t *d_out);                                        /* ffi.rs:642-647 */
/* Multifolds: out_j = sum_m z_m,j r_m as int16 [N_ring][64] (binary instances [M][N_ring];
 * the mixed forms add ternary instances pos - neg after the binary ones; challenges
 * [(K + T)][64] on the device), and out_i = sum_m r_m c_m,i (commitments [M][42][64]). The _tc
 * forms compute the same result. */
int ajtai_multifold_witness_ffi(const uint64_t *d_z_packed, const int8_t *d_r_all, int16_t *d_out,
                                int num_instances, uint64_t N_ring, uint64_t chunk_size); /* ffi.rs:444-451 */
int ajtai_multifold_mixed_witness_ffi(const uint64_t *d_z_bin_packed, const uint64_t *d_pos_packed,
                                      const uint64_t *d_neg_packed, const int8_t *d_r_all, int16_t *d_out,
                                      int num_binary, int num_ternary, uint64_t N_ring,
                                      uint64_t chunk_size);                            /* ffi.rs:453-463 */
int ajtai_multifold_mixed_witness_tc_ffi(const uint64_t *d_z_bin_packed, const uint64_t *d_pos_packed,
                                         const uint64_t *d_neg_packed, const int8_t *d_r_all, int16_t *d_out,
                                         int num_binary, int num_ternary, uint64_t N_ring); /* ffi.rs:465-474 */
int ajtai_multifold_mixed_witness_tc_fused_ffi(const uint64_t *d_z_bin_packed, const uint64_t *d_pos_packed,
                                               const uint64_t *d_neg_packed, const int8_t *d_r_all,
                                               int16_t *d_out, int num_binary, int num_ternary,
                                               uint64_t N_ring);                       /* ffi.rs:476-485 */
int ajtai_multifold_commitment_ffi(const uint64_t *d_c_packed, const int8_t *d_r_all, int num_instances,
                                   uint64_t *d_out);                                   /* ffi.rs:635-640 */
/* Split of a wide int16 witness [N_ring][64] into 13 ternary chunks: v = sum_i 2^i (pos_i - neg_i),
 * pos, neg [13][N_ring] bit words with no common bit. */
int ajtai_split_witness_ffi(const int16_t *d_z_wide, uint64_t *d_pos_chunks, uint64_t *d_neg_chunks,
                            uint64_t N_ring);                                          /* ffi.rs:487-492 */
/* No reference counterpart: host copy of M_{row, col} (64 coefficients) for checks. */
int mam_ajtai_ring_matrix_entry(const uint8_t *key32, uint32_t row, uint64_t col, uint64_t *out64);

/* Tensor-core commit probe (ffi.rs:627-633): int8 GEMM partials
 * partial[kb][nt][r * 16 + c] = sum over k in [1024 kb, min(1024 (kb + 1), K_total)) of
 * z[r * K_total + k] * M[(nt * 16 + c) * K_total + k], nt < 60, r, c < 16 (z row-major 16 x K,
 * M column-major K x 960); K_total a multiple of 16. Asynchronous; 0 / -1. */
int ajtai_tc_commit_probe_ffi(const int8_t *d_z_int8, const int8_t *d_M_int8, int32_t *d_partial, int K_total,
                              int num_K_chunks);

#ifdef __cplusplus
}
#endif
#endif /* MORPHEUS_AGL_H */
