/*
 * morpheus_gl.h -- C-ABI of libmorpheus_amd.so under the reference's symbol names.
 *
 * These are the hot-path entry points of the reference crate goldilocks-cuda-rs
 * (static library `goldilocks_cuda_wrapper`, goldilocks-cuda-rs/src/ffi.rs:5); each
 * declaration cites the ffi.rs line of the binding it replaces. Names, argument order,
 * argument meaning and return codes are unchanged, so the crate's `ffi.rs` links
 * against this library as it stands (see INTEGRATION.md for the build.rs change).
 *
 * Conventions (goldilocks-cuda-rs/cuda/wrapper.cu:27-97):
 *   - d_* are device pointers to raw u64 words; Ext2 values are interleaved [c0, c1].
 *   - return 0 on success, -1 on a runtime error, -2 for a rejected argument.
 *   - work is queued on the calling thread's per-thread stream; entry points the
 *     reference documents as synchronizing (interpolation, encode, Merkle builds,
 *     sumcheck round messages and folds, copies) complete before returning.
 *   - outputs are canonical residues in [0, p); inputs may be any 64-bit words.
 *   - p = 2^64 - 2^32 + 1, Ext2 = F_p[X]/(X^2 - 7).
 */
#ifndef MORPHEUS_GL_H
#define MORPHEUS_GL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- initialization and memory (ffi.rs:11-26, 306-309) ---- */
int goldilocks_cuda_init(void);                                     /* ffi.rs:11 */
int poseidon2_cuda_init(void);                                      /* ffi.rs:12 */
int monolith_cuda_init(void);                                       /* ffi.rs:565 */
int cuda_malloc(void **ptr, size_t size);                           /* ffi.rs:18 */
int cuda_free(void *ptr);                                           /* ffi.rs:19; caching, see morpheus_agl.h */
int cuda_memcpy_htod(void *dst, const void *src, size_t size);      /* ffi.rs:20 */
int cuda_memcpy_dtoh(void *dst, const void *src, size_t size);      /* ffi.rs:21 */
int cuda_memcpy_dtod(void *dst, const void *src, size_t size);      /* ffi.rs:22 */
int cuda_device_synchronize(void);                                  /* ffi.rs:23 */
int cuda_get_last_error(void);                                      /* ffi.rs:24 */
int cuda_peek_at_last_error(void);                                  /* ffi.rs:25 */
int cuda_mem_get_info(size_t *free_bytes, size_t *total_bytes);     /* ffi.rs:26 */
int cuda_set_device(int device);                                    /* ffi.rs:306 */
int cuda_get_device(int *device);                                   /* ffi.rs:307 */
int cuda_get_device_count(int *count);                              /* ffi.rs:308 */
int cuda_get_device_name(int device, char *name, int max_len);      /* ffi.rs:309 */
int gl_memset_zero(uint64_t *d_buf, size_t n_u64);                  /* ffi.rs:105 */
int gl_zero_buffer(uint64_t *d_buf, int n);                         /* ffi.rs:131 */

/* ---- base-field batches (ffi.rs:32-53, 163-178) ---- */
int gl_batch_add(const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_result, int n);   /* ffi.rs:32 */
int gl_batch_sub(const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_result, int n);   /* ffi.rs:39 */
int gl_batch_mul(const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_result, int n);   /* ffi.rs:46 */
int gl_batch_inverse(const uint64_t *d_a, uint64_t *d_result, int n);                    /* ffi.rs:53 */
int gl_batch_mul_scalar(uint64_t scalar, const uint64_t *d_a, uint64_t *d_result, int n); /* ffi.rs:163 */
int gl_batch_neg(const uint64_t *d_a, uint64_t *d_result, int n);                        /* ffi.rs:170 */
int gl_batch_square(const uint64_t *d_a, uint64_t *d_result, int n);                     /* ffi.rs:172 */
int gl_batch_double(const uint64_t *d_a, uint64_t *d_result, int n);                     /* ffi.rs:174 */
int gl_batch_exp(const uint64_t *d_a, uint64_t exp, uint64_t *d_result, int n);          /* ffi.rs:176 */
int gl_batch_div(const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_result, int n);   /* ffi.rs:178 */

/* ---- Ext2 batches (ffi.rs:189-227, 275-277); n counts Ext2 elements ---- */
int ext2_batch_add_ffi(const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_result, int n);   /* ffi.rs:189 */
int ext2_batch_sub_ffi(const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_result, int n);   /* ffi.rs:196 */
int ext2_batch_mul_ffi(const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_result, int n);   /* ffi.rs:203 */
int ext2_batch_inverse_ffi(const uint64_t *d_a, uint64_t *d_result, int n);                    /* ffi.rs:210 */
int ext2_batch_mul_scalar_ffi(uint64_t scalar, const uint64_t *d_a, uint64_t *d_result, int n); /* ffi.rs:212 */
int ext2_batch_neg_ffi(const uint64_t *d_a, uint64_t *d_result, int n);                        /* ffi.rs:219 */
int ext2_batch_square_ffi(const uint64_t *d_a, uint64_t *d_result, int n);                     /* ffi.rs:221 */
int ext2_batch_frobenius_ffi(const uint64_t *d_a, uint64_t *d_result, int n);                  /* ffi.rs:223 */
int ext2_batch_conjugate_ffi(const uint64_t *d_a, uint64_t *d_result, int n);                  /* ffi.rs:225 */
int ext2_batch_exp_ffi(const uint64_t *d_a, uint64_t exp, uint64_t *d_result, int n);          /* ffi.rs:227 */
int gl_to_ext2_batch_ffi(const uint64_t *d_input, uint64_t *d_output, int n);                  /* ffi.rs:275 */
int ext2_to_gl_batch_ffi(const uint64_t *d_input, uint64_t *d_output, int n);                  /* ffi.rs:277 */

/* ---- Poseidon2-w8 batches (ffi.rs:287-296) ---- */
int poseidon2_hash_batch_ffi(const uint64_t *d_input, uint64_t *d_output, int n);              /* ffi.rs:287 */
int poseidon2_compress_batch_ffi(const uint64_t *d_left, const uint64_t *d_right, uint64_t *d_output,
                                 int n);                                                       /* ffi.rs:289 */
int poseidon2_merkle_layer_ffi(const uint64_t *d_input, uint64_t *d_output, int n);            /* ffi.rs:296 */

/* ---- partial evaluation and eq tables (ffi.rs:377-410) ---- */
int partial_eval_gl_ffi(uint64_t *d_data, const uint64_t *d_r, int log_n, int m);              /* ffi.rs:377 */
/* d_output holds 2^(log_n - 1) Ext2 entries (the first layer's output, as partial_eval.rs:116-117
 * allocates it); the result is its first 2^(log_n - m) entries. */
int partial_eval_ext2_from_gl_ffi(const uint64_t *d_input, uint64_t *d_output, const uint64_t *d_r, int log_n,
                                  int m);                                                      /* ffi.rs:384 */
/* The result buffer is reported through d_result (always d_buf_a here). */
int eq_dp_all_ffi(const uint64_t *d_r, uint64_t *d_buf_a, uint64_t *d_buf_b, int log_n,
                  uint64_t **d_result);                                                        /* ffi.rs:396 */
int ext2_eq_dp_all_ffi(const uint64_t *d_r, uint64_t *d_buf_a, uint64_t *d_buf_b, int log_n,
                       uint64_t **d_result);                                                   /* ffi.rs:404 */

/* ---- Basefold (ffi.rs:416-542) ---- */
int basefold_bit_reverse_gl_ffi(uint64_t *d_data, int log_n);                                  /* ffi.rs:416 */
int basefold_bit_reverse_ext2_ffi(uint64_t *d_data, int log_n);                                /* ffi.rs:417 */
int basefold_bhc_interpolate_ffi(const uint64_t *d_evals, uint64_t *d_coeffs, uint64_t *d_bh_evals,
                                 int num_vars);                                                /* ffi.rs:419 */
int basefold_encode_ffi(const uint64_t *d_coeffs, uint64_t *d_codeword, int num_vars, int log_rate); /* ffi.rs:426 */
int basefold_fold_gl_ffi(const uint64_t *d_codeword, const uint64_t *d_table, uint64_t challenge,
                         uint64_t *d_output, int pair_count);                                  /* ffi.rs:433 */
int basefold_fold_mixed_ffi(const uint64_t *d_codeword, const uint64_t *d_table, uint64_t challenge_c0,
                            uint64_t challenge_c1, uint64_t *d_output, int pair_count);        /* ffi.rs:441 */
int basefold_fold_ext2_ffi(const uint64_t *d_codeword, const uint64_t *d_table, uint64_t challenge_c0,
                           uint64_t challenge_c1, uint64_t *d_output, int pair_count);         /* ffi.rs:450 */
int basefold_sumcheck_interp_gl_ffi(uint64_t *d_data, int pair_count);                         /* ffi.rs:459 */
int basefold_sumcheck_interp_ext2_ffi(uint64_t *d_data, int pair_count);                       /* ffi.rs:460 */
int basefold_sumcheck_product_gl_ffi(const uint64_t *d_eq, const uint64_t *d_bh, uint64_t *d_partial_c0,
                                     uint64_t *d_partial_c1, uint64_t *d_partial_c2, int pair_count,
                                     int num_blocks);                                          /* ffi.rs:462 */
int basefold_sumcheck_product_mixed_ffi(const uint64_t *d_eq, const uint64_t *d_bh, uint64_t *d_partial_c0,
                                        uint64_t *d_partial_c1, uint64_t *d_partial_c2, int pair_count,
                                        int num_blocks);                                       /* ffi.rs:472 */
int basefold_sumcheck_product_ext2_ffi(const uint64_t *d_eq, const uint64_t *d_bh, uint64_t *d_partial_c0,
                                       uint64_t *d_partial_c1, uint64_t *d_partial_c2, int pair_count,
                                       int num_blocks);                                        /* ffi.rs:482 */
int basefold_sumcheck_eval_gl_ffi(const uint64_t *d_data, uint64_t challenge, uint64_t *d_output,
                                  int pair_count);                                             /* ffi.rs:492 */
int basefold_sumcheck_eval_mixed_ffi(const uint64_t *d_data, uint64_t challenge_c0, uint64_t challenge_c1,
                                     uint64_t *d_output, int pair_count);                      /* ffi.rs:499 */
int basefold_sumcheck_eval_ext2_ffi(const uint64_t *d_data, uint64_t challenge_c0, uint64_t challenge_c1,
                                    uint64_t *d_output, int pair_count);                       /* ffi.rs:507 */
int fused_sumcheck_round_ext2_ffi(const uint64_t *d_eq_in, const uint64_t *d_bh_in, uint64_t challenge_c0,
                                  uint64_t challenge_c1, uint64_t *d_eq_out, uint64_t *d_bh_out,
                                  uint64_t *d_partial_c0, uint64_t *d_partial_c1, uint64_t *d_partial_c2,
                                  int pair_count, int num_blocks);                             /* ffi.rs:515 */
int basefold_dot_product_gl_ffi(const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_partial, int n,
                                int num_blocks);                                               /* ffi.rs:529 */
int basefold_dot_product_mixed_ffi(const uint64_t *d_a, const uint64_t *d_b, uint64_t *d_partial, int n,
                                   int num_blocks);                                            /* ffi.rs:537 */

/* ---- Merkle trees, layout [N leaves][N/2]...[root] x 4 words (ffi.rs:549-582) ---- */
int poseidon2_merkle_tree_gl_ffi(const uint64_t *d_codeword, uint64_t *d_tree, int num_leaves);   /* ffi.rs:549 */
int poseidon2_merkle_tree_ext2_ffi(const uint64_t *d_codeword, uint64_t *d_tree, int num_leaves); /* ffi.rs:555 */
int monolith_merkle_tree_gl_ffi(const uint64_t *d_codeword, uint64_t *d_tree, int num_leaves);    /* ffi.rs:567 */
int monolith_merkle_tree_ext2_ffi(const uint64_t *d_codeword, uint64_t *d_tree, int num_leaves);  /* ffi.rs:573 */
int monolith_permute_test_ffi(const uint64_t *h_input, uint64_t *h_output);                      /* ffi.rs:579 */

/* ---- linear sumcheck (ffi.rs:588-627); partials: num_blocks * (d + 1) [* 2] words ---- */
int sumcheck_round_message_ffi(const uint64_t *d_polys, uint64_t *d_partial, int d, size_t original_size,
                               size_t half, int num_blocks);                                   /* ffi.rs:588 */
int sumcheck_fold_ffi(const uint64_t *d_input, uint64_t *d_output, uint64_t challenge, int d,
                      size_t original_size, size_t half);                                      /* ffi.rs:597 */
int sumcheck_round_message_ext2_ffi(const uint64_t *d_polys, uint64_t *d_partial, int d, size_t original_size,
                                    size_t half, int num_blocks);                              /* ffi.rs:610 */
int sumcheck_fold_ext2_ffi(const uint64_t *d_input, uint64_t *d_output, uint64_t challenge_c0,
                           uint64_t challenge_c1, int d, size_t original_size, size_t half);   /* ffi.rs:619 */

/* ---- Fiat-Shamir duplex challengers, n independent Poseidon2-w8 sponges (rate 4) in device
 * memory, one per thread of each batch call (ffi.rs:315-371; cuda/challenger.cuh:43-259).
 * State: 136 bytes = sponge[8], in[4], out[4] (u64), in_count, out_count (int32). Values
 * layout: challenger i uses values[i * count ..]; Ext2 = 2 consecutive words, Ext5 = 5.
 * Launches are asynchronous on the calling thread's stream; the copies synchronize it. ---- */
int challenger_state_size(void);                                                            /* ffi.rs:315 */
int challenger_alloc_states(void **d_states, int n);                                        /* ffi.rs:316 */
int challenger_init_states(void *d_states, int n);                                          /* ffi.rs:317 */
int challenger_observe_ffi(void *d_states, const uint64_t *d_values, int n);                /* ffi.rs:319 */
int challenger_observe_slice_ffi(void *d_states, const uint64_t *d_values, int count, int n); /* ffi.rs:325 */
int challenger_sample_ffi(void *d_states, uint64_t *d_outputs, int n);                      /* ffi.rs:332 */
int challenger_sample_array_ffi(void *d_states, uint64_t *d_outputs, int count, int n);     /* ffi.rs:338 */
int challenger_sample_ext2_ffi(void *d_states, uint64_t *d_outputs, int n);                 /* ffi.rs:345 */
int challenger_sample_ext5_ffi(void *d_states, uint64_t *d_outputs, int n);                 /* ffi.rs:351 */
int challenger_observe_ext2_ffi(void *d_states, const uint64_t *d_values, int n);           /* ffi.rs:357 */
int challenger_copy_state_to_host(void *h_state, const void *d_state);                      /* ffi.rs:363 */
int challenger_copy_state_to_device(void *d_state, const void *h_state);                    /* ffi.rs:368 */

/* acc[i] += (s0 + s1 X) * src[i] over Ext2 X^2 - 7, n Ext2 entries (ffi.rs:633) */
int ext2_scale_accumulate_ffi(uint64_t scalar_c0, uint64_t scalar_c1, const uint64_t *d_src, uint64_t *d_acc,
                              int n);
/* out[x] = in[y] with bit b of x = bit d_perm_map[b] of y; total = 2^n_bits (ffi.rs:645).
 * d_perm_map is the inverse variable map (new position -> old), as field.rs:205-214 passes it. */
int bit_permute_gl_ffi(const uint64_t *d_input, uint64_t *d_output, const int *d_perm_map, int n_bits, int total);

/* fused variable permutation + partial evaluation, ffi.rs:657-668:
 * output[j] = sum_{b < 2^m} evals[lo_lut[x & (2^half-1)] | hi_lut[x >> half]] * eq[b], x = j 2^m + b */
int fused_permute_partial_eval_ffi(const uint64_t *d_evals, uint64_t *d_output, const uint64_t *d_eq_table,
                                   const uint32_t *d_lo_lut, const uint32_t *d_hi_lut, int n, int m, int half,
                                   int output_size, int smem_bytes);

#ifdef __cplusplus
}
#endif
#endif /* MORPHEUS_GL_H */
