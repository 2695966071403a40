"""Loader for libmorpheus_amd.so and the ctypes signatures of its C-ABI.

The library is built in-tree (``morpheus_amd/libmorpheus_amd.so``) by
``morpheus_amd.build.build()``. There is no fallback: importing this module without the
built library raises ``ImportError``.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmorpheus_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()' "
        "or make -C morpheus_amd/csrc)")

lib = ctypes.CDLL(LIB_PATH)

u64 = ctypes.c_uint64
u64p = ctypes.POINTER(ctypes.c_uint64)
vp = ctypes.c_void_p
ci = ctypes.c_int
cl = ctypes.c_long
csz = ctypes.c_size_t
u8p = ctypes.c_char_p


class Transcript_t(ctypes.Structure):
    _fields_ = [("state", u64 * 8), ("in_buf", u64 * 4), ("out_buf", u64 * 4),
                ("in_count", ctypes.c_int64), ("out_count", ctypes.c_int64)]


trp = ctypes.POINTER(Transcript_t)


class QTranscript_t(ctypes.Structure):
    _fields_ = [("state", u64 * 12), ("in_buf", u64 * 8), ("out_buf", u64 * 8),
                ("in_count", ctypes.c_int64), ("out_count", ctypes.c_int64)]


qtrp = ctypes.POINTER(QTranscript_t)


class MatmulInstance_t(ctypes.Structure):
    _fields_ = [("d_a", vp), ("d_b", vp), ("d_c", vp), ("lm", ci), ("lk", ci), ("ln", ci),
                ("label", ctypes.c_char_p), ("label_len", cl), ("roots", u64p), ("ajtai", u64p), ("point", u64p),
                ("claim", u64p), ("round_msgs", u64p), ("challenges", u64p), ("final_evals", u64p),
                ("proof_a", u64p), ("proof_b", u64p), ("proof_c", u64p), ("start_ms", ctypes.c_double),
                ("end_ms", ctypes.c_double), ("status", ci)]


class EvalClaim_t(ctypes.Structure):
    _fields_ = [("poly", ci), ("point", ci), ("value", u64 * 2)]


class ServeRequest_t(ctypes.Structure):
    _fields_ = [("inst", ctypes.POINTER(MatmulInstance_t)), ("arrival_ms", ctypes.c_double),
                ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double), ("status", ci)]


class BatchInstance_t(ctypes.Structure):
    _fields_ = [("d_fg", vp), ("num_vars", ci), ("label", ctypes.c_char_p), ("label_len", cl),
                ("round_msgs", u64p), ("challenges", u64p), ("final_evals", u64p), ("root", u64p),
                ("proof", u64p), ("start_ms", ctypes.c_double), ("end_ms", ctypes.c_double), ("status", ci)]

_SIGS = {
    # reference-named (morpheus_gl.h)
    "goldilocks_cuda_init": (ci, []),
    "poseidon2_cuda_init": (ci, []),
    "monolith_cuda_init": (ci, []),
    "cuda_malloc": (ci, [ctypes.POINTER(vp), csz]),
    "cuda_free": (ci, [vp]),
    "cuda_memcpy_htod": (ci, [vp, vp, csz]),
    "cuda_memcpy_dtoh": (ci, [vp, vp, csz]),
    "cuda_memcpy_dtod": (ci, [vp, vp, csz]),
    "cuda_device_synchronize": (ci, []),
    "cuda_get_last_error": (ci, []),
    "cuda_peek_at_last_error": (ci, []),
    "cuda_mem_get_info": (ci, [ctypes.POINTER(csz), ctypes.POINTER(csz)]),
    "cuda_set_device": (ci, [ci]),
    "cuda_get_device": (ci, [ctypes.POINTER(ci)]),
    "cuda_get_device_count": (ci, [ctypes.POINTER(ci)]),
    "cuda_get_device_name": (ci, [ci, ctypes.c_char_p, ci]),
    "gl_memset_zero": (ci, [vp, csz]),
    "gl_zero_buffer": (ci, [vp, ci]),
    "gl_batch_add": (ci, [vp, vp, vp, ci]),
    "gl_batch_sub": (ci, [vp, vp, vp, ci]),
    "gl_batch_mul": (ci, [vp, vp, vp, ci]),
    "gl_batch_div": (ci, [vp, vp, vp, ci]),
    "gl_batch_inverse": (ci, [vp, vp, ci]),
    "gl_batch_neg": (ci, [vp, vp, ci]),
    "gl_batch_square": (ci, [vp, vp, ci]),
    "gl_batch_double": (ci, [vp, vp, ci]),
    "gl_batch_exp": (ci, [vp, u64, vp, ci]),
    "gl_batch_mul_scalar": (ci, [u64, vp, vp, ci]),
    "ext2_batch_add_ffi": (ci, [vp, vp, vp, ci]),
    "ext2_batch_sub_ffi": (ci, [vp, vp, vp, ci]),
    "ext2_batch_mul_ffi": (ci, [vp, vp, vp, ci]),
    "ext2_batch_inverse_ffi": (ci, [vp, vp, ci]),
    "ext2_batch_mul_scalar_ffi": (ci, [u64, vp, vp, ci]),
    "ext2_batch_neg_ffi": (ci, [vp, vp, ci]),
    "ext2_batch_square_ffi": (ci, [vp, vp, ci]),
    "ext2_batch_frobenius_ffi": (ci, [vp, vp, ci]),
    "ext2_batch_conjugate_ffi": (ci, [vp, vp, ci]),
    "ext2_batch_exp_ffi": (ci, [vp, u64, vp, ci]),
    "gl_to_ext2_batch_ffi": (ci, [vp, vp, ci]),
    "ext2_to_gl_batch_ffi": (ci, [vp, vp, ci]),
    "poseidon2_hash_batch_ffi": (ci, [vp, vp, ci]),
    "poseidon2_compress_batch_ffi": (ci, [vp, vp, vp, ci]),
    "poseidon2_merkle_layer_ffi": (ci, [vp, vp, ci]),
    "partial_eval_gl_ffi": (ci, [vp, vp, ci, ci]),
    "partial_eval_ext2_from_gl_ffi": (ci, [vp, vp, vp, ci, ci]),
    "eq_dp_all_ffi": (ci, [vp, vp, vp, ci, ctypes.POINTER(vp)]),
    "ext2_eq_dp_all_ffi": (ci, [vp, vp, vp, ci, ctypes.POINTER(vp)]),
    "basefold_bit_reverse_gl_ffi": (ci, [vp, ci]),
    "basefold_bit_reverse_ext2_ffi": (ci, [vp, ci]),
    "basefold_bhc_interpolate_ffi": (ci, [vp, vp, vp, ci]),
    "basefold_encode_ffi": (ci, [vp, vp, ci, ci]),
    "basefold_fold_gl_ffi": (ci, [vp, vp, u64, vp, ci]),
    "basefold_fold_mixed_ffi": (ci, [vp, vp, u64, u64, vp, ci]),
    "basefold_fold_ext2_ffi": (ci, [vp, vp, u64, u64, vp, ci]),
    "basefold_sumcheck_interp_gl_ffi": (ci, [vp, ci]),
    "basefold_sumcheck_interp_ext2_ffi": (ci, [vp, ci]),
    "basefold_sumcheck_product_gl_ffi": (ci, [vp, vp, vp, vp, vp, ci, ci]),
    "basefold_sumcheck_product_mixed_ffi": (ci, [vp, vp, vp, vp, vp, ci, ci]),
    "basefold_sumcheck_product_ext2_ffi": (ci, [vp, vp, vp, vp, vp, ci, ci]),
    "basefold_sumcheck_eval_gl_ffi": (ci, [vp, u64, vp, ci]),
    "basefold_sumcheck_eval_mixed_ffi": (ci, [vp, u64, u64, vp, ci]),
    "basefold_sumcheck_eval_ext2_ffi": (ci, [vp, u64, u64, vp, ci]),
    "fused_sumcheck_round_ext2_ffi": (ci, [vp, vp, u64, u64, vp, vp, vp, vp, vp, ci, ci]),
    "basefold_dot_product_gl_ffi": (ci, [vp, vp, vp, ci, ci]),
    "basefold_dot_product_mixed_ffi": (ci, [vp, vp, vp, ci, ci]),
    "poseidon2_merkle_tree_gl_ffi": (ci, [vp, vp, ci]),
    "poseidon2_merkle_tree_ext2_ffi": (ci, [vp, vp, ci]),
    "monolith_merkle_tree_gl_ffi": (ci, [vp, vp, ci]),
    "monolith_merkle_tree_ext2_ffi": (ci, [vp, vp, ci]),
    "monolith_permute_test_ffi": (ci, [u64p, u64p]),
    "sumcheck_round_message_ffi": (ci, [vp, vp, ci, csz, csz, ci]),
    "sumcheck_fold_ffi": (ci, [vp, vp, u64, ci, csz, csz]),
    "sumcheck_round_message_ext2_ffi": (ci, [vp, vp, ci, csz, csz, ci]),
    "sumcheck_fold_ext2_ffi": (ci, [vp, vp, u64, u64, ci, csz, csz]),
    # prover level (morpheus_prover.h)
    "mam_transcript_init": (None, [trp, u8p, cl]),
    "mam_transcript_append_u64": (None, [trp, u8p, cl, u64]),
    "mam_transcript_append_scalars": (None, [trp, u8p, cl, u64p, cl]),
    "mam_transcript_append_ext2": (None, [trp, u8p, cl, u64p]),
    "mam_transcript_challenge_scalar": (u64, [trp, u8p, cl]),
    "mam_transcript_challenge_ext2": (None, [trp, u8p, cl, u64p]),
    "mam_poseidon2_permute_host": (None, [u64p]),
    "mam_monolith_permute_host": (None, [u64p]),
    "mam_sumcheck_prove": (ci, [trp, vp, ci, ci, ci, u64p, u64p, u64p, u64p]),
    "mam_sc_state_new": (ci, [vp, ci, ci, ci, ctypes.POINTER(vp)]),
    "mam_sc_state_round_message": (ci, [vp, u64p]),
    "mam_sc_state_fold": (ci, [vp, u64, u64]),
    "mam_sc_state_final_evaluations": (ci, [vp, u64p]),
    "mam_sc_state_free": (ci, [vp]),
    "mam_basefold_table_len": (cl, [ci, ci, ci]),
    "mam_basefold_table_generate": (ci, [ci, ci, ci, u64, u64p]),
    "mam_basefold_commit": (ci, [vp, ci, ci, ci, ctypes.POINTER(vp), u64p]),
    "mam_basefold_commit_into": (ci, [vp, ci, ci, ci, vp, vp, vp, u64p]),
    "mam_basefold_commitment_wrap": (ci, [vp, vp, vp, u64p, ci, ci, ci, ctypes.POINTER(vp)]),
    "mam_basefold_commitment_buffers": (ci, [vp, ctypes.POINTER(vp), ctypes.POINTER(vp), ctypes.POINTER(vp), u64p]),
    "mam_basefold_commitment_free": (ci, [vp]),
    "mam_basefold_proof_len": (cl, [ci, ci, ci]),
    "mam_basefold_open_ext2": (ci, [vp, u64p, vp, ci, trp, ci, u64p]),
    "mam_basefold_verify_ext2": (ci, [u64p, u64p, ci, ci, ci, u64p, ci, u64p, ci, trp]),
    "mam_monolith_permute_batch": (ci, [vp, ci]),
    "mam_merkle_tree_async": (ci, [vp, vp, csz, ci, ci]),
    "mam_ajtai_i8_setup": (ci, [ctypes.c_char_p, ci, u64, ctypes.POINTER(vp)]),
    "mam_ajtai_i8_commit": (ci, [vp, vp, ci, vp]),
    "mam_ajtai_i8_commit_prefix": (ci, [vp, vp, ci, u64, vp]),
    "mam_ajtai_i8_commit_noalloc": (ci, [vp, vp, ci, u64, vp]),
    "mam_ajtai_i8_matrix_row": (ci, [vp, ci, u64, u64, ctypes.POINTER(ctypes.c_int8)]),
    "mam_ajtai_i8_free": (ci, [vp]),
    # almost-Goldilocks crate (morpheus_agl.h)
    "almost_goldilocks_cuda_init": (ci, []),
    **{f"agl_batch_{op}_ffi": (ci, [vp, vp, vp, ci]) for op in ("add", "sub", "mul", "div")},
    **{f"agl_batch_{op}_ffi": (ci, [vp, vp, ci]) for op in ("neg", "square", "double", "inverse")},
    "agl_batch_exp_ffi": (ci, [vp, u64, vp, ci]),
    "agl_batch_mul_scalar_ffi": (ci, [u64, vp, vp, ci]),
    **{f"aext2_batch_{op}_ffi": (ci, [vp, vp, vp, ci]) for op in ("add", "sub", "mul")},
    **{f"aext2_batch_{op}_ffi": (ci, [vp, vp, ci])
       for op in ("inverse", "neg", "square", "frobenius", "conjugate")},
    "aext2_batch_exp_ffi": (ci, [vp, u64, vp, ci]),
    "aext2_batch_mul_scalar_ffi": (ci, [u64, vp, vp, ci]),
    "aext2_scale_accumulate_ffi": (ci, [u64, u64, vp, vp, ci]),
    "agl_to_aext2_batch_ffi": (ci, [vp, vp, ci]),
    "aext2_to_agl_batch_ffi": (ci, [vp, vp, ci]),
    "agl_eq_dp_all_ffi": (ci, [vp, vp, vp, ci, ctypes.POINTER(vp)]),
    "aext2_eq_dp_all_ffi": (ci, [vp, vp, vp, ci, ctypes.POINTER(vp)]),
    "aext2_eq_dp_all_stream_ffi": (ci, [vp, vp, vp, ci, ctypes.POINTER(vp), vp]),
    "aext2_eq_dp_all_batched_ffi": (ci, [vp, vp, vp, ci, ci, csz, ctypes.POINTER(vp), vp]),
    "agl_partial_eval_ffi": (ci, [vp, vp, ci, ci]),
    "agl_partial_eval_ext2_from_base_ffi": (ci, [vp, vp, vp, ci, ci]),
    "agl_sumcheck_round_message_ffi": (ci, [vp, vp, ci, csz, csz, ci]),
    "agl_sumcheck_fold_ffi": (ci, [vp, vp, u64, ci, csz, csz]),
    "aext2_sumcheck_round_message_ffi": (ci, [vp, vp, ci, csz, csz, ci]),
    "aext2_sumcheck_fold_ffi": (ci, [vp, vp, u64, u64, ci, csz, csz]),
    "aext2_sumcheck_batched_round_message_ffi": (ci, [vp, vp, csz, csz, ci, ci]),
    "aext2_sumcheck_batched_fold_ffi": (ci, [vp, vp, u64, u64, csz, csz, ci]),
    "aext2_sumcheck_batched_round0_binary_msg_ffi": (ci, [vp, vp, vp, csz, csz, ci, ci, csz]),
    "aext2_sumcheck_batched_round0_binary_fold_ffi": (ci, [vp, vp, vp, u64, u64, csz, csz, ci, csz]),
    "aext2_batched_lift_binary_ffi": (ci, [vp, vp, csz, ci, csz]),
    "aext2_batched_lift_ternary_single_ffi": (ci, [vp, vp, vp, csz, ci, csz]),
    "agl_bit_permute_ffi": (ci, [vp, vp, vp, ci, ci]),
    "agl_cached_bytes": (csz, []),
    "agl_d2h_staged": (ci, [vp, vp, csz]),
    "agl_h2d_staged": (ci, [vp, vp, csz]),
    "agl_cuda_error_string": (ctypes.c_char_p, [ci]),
    "cuda_memset": (ci, [vp, ci, csz]),
    "cuda_pool_trim": (ci, [csz]),
    "cuda_sync_mine": (ci, []),
    "cuda_stream_create": (ci, [ctypes.POINTER(vp)]),
    "cuda_stream_destroy": (ci, [vp]),
    "cuda_stream_synchronize": (ci, [vp]),
    "cuda_memcpy_dtod_async": (ci, [vp, vp, csz, vp]),
    "cuda_memcpy_htod_async": (ci, [vp, vp, csz, vp]),
    "cuda_memcpy_peer": (ci, [vp, ci, vp, ci, csz]),
    "cuda_enable_peer_access": (ci, [ci]),
    "agl_zero_buffer_ffi": (ci, [vp, csz]),
    "ajtai_commit_dense_batched_ffi": (ci, [vp, vp, u64, ci, ci, vp]),
    "ajtai_commit_dense_batched_at_ffi": (ci, [vp, vp, u64, ci, ci, u64, vp]),
    "ajtai_commit_wide_ffi": (ci, [vp, vp, u64, u64, ci, vp]),
    "ajtai_commit_sparse_ffi": (ci, [vp, vp, u64, ci, vp]),
    "ajtai_commit_ternary_ffi": (ci, [vp, vp, vp, u64, ci, vp]),
    "ajtai_materialize_m_ffi": (ci, [vp, vp, u64]),
    "ajtai_commit_ternary_premat_ffi": (ci, [vp, vp, vp, u64, ci, vp]),
    "ajtai_fold_witness_ffi": (ci, [vp, vp, vp, u64, ci, vp]),
    "ajtai_fold_commitment_ffi": (ci, [vp, vp, vp, vp]),
    "ajtai_multifold_witness_ffi": (ci, [vp, vp, vp, ci, u64, u64]),
    "ajtai_multifold_mixed_witness_ffi": (ci, [vp, vp, vp, vp, vp, ci, ci, u64, u64]),
    "ajtai_multifold_mixed_witness_tc_ffi": (ci, [vp, vp, vp, vp, vp, ci, ci, u64]),
    "ajtai_multifold_mixed_witness_tc_fused_ffi": (ci, [vp, vp, vp, vp, vp, ci, ci, u64]),
    "ajtai_multifold_commitment_ffi": (ci, [vp, vp, ci, vp]),
    "ajtai_split_witness_ffi": (ci, [vp, vp, vp, u64]),
    "aext2_selective_add_batched_planes_ffi": (ci, [vp, vp, vp, csz, ci, csz, ci]),
    "mam_ajtai_ring_matrix_entry": (ci, [ctypes.c_char_p, ctypes.c_uint32, u64, u64p]),
    "mam_prove_batch": (ci, [ctypes.POINTER(BatchInstance_t), ci, vp, ci, ci, ci, ci, ci]),
    "mam_prove_batch_graph": (ci, [ctypes.POINTER(BatchInstance_t), ci, vp, ci, ci, ci, ci, ci, ci, ci]),
    "mam_batch_graph_release": (ci, []),
    "mam_batch_graph_phase_ms": (ci, [ctypes.POINTER(ctypes.c_double)]),
    "mam_matmul_prove": (ci, [ctypes.POINTER(MatmulInstance_t), vp, ci, ci, ci, ci, vp, ci]),
    "challenger_state_size": (ci, []), "challenger_alloc_states": (ci, [ctypes.POINTER(vp), ci]),
    "challenger_init_states": (ci, [vp, ci]), "challenger_observe_ffi": (ci, [vp, vp, ci]),
    "challenger_observe_slice_ffi": (ci, [vp, vp, ci, ci]), "challenger_sample_ffi": (ci, [vp, vp, ci]),
    "challenger_sample_array_ffi": (ci, [vp, vp, ci, ci]), "challenger_sample_ext2_ffi": (ci, [vp, vp, ci]),
    "challenger_sample_ext5_ffi": (ci, [vp, vp, ci]), "challenger_observe_ext2_ffi": (ci, [vp, vp, ci]),
    "challenger_copy_state_to_host": (ci, [vp, vp]), "challenger_copy_state_to_device": (ci, [vp, vp]),
    "ext2_scale_accumulate_ffi": (ci, [u64, u64, vp, vp, ci]), "bit_permute_gl_ffi": (ci, [vp, vp, vp, ci, ci]),
    "mam_basefold_batch_proof_len": (ctypes.c_long, [ci, ci, ci, ci]),
    "mam_basefold_batch_open_ext2": (ci, [vp, ci, vp, ci, vp, ci, vp, ci, vp, ci, vp]),
    "mam_basefold_batch_verify_ext2": (ci, [vp, ci, ci, ci, ci, vp, ci, vp, ci, vp, ci, vp, ci, vp]),
    "mam_sumcheck_proof_bincode_len": (ctypes.c_long, [ci, ci]),
    "mam_sumcheck_proof_bincode": (ctypes.c_long, [vp, ci, ci, vp, vp, ctypes.c_long]),
    "ajtai_tc_commit_probe_ffi": (ci, [vp, vp, vp, ci, ci]),
    "mam_serve_matmul": (ci, [ctypes.POINTER(ServeRequest_t), ci, vp, ci, ci, ci, ci, vp, ci, ci, u64]),
    "fused_permute_partial_eval_ffi": (ci, [vp, vp, vp, vp, vp, ci, ci, ci, ci, ci]),
    "agl_fused_permute_partial_eval_ffi": (ci, [vp, vp, vp, vp, vp, ci, ci, ci, ci, ci]),
    "mam_profile_enable": (ci, [ci]), "mam_profile_enable_mask": (ci, [ctypes.c_uint]),
    "mam_valu_probe": (ci, [ci, ci, ci, ctypes.POINTER(ctypes.c_double)]),
    "mam_copy_probe": (ci, [ctypes.c_size_t, ci, ctypes.POINTER(ctypes.c_double)]),
    "mam_release_scratch": (ci, []),
    "mam_host_alloc": (ci, [ctypes.c_size_t, ctypes.POINTER(vp)]),
    "mam_host_free": (ci, [vp]),
    "mam_qtranscript_init": (None, [qtrp, ctypes.c_char_p, cl]),
    "mam_qtranscript_append_u64s": (None, [qtrp, ctypes.c_char_p, cl, u64p, cl]),
    "mam_qtranscript_append_ext2": (None, [qtrp, ctypes.c_char_p, cl, u64p]),
    "mam_qtranscript_challenge": (None, [qtrp, ctypes.c_char_p, cl, u64p, cl]),
    "mam_qtranscript_fork": (None, [qtrp, ctypes.c_char_p, cl, u64, qtrp]),
    "mam_qtranscript_fingerprint": (u64, [qtrp]),
    "mam_monolith_q_permute": (None, [u64p]),
    "mam_q_sumcheck_prove": (ci, [qtrp, vp, ci, ci, ci, u64p, u64p, u64p, u64p]),
    "mam_dtranscript_run": (ci, [vp, vp, cl, vp, cl, vp, cl, vp]),
    "mam_poseidon2_wide_permute": (ci, [vp, cl]),
    "mam_profile_collect": (ci, [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(cl), ctypes.POINTER(ctypes.c_double)]),
}

for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(_SIGS)
