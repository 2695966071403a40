// Multilinear (linear) sumcheck kernels: round messages and folds.
//
// Semantics (cuda/sumcheck_prover.cuh:20-292, goldilocks-cuda-rs/cuda/wrapper.cu:1716-1785):
//   message g(c) = sum_y prod_i (p_i[2y] + c (p_i[2y+1] - p_i[2y])) for c = 0..d
//   fold        p_i'[y] = p_i[2y] + r (p_i[2y+1] - p_i[2y])
// Polynomials are packed [d][original_size] (Ext2: interleaved [c0, c1] words).
//
// MI355X design. The reference computes the message of round k and the fold of round k
// in separate passes over the data. The fused kernel k_sc_fold_msg reads the round-k
// arrays once, folds them at the round-k challenge, writes the halved arrays once, and
// accumulates the round-(k+1) message from the pair it just produced; the message is
// reduced inside the launch (wave shuffles, LDS, last-workgroup finish), so the host
// reads back (d+1) Ext2 values per round and never reduces block partials itself.
// Evaluation points c = 0..d are walked with additions (v_{c+1} = v_c + d).
#include "common.h"
#include "profile.h"
#include "reduce.h"
#include "sumcheck.h"

// ---------------------------------------------------------------------------
// Reference-named kernels (per-workgroup partial sums; caller reduces)
// ---------------------------------------------------------------------------
template <bool EXT>
__global__ __launch_bounds__(MAM_BLOCK) void k_round_message_ref(const u64 *__restrict__ polys, u64 *__restrict__ partial,
                                                                 int d, size_t original_size, size_t half) {
    __shared__ u64 lds[(MAM_MAX_POLYS + 1) * 2 * (MAM_BLOCK / WAVE)];
    u64 acc[(MAM_MAX_POLYS + 1) * 2];
    for (int k = 0; k < (MAM_MAX_POLYS + 1) * 2; k++) acc[k] = 0;
    const size_t stride = EXT ? 2 * original_size : original_size;
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half; y += (size_t)gridDim.x * blockDim.x) {
        ext2 e[MAM_MAX_POLYS], df[MAM_MAX_POLYS];
        for (int i = 0; i < d; i++) {
            const u64 *p = polys + i * stride;
            ext2 a = EXT ? e2_load(p, 2 * y) : e2_base(gl_canon(p[2 * y]));
            ext2 b = EXT ? e2_load(p, 2 * y + 1) : e2_base(gl_canon(p[2 * y + 1]));
            e[i] = a;
            df[i] = e2_sub(b, a);
        }
        for (int c = 0; c <= d; c++) {
            ext2 prod = e[0];
            for (int i = 1; i < d; i++) prod = e2_mul(prod, e[i]);
            acc[2 * c] = gl_add(acc[2 * c], prod.c0);
            acc[2 * c + 1] = gl_add(acc[2 * c + 1], prod.c1);
            for (int i = 0; i < d; i++) e[i] = e2_add(e[i], df[i]);
        }
    }
    block_sum<(MAM_MAX_POLYS + 1) * 2>(acc, lds);
    if (threadIdx.x == 0) {
        for (int c = 0; c <= d; c++) {
            if (EXT) {
                partial[((size_t)blockIdx.x * (d + 1) + c) * 2] = acc[2 * c];
                partial[((size_t)blockIdx.x * (d + 1) + c) * 2 + 1] = acc[2 * c + 1];
            } else {
                partial[(size_t)blockIdx.x * (d + 1) + c] = acc[2 * c];
            }
        }
    }
}

template <bool EXT>
__global__ void k_fold_ref(const u64 *__restrict__ in, u64 *__restrict__ out, ext2 ch, int d, size_t original_size,
                           size_t half) {
    const size_t stride = EXT ? 2 * original_size : original_size;
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half; y += (size_t)gridDim.x * blockDim.x) {
        for (int i = 0; i < d; i++) {
            const u64 *p = in + i * stride;
            u64 *q = out + i * stride;
            if (EXT) {
                ext2 a = e2_load(p, 2 * y), b = e2_load(p, 2 * y + 1);
                e2_store(q, y, e2_add(a, e2_mul(ch, e2_sub(b, a))));
            } else {
                u64 a = gl_canon(p[2 * y]), b = gl_canon(p[2 * y + 1]);
                q[y] = gl_add(a, gl_mul(ch.c0, gl_sub(b, a)));
            }
        }
    }
}

MAM_API int sumcheck_round_message_ffi(const u64 *d_polys, u64 *d_partial, int d, size_t original_size, size_t half,
                                       int num_blocks) {
    if (d < 1 || d > MAM_MAX_POLYS || num_blocks <= 0) return -1;
    k_round_message_ref<false><<<num_blocks, MAM_BLOCK, 0, mam_stream()>>>(d_polys, d_partial, d, original_size, half);
    return mam_sync_status();
}
MAM_API int sumcheck_round_message_ext2_ffi(const u64 *d_polys, u64 *d_partial, int d, size_t original_size,
                                            size_t half, int num_blocks) {
    if (d < 1 || d > MAM_MAX_POLYS || num_blocks <= 0) return -1;
    k_round_message_ref<true><<<num_blocks, MAM_BLOCK, 0, mam_stream()>>>(d_polys, d_partial, d, original_size, half);
    return mam_sync_status();
}
MAM_API int sumcheck_fold_ffi(const u64 *d_input, u64 *d_output, u64 challenge, int d, size_t original_size,
                              size_t half) {
    if (d < 1 || half == 0) return half == 0 ? 0 : -1;
    k_fold_ref<false><<<grid_stride(half, MAM_BLOCK, 8192), MAM_BLOCK, 0, mam_stream()>>>(
        d_input, d_output, e2_base(gl_canon(challenge)), d, original_size, half);
    return mam_sync_status();
}
MAM_API int sumcheck_fold_ext2_ffi(const u64 *d_input, u64 *d_output, u64 challenge_c0, u64 challenge_c1, int d,
                                   size_t original_size, size_t half) {
    if (d < 1 || half == 0) return half == 0 ? 0 : -1;
    k_fold_ref<true><<<grid_stride(half, MAM_BLOCK, 8192), MAM_BLOCK, 0, mam_stream()>>>(
        d_input, d_output, e2(gl_canon(challenge_c0), gl_canon(challenge_c1)), d, original_size, half);
    return mam_sync_status();
}

// ---------------------------------------------------------------------------
// Fused prover kernels
// ---------------------------------------------------------------------------

// Round-0 message straight from the input polynomials (base field or Ext2).
template <int NP, bool BASE>
__global__ __launch_bounds__(MAM_BLOCK) void k_sc_msg(const u64 *__restrict__ in, size_t stride, size_t half,
                                                     u64 *partials, u32 *counter, u64 *msg) {
    __shared__ u64 lds[2 * (NP + 1) * (MAM_BLOCK / WAVE)];
    u64 acc[2 * (NP + 1)];
#pragma unroll
    for (int k = 0; k < 2 * (NP + 1); k++) acc[k] = 0;
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half; y += (size_t)gridDim.x * blockDim.x) {
        if (BASE) {
            u64 e[NP], df[NP];
#pragma unroll
            for (int i = 0; i < NP; i++) {
                const u64 *p = in + i * stride;
                u64 a = gl_canon(p[2 * y]), b = gl_canon(p[2 * y + 1]);
                e[i] = a;
                df[i] = gl_sub(b, a);
            }
#pragma unroll
            for (int c = 0; c <= NP; c++) {
                u64 prod = e[0];
#pragma unroll
                for (int i = 1; i < NP; i++) prod = gl_mul(prod, e[i]);
                acc[2 * c] = gl_add(acc[2 * c], prod);
                if (c < NP) {
#pragma unroll
                    for (int i = 0; i < NP; i++) e[i] = gl_add(e[i], df[i]);
                }
            }
        } else {
            ext2 e[NP], df[NP];
#pragma unroll
            for (int i = 0; i < NP; i++) {
                const u64 *p = in + i * stride;
                ext2 a = e2_load(p, 2 * y), b = e2_load(p, 2 * y + 1);
                e[i] = a;
                df[i] = e2_sub(b, a);
            }
#pragma unroll
            for (int c = 0; c <= NP; c++) {
                ext2 prod = e[0];
#pragma unroll
                for (int i = 1; i < NP; i++) prod = e2_mul(prod, e[i]);
                acc[2 * c] = gl_add(acc[2 * c], prod.c0);
                acc[2 * c + 1] = gl_add(acc[2 * c + 1], prod.c1);
                if (c < NP) {
#pragma unroll
                    for (int i = 0; i < NP; i++) e[i] = e2_add(e[i], df[i]);
                }
            }
        }
    }
    block_sum<2 * (NP + 1)>(acc, lds);
    grid_finish<2 * (NP + 1)>(acc, lds, partials, counter, msg);
}

// Fold the current arrays (each 4*half_out elements) at `ch` into `out` (2*half_out
// Ext2 elements each) and accumulate the next round's message from each output pair.
template <int NP, bool BASE_IN>
__global__ __launch_bounds__(MAM_BLOCK) void k_sc_fold_msg(const u64 *__restrict__ in, size_t in_stride,
                                                          u64 *__restrict__ out, size_t out_stride, size_t half_out,
                                                          ext2 ch, u64 *partials, u32 *counter, u64 *msg) {
    __shared__ u64 lds[2 * (NP + 1) * (MAM_BLOCK / WAVE)];
    u64 acc[2 * (NP + 1)];
#pragma unroll
    for (int k = 0; k < 2 * (NP + 1); k++) acc[k] = 0;
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half_out;
         y += (size_t)gridDim.x * blockDim.x) {
        ext2 e[NP], df[NP];
#pragma unroll
        for (int i = 0; i < NP; i++) {
            const u64 *p = in + i * in_stride;
            ext2 f0, f1;
            if (BASE_IN) {
                // 4 consecutive base words; ch * (b - a) costs two base products
                const u64 a0 = gl_canon(p[4 * y]), a1 = gl_canon(p[4 * y + 1]);
                const u64 a2 = gl_canon(p[4 * y + 2]), a3 = gl_canon(p[4 * y + 3]);
                f0 = e2_add(e2_base(a0), e2_scale(gl_sub(a1, a0), ch));
                f1 = e2_add(e2_base(a2), e2_scale(gl_sub(a3, a2), ch));
            } else {
                const ext2 a0 = e2_load(p, 4 * y), a1 = e2_load(p, 4 * y + 1);
                const ext2 a2 = e2_load(p, 4 * y + 2), a3 = e2_load(p, 4 * y + 3);
                f0 = e2_add(a0, e2_mul(ch, e2_sub(a1, a0)));
                f1 = e2_add(a2, e2_mul(ch, e2_sub(a3, a2)));
            }
            u64 *q = out + i * out_stride;
            e2_store(q, 2 * y, f0);
            e2_store(q, 2 * y + 1, f1);
            e[i] = f0;
            df[i] = e2_sub(f1, f0);
        }
#pragma unroll
        for (int c = 0; c <= NP; c++) {
            ext2 prod = e[0];
#pragma unroll
            for (int i = 1; i < NP; i++) prod = e2_mul(prod, e[i]);
            acc[2 * c] = gl_add(acc[2 * c], prod.c0);
            acc[2 * c + 1] = gl_add(acc[2 * c + 1], prod.c1);
            if (c < NP) {
#pragma unroll
                for (int i = 0; i < NP; i++) e[i] = e2_add(e[i], df[i]);
            }
        }
    }
    block_sum<2 * (NP + 1)>(acc, lds);
    grid_finish<2 * (NP + 1)>(acc, lds, partials, counter, msg);
}

// Last fold: two elements per polynomial -> one final value each, written to `finals`.
template <bool BASE_IN>
__global__ void k_sc_fold_final(const u64 *__restrict__ in, size_t in_stride, int np, ext2 ch, u64 *finals) {
    int i = threadIdx.x;
    if (i >= np) return;
    const u64 *p = in + i * in_stride;
    ext2 a = BASE_IN ? e2_base(gl_canon(p[0])) : e2_load(p, 0);
    ext2 b = BASE_IN ? e2_base(gl_canon(p[1])) : e2_load(p, 1);
    e2_store(finals, i, e2_add(a, e2_mul(ch, e2_sub(b, a))));
}

template <int NP>
static int launch_msg(const u64 *in, size_t stride, size_t half, bool base, const ScScratch &s, hipStream_t st) {
    unsigned g = grid_reduce(half);
    if (base)
        k_sc_msg<NP, true><<<g, MAM_BLOCK, 0, st>>>(in, stride, half, s.partials, s.counter, s.msg);
    else
        k_sc_msg<NP, false><<<g, MAM_BLOCK, 0, st>>>(in, stride, half, s.partials, s.counter, s.msg);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NP>
static int launch_fold_msg(const u64 *in, size_t in_stride, u64 *out, size_t out_stride, size_t half_out, bool base_in,
                           ext2 ch, const ScScratch &s, hipStream_t st) {
    unsigned g = grid_reduce(half_out);
    // per output pair and polynomial: read 4 inputs (base or Ext2), write 2 Ext2
    ProfScope prof(PROF_SUMCHECK, (double)half_out * NP * ((base_in ? 32.0 : 64.0) + 32.0), st);
    if (base_in)
        k_sc_fold_msg<NP, true><<<g, MAM_BLOCK, 0, st>>>(in, in_stride, out, out_stride, half_out, ch, s.partials,
                                                         s.counter, s.msg);
    else
        k_sc_fold_msg<NP, false><<<g, MAM_BLOCK, 0, st>>>(in, in_stride, out, out_stride, half_out, ch, s.partials,
                                                          s.counter, s.msg);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mam_sc_launch_msg(int np, const u64 *in, size_t stride, size_t half, bool base, const ScScratch &s,
                      hipStream_t st) {
    switch (np) {
        case 1: return launch_msg<1>(in, stride, half, base, s, st);
        case 2: return launch_msg<2>(in, stride, half, base, s, st);
        case 3: return launch_msg<3>(in, stride, half, base, s, st);
        case 4: return launch_msg<4>(in, stride, half, base, s, st);
        case 5: return launch_msg<5>(in, stride, half, base, s, st);
        case 6: return launch_msg<6>(in, stride, half, base, s, st);
        case 7: return launch_msg<7>(in, stride, half, base, s, st);
        case 8: return launch_msg<8>(in, stride, half, base, s, st);
        default: return -2;
    }
}

int mam_sc_launch_fold_msg(int np, const u64 *in, size_t in_stride, u64 *out, size_t out_stride, size_t half_out,
                           bool base_in, ext2 ch, const ScScratch &s, hipStream_t st) {
    switch (np) {
        case 1: return launch_fold_msg<1>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 2: return launch_fold_msg<2>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 3: return launch_fold_msg<3>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 4: return launch_fold_msg<4>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 5: return launch_fold_msg<5>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 6: return launch_fold_msg<6>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 7: return launch_fold_msg<7>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 8: return launch_fold_msg<8>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        default: return -2;
    }
}

int mam_sc_launch_fold_final(int np, const u64 *in, size_t in_stride, bool base_in, ext2 ch, u64 *finals,
                             hipStream_t st) {
    if (np < 1 || np > MAM_MAX_POLYS) return -2;
    if (base_in)
        k_sc_fold_final<true><<<1, 64, 0, st>>>(in, in_stride, np, ch, finals);
    else
        k_sc_fold_final<false><<<1, 64, 0, st>>>(in, in_stride, np, ch, finals);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
