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
#include <stdlib.h>

#include "common.h"
#include "dtranscript.h"
#include "profile.h"
#include "reduce.h"
#include "sumcheck.h"
#include "sync.h"

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

// The kernels below take the field policy F (FieldGL: the Goldilocks prover; FieldQ: the
// almost-Goldilocks prover of zk-torch-4), E = F::E its Ext2.

// Round-0 message straight from the input polynomials (base field or Ext2).
template <class F, int NP, bool BASE>
__global__ __launch_bounds__(MAM_BLOCK) void k_sc_msg(const u64 *__restrict__ in, size_t stride, size_t half,
                                                     u64 *partials, u64 *msg) {
    using E = typename F::E;
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
                u64 a = F::canon(p[2 * y]), b = F::canon(p[2 * y + 1]);
                e[i] = a;
                df[i] = F::sub(b, a);
            }
#pragma unroll
            for (int c = 0; c <= NP; c++) {
                u64 prod = e[0];
#pragma unroll
                for (int i = 1; i < NP; i++) prod = F::mul(prod, e[i]);
                acc[2 * c] = F::add(acc[2 * c], prod);
                if (c < NP) {
#pragma unroll
                    for (int i = 0; i < NP; i++) e[i] = F::add(e[i], df[i]);
                }
            }
        } else {
            E e[NP], df[NP];
#pragma unroll
            for (int i = 0; i < NP; i++) {
                const u64 *p = in + i * stride;
                E a = F::e_load(p, 2 * y), b = F::e_load(p, 2 * y + 1);
                e[i] = a;
                df[i] = F::e_sub(b, a);
            }
#pragma unroll
            for (int c = 0; c <= NP; c++) {
                E prod = e[0];
#pragma unroll
                for (int i = 1; i < NP; i++) prod = F::e_mul(prod, e[i]);
                acc[2 * c] = F::add(acc[2 * c], prod.c0);
                acc[2 * c + 1] = F::add(acc[2 * c + 1], prod.c1);
                if (c < NP) {
#pragma unroll
                    for (int i = 0; i < NP; i++) e[i] = F::e_add(e[i], df[i]);
                }
            }
        }
    }
    block_sum<2 * (NP + 1), F>(acc, lds);
    grid_partials<2 * (NP + 1)>(acc, partials, msg);
}

// Fold the current arrays (each 4*half_out elements) at `ch` into `out` (2*half_out
// Ext2 elements each) and accumulate the next round's message from each output pair.
template <class F, int NP, bool BASE_IN>
__global__ __launch_bounds__(MAM_BLOCK) void k_sc_fold_msg(const u64 *__restrict__ in, size_t in_stride,
                                                          u64 *__restrict__ out, size_t out_stride, size_t half_out,
                                                          typename F::E ch, u64 *partials, u64 *msg) {
    using E = typename F::E;
    __shared__ u64 lds[2 * (NP + 1) * (MAM_BLOCK / WAVE)];
    u64 acc[2 * (NP + 1)];
#pragma unroll
    for (int k = 0; k < 2 * (NP + 1); k++) acc[k] = 0;
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half_out;
         y += (size_t)gridDim.x * blockDim.x) {
        E e[NP], df[NP];
#pragma unroll
        for (int i = 0; i < NP; i++) {
            const u64 *p = in + i * in_stride;
            E f0, f1;
            if (BASE_IN) {
                // 4 consecutive base words; ch * (b - a) costs two base products
                const u64 a0 = F::canon(p[4 * y]), a1 = F::canon(p[4 * y + 1]);
                const u64 a2 = F::canon(p[4 * y + 2]), a3 = F::canon(p[4 * y + 3]);
                f0 = F::e_add(F::e_base(a0), F::e_scale(F::sub(a1, a0), ch));
                f1 = F::e_add(F::e_base(a2), F::e_scale(F::sub(a3, a2), ch));
            } else {
                const E a0 = F::e_load(p, 4 * y), a1 = F::e_load(p, 4 * y + 1);
                const E a2 = F::e_load(p, 4 * y + 2), a3 = F::e_load(p, 4 * y + 3);
                f0 = F::e_add(a0, F::e_mul(ch, F::e_sub(a1, a0)));
                f1 = F::e_add(a2, F::e_mul(ch, F::e_sub(a3, a2)));
            }
            u64 *q = out + i * out_stride;
            F::e_store(q, 2 * y, f0);
            F::e_store(q, 2 * y + 1, f1);
            e[i] = f0;
            df[i] = F::e_sub(f1, f0);
        }
#pragma unroll
        for (int c = 0; c <= NP; c++) {
            E prod = e[0];
#pragma unroll
            for (int i = 1; i < NP; i++) prod = F::e_mul(prod, e[i]);
            acc[2 * c] = F::add(acc[2 * c], prod.c0);
            acc[2 * c + 1] = F::add(acc[2 * c + 1], prod.c1);
            if (c < NP) {
#pragma unroll
                for (int i = 0; i < NP; i++) e[i] = F::e_add(e[i], df[i]);
            }
        }
    }
    block_sum<2 * (NP + 1), F>(acc, lds);
    grid_partials<2 * (NP + 1)>(acc, partials, msg);
}

// Last fold: two elements per polynomial -> one final value each, written to `finals`.
template <class F, bool BASE_IN>
__global__ void k_sc_fold_final(const u64 *__restrict__ in, size_t in_stride, int np, typename F::E ch, u64 *finals) {
    using E = typename F::E;
    int i = threadIdx.x;
    if (i >= np) return;
    const u64 *p = in + i * in_stride;
    E a = BASE_IN ? F::e_base(F::canon(p[0])) : F::e_load(p, 0);
    E b = BASE_IN ? F::e_base(F::canon(p[1])) : F::e_load(p, 1);
    F::e_store(finals, i, F::e_add(a, F::e_mul(ch, F::e_sub(b, a))));
}

template <class F, int NP>
static int launch_msg(const u64 *in, size_t stride, size_t half, bool base, const ScScratch &s, hipStream_t st) {
    unsigned g = grid_reduce(half);
    if (base)
        k_sc_msg<F, NP, true><<<g, MAM_BLOCK, 0, st>>>(in, stride, half, s.partials, s.msg);
    else
        k_sc_msg<F, NP, false><<<g, MAM_BLOCK, 0, st>>>(in, stride, half, s.partials, s.msg);
    launch_sum_partials<2 * (NP + 1), F>(s.partials, g, s.msg, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <class F, int NP>
static int launch_fold_msg(const u64 *in, size_t in_stride, u64 *out, size_t out_stride, size_t half_out, bool base_in,
                           typename F::E ch, const ScScratch &s, hipStream_t st) {
    unsigned g = grid_reduce(half_out);
    // per output pair and polynomial: read 4 inputs (base or Ext2), write 2 Ext2
    ProfScope prof(PROF_SUMCHECK, (double)half_out * NP * ((base_in ? 32.0 : 64.0) + 32.0), st);
    if (base_in)
        k_sc_fold_msg<F, NP, true><<<g, MAM_BLOCK, 0, st>>>(in, in_stride, out, out_stride, half_out, ch, s.partials,
                                                            s.msg);
    else
        k_sc_fold_msg<F, NP, false><<<g, MAM_BLOCK, 0, st>>>(in, in_stride, out, out_stride, half_out, ch, s.partials,
                                                             s.msg);
    launch_sum_partials<2 * (NP + 1), F>(s.partials, g, s.msg, st);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <class F>
static int sc_launch_msg(int np, const u64 *in, size_t stride, size_t half, bool base, const ScScratch &s,
                         hipStream_t st) {
    switch (np) {
        case 1: return launch_msg<F, 1>(in, stride, half, base, s, st);
        case 2: return launch_msg<F, 2>(in, stride, half, base, s, st);
        case 3: return launch_msg<F, 3>(in, stride, half, base, s, st);
        case 4: return launch_msg<F, 4>(in, stride, half, base, s, st);
        case 5: return launch_msg<F, 5>(in, stride, half, base, s, st);
        case 6: return launch_msg<F, 6>(in, stride, half, base, s, st);
        case 7: return launch_msg<F, 7>(in, stride, half, base, s, st);
        case 8: return launch_msg<F, 8>(in, stride, half, base, s, st);
        default: return -2;
    }
}

template <class F>
static int sc_launch_fold_msg(int np, const u64 *in, size_t in_stride, u64 *out, size_t out_stride, size_t half_out,
                              bool base_in, typename F::E ch, const ScScratch &s, hipStream_t st) {
    switch (np) {
        case 1: return launch_fold_msg<F, 1>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 2: return launch_fold_msg<F, 2>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 3: return launch_fold_msg<F, 3>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 4: return launch_fold_msg<F, 4>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 5: return launch_fold_msg<F, 5>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 6: return launch_fold_msg<F, 6>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 7: return launch_fold_msg<F, 7>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        case 8: return launch_fold_msg<F, 8>(in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
        default: return -2;
    }
}

template <class F>
static int sc_launch_fold_final(int np, const u64 *in, size_t in_stride, bool base_in, typename F::E ch, u64 *finals,
                                hipStream_t st) {
    if (np < 1 || np > MAM_MAX_POLYS) return -2;
    if (base_in)
        k_sc_fold_final<F, true><<<1, 64, 0, st>>>(in, in_stride, np, ch, finals);
    else
        k_sc_fold_final<F, false><<<1, 64, 0, st>>>(in, in_stride, np, ch, finals);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mam_sc_launch_msg(int np, const u64 *in, size_t stride, size_t half, bool base, const ScScratch &s,
                      hipStream_t st) {
    return sc_launch_msg<FieldGL>(np, in, stride, half, base, s, st);
}
int mam_sc_launch_fold_msg(int np, const u64 *in, size_t in_stride, u64 *out, size_t out_stride, size_t half_out,
                           bool base_in, ext2 ch, const ScScratch &s, hipStream_t st) {
    return sc_launch_fold_msg<FieldGL>(np, in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
}
int mam_sc_launch_fold_final(int np, const u64 *in, size_t in_stride, bool base_in, ext2 ch, u64 *finals,
                             hipStream_t st) {
    return sc_launch_fold_final<FieldGL>(np, in, in_stride, base_in, ch, finals, st);
}
int mam_qsc_launch_msg(int np, const u64 *in, size_t stride, size_t half, bool base, const ScScratch &s,
                       hipStream_t st) {
    return sc_launch_msg<FieldQ>(np, in, stride, half, base, s, st);
}
int mam_qsc_launch_fold_msg(int np, const u64 *in, size_t in_stride, u64 *out, size_t out_stride, size_t half_out,
                            bool base_in, qe2 ch, const ScScratch &s, hipStream_t st) {
    return sc_launch_fold_msg<FieldQ>(np, in, in_stride, out, out_stride, half_out, base_in, ch, s, st);
}
int mam_qsc_launch_fold_final(int np, const u64 *in, size_t in_stride, bool base_in, qe2 ch, u64 *finals,
                              hipStream_t st) {
    return sc_launch_fold_final<FieldQ>(np, in, in_stride, base_in, ch, finals, st);
}

// ---------------------------------------------------------------------------
// Two rounds per device pass (whole-proof prover, up to 3 polynomials).
//
// The round-(r+1) message is a polynomial in the round-r challenge: with the round-r
// arrays grouped in fours (e0, e1, e2, e3), folding at ch gives the pair
// (e0 + ch (e1 - e0), e2 + ch (e3 - e2)), and its factor at point c is
//   u(c) + ch v(c),  u(c) = e0 + c (e2 - e0),  v(c) = (e1 - e0) + c ((e3 - e2) - (e1 - e0)).
// One pass over the round-r arrays therefore yields the round-r message (from the pairs
// (e0, e1), (e2, e3)) and the coefficients sym[c][k] of ch^k in the round-(r+1) message.
// The host evaluates the round-(r+1) message itself once it has drawn ch_r, draws ch_{r+1},
// and the next pass applies both challenges on read (four input elements per element).
// The transcript is unchanged; the device is visited once per two rounds.
//
// k_sc_pass<NP, BASE_IN, FOLDS, LOOK>: FOLDS (0 or 2) pending challenges are applied on
// read; each item is a group of G = 4 (LOOK) or 2 elements of the folded arrays, which are
// written to `out` when FOLDS > 0. Reduced words: msg[c] (NP + 1 Ext2), then with LOOK
// sym[c][k] (c, k = 0..NP).
// ---------------------------------------------------------------------------
template <int NP, bool LOOK>
struct PassWords {
    static constexpr int value = 2 * (NP + 1) + (LOOK ? 2 * (NP + 1) * (NP + 1) : 0);
};

// The pending challenges from device memory (a pre-queued pass): ca at chp[0..1], cb at
// chp[2..3].
__device__ __forceinline__ void pass_challenges(const u64 *chp, ext2 &ca, ext2 &cb) {
    ca = ext2{gl_canon(chp[0]), gl_canon(chp[1])};
    cb = ext2{gl_canon(chp[2]), gl_canon(chp[3])};
}

// ca7 = 7 ca.c1, cb7 = 7 cb.c1 (formed once per thread, see e2_mul_fixed)
template <bool BASE_IN, int FOLDS>
__device__ __forceinline__ ext2 pass_load(const u64 *p, size_t j, ext2 ca, u64 ca7, ext2 cb, u64 cb7) {
    if (FOLDS == 0) return BASE_IN ? e2_base(gl_canon(p[j])) : e2_load(p, j);
    ext2 v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = BASE_IN ? e2_base(gl_canon(p[4 * j + k])) : e2_load(p, 4 * j + k);
    const ext2 f0 = e2_add(v[0], e2_mul_fixed(ca, ca7, e2_sub(v[1], v[0])));
    const ext2 f1 = e2_add(v[2], e2_mul_fixed(ca, ca7, e2_sub(v[3], v[2])));
    return e2_add(f0, e2_mul_fixed(cb, cb7, e2_sub(f1, f0)));
}

// LDS words of sc_pass_body's reduction (block_sum_lds up to 24 words, shuffles above)
template <int W>
struct PassLds {
    static constexpr int value = W <= 24 ? block_sum_lds_words(W, MAM_BLOCK) : W * (MAM_BLOCK / WAVE);
};

// The pass as workgroup `bid` of `nblk` (the grid-stride loop runs over the nblk workgroups
// given); the workgroup's W reduced words are left in acc of thread 0 after block_sum.
template <int NP, bool BASE_IN, int FOLDS, bool LOOK>
__device__ __forceinline__ void sc_pass_body(const u64 *__restrict__ in, size_t in_stride, u64 *__restrict__ out,
                                             size_t out_stride, size_t items, ext2 ca, ext2 cb, unsigned bid,
                                             unsigned nblk, u64 (&acc)[PassWords<NP, LOOK>::value], u64 *lds) {
    constexpr int G = LOOK ? 4 : 2;
    constexpr int W = PassWords<NP, LOOK>::value;
#pragma unroll
    for (int k = 0; k < W; k++) acc[k] = 0;
    const u64 ca7 = gl_mul7(ca.c1), cb7 = gl_mul7(cb.c1);
    if constexpr (BASE_IN && FOLDS == 0) {
        // unfolded base inputs: every product is a base-field product (the c1 words stay 0)
        for (size_t z = bid * (size_t)blockDim.x + threadIdx.x; z < items; z += (size_t)nblk * blockDim.x) {
            u64 e[NP][G];
#pragma unroll
            for (int i = 0; i < NP; i++) {
#pragma unroll
                for (int t = 0; t < G; t++) e[i][t] = gl_canon(in[i * in_stride + G * z + t]);
            }
#pragma unroll
            for (int q = 0; q < G / 2; q++) {
                u64 v[NP], d[NP];
#pragma unroll
                for (int i = 0; i < NP; i++) {
                    v[i] = e[i][2 * q];
                    d[i] = gl_sub(e[i][2 * q + 1], e[i][2 * q]);
                }
#pragma unroll
                for (int c = 0; c <= NP; c++) {
                    u64 prod = v[0];
#pragma unroll
                    for (int i = 1; i < NP; i++) prod = gl_mul(prod, v[i]);
                    acc[2 * c] = gl_add(acc[2 * c], prod);
                    if (c < NP) {
#pragma unroll
                        for (int i = 0; i < NP; i++) v[i] = gl_add(v[i], d[i]);
                    }
                }
            }
            if (LOOK) {
                u64 u[NP], du[NP], w[NP], dw[NP];
#pragma unroll
                for (int i = 0; i < NP; i++) {
                    const u64 d01 = gl_sub(e[i][1], e[i][0]), d23 = gl_sub(e[i][G - 1], e[i][G - 2]);
                    u[i] = e[i][0];
                    du[i] = gl_sub(e[i][G - 2], e[i][0]);
                    w[i] = d01;
                    dw[i] = gl_sub(d23, d01);
                }
#pragma unroll
                for (int c = 0; c <= NP; c++) {
                    u64 poly[NP + 1];
                    poly[0] = u[0];
                    poly[1] = w[0];
#pragma unroll
                    for (int i = 1; i < NP; i++) {
                        poly[i + 1] = gl_mul(poly[i], w[i]);
#pragma unroll
                        for (int k = i; k >= 1; k--) poly[k] = gl_add(gl_mul(poly[k], u[i]), gl_mul(poly[k - 1], w[i]));
                        poly[0] = gl_mul(poly[0], u[i]);
                    }
#pragma unroll
                    for (int k = 0; k <= NP; k++) {
                        const int o = 2 * (NP + 1) + 2 * (c * (NP + 1) + k);
                        acc[o] = gl_add(acc[o], poly[k]);
                    }
                    if (c < NP) {
#pragma unroll
                        for (int i = 0; i < NP; i++) {
                            u[i] = gl_add(u[i], du[i]);
                            w[i] = gl_add(w[i], dw[i]);
                        }
                    }
                }
            }
        }
    } else
    for (size_t z = bid * (size_t)blockDim.x + threadIdx.x; z < items; z += (size_t)nblk * blockDim.x) {
        ext2 e[NP][G];
#pragma unroll
        for (int i = 0; i < NP; i++) {
            const u64 *p = in + i * in_stride;
#pragma unroll
            for (int t = 0; t < G; t++) e[i][t] = pass_load<BASE_IN, FOLDS>(p, G * z + t, ca, ca7, cb, cb7);
            if (FOLDS > 0) {
#pragma unroll
                for (int t = 0; t < G; t++) e2_store(out + i * out_stride, G * z + t, e[i][t]);
            }
        }
        // round-r message over the G / 2 pairs of the group
#pragma unroll
        for (int q = 0; q < G / 2; q++) {
            ext2 v[NP], d[NP];
#pragma unroll
            for (int i = 0; i < NP; i++) {
                v[i] = e[i][2 * q];
                d[i] = e2_sub(e[i][2 * q + 1], e[i][2 * q]);
            }
#pragma unroll
            for (int c = 0; c <= NP; c++) {
                ext2 prod = v[0];
#pragma unroll
                for (int i = 1; i < NP; i++) prod = e2_mul(prod, v[i]);
                acc[2 * c] = gl_add(acc[2 * c], prod.c0);
                acc[2 * c + 1] = gl_add(acc[2 * c + 1], prod.c1);
                if (c < NP) {
#pragma unroll
                    for (int i = 0; i < NP; i++) v[i] = e2_add(v[i], d[i]);
                }
            }
        }
        if (LOOK) {
            // sym[c][k]: coefficient of ch^k in prod_i (u_i(c) + ch v_i(c))
            ext2 u[NP], du[NP], w[NP], dw[NP];
#pragma unroll
            for (int i = 0; i < NP; i++) {
                const ext2 d01 = e2_sub(e[i][1], e[i][0]), d23 = e2_sub(e[i][G - 1], e[i][G - 2]);
                u[i] = e[i][0];
                du[i] = e2_sub(e[i][G - 2], e[i][0]);
                w[i] = d01;
                dw[i] = e2_sub(d23, d01);
            }
#pragma unroll
            for (int c = 0; c <= NP; c++) {
                ext2 poly[NP + 1];
                poly[0] = u[0];
                poly[1] = w[0];
#pragma unroll
                for (int i = 1; i < NP; i++) {
                    poly[i + 1] = e2_mul(poly[i], w[i]);
#pragma unroll
                    for (int k = i; k >= 1; k--) poly[k] = e2_add(e2_mul(poly[k], u[i]), e2_mul(poly[k - 1], w[i]));
                    poly[0] = e2_mul(poly[0], u[i]);
                }
#pragma unroll
                for (int k = 0; k <= NP; k++) {
                    const int o = 2 * (NP + 1) + 2 * (c * (NP + 1) + k);
                    acc[o] = gl_add(acc[o], poly[k].c0);
                    acc[o + 1] = gl_add(acc[o + 1], poly[k].c1);
                }
                if (c < NP) {
#pragma unroll
                    for (int i = 0; i < NP; i++) {
                        u[i] = e2_add(u[i], du[i]);
                        w[i] = e2_add(w[i], dw[i]);
                    }
                }
            }
        }
    }
    if constexpr (BASE_IN && FOLDS == 0) {
        // base products only: the c1 words are zero, so only the c0 words are reduced
        u64 c0w[W / 2];
#pragma unroll
        for (int k = 0; k < W / 2; k++) c0w[k] = acc[2 * k];
        block_sum_lds<W / 2>(c0w, lds);
#pragma unroll
        for (int k = 0; k < W / 2; k++) {
            acc[2 * k] = c0w[k];
            acc[2 * k + 1] = 0;
        }
    } else if constexpr (W <= 24) {
        block_sum_lds<W>(acc, lds);
    } else {
        block_sum<W>(acc, lds);  // three polynomials: the LDS rows would take 80 KB
    }
}

template <int NP, bool BASE_IN, int FOLDS, bool LOOK>
__global__ __launch_bounds__(MAM_BLOCK) void k_sc_pass(const u64 *__restrict__ in, size_t in_stride,
                                                      u64 *__restrict__ out, size_t out_stride, size_t items,
                                                      ext2 ca, ext2 cb, const u64 *chp, u64 *partials, u64 *res) {
    constexpr int W = PassWords<NP, LOOK>::value;
    __shared__ u64 lds[PassLds<W>::value];
    u64 acc[W];
    if (chp) pass_challenges(chp, ca, cb);
    sc_pass_body<NP, BASE_IN, FOLDS, LOOK>(in, in_stride, out, out_stride, items, ca, cb, blockIdx.x, gridDim.x, acc,
                                           lds);
    grid_partials<W>(acc, partials, res);
}

// ---------------------------------------------------------------------------
// k_sc_pass_lp: the same pass for small arrays, one item spread over LP_LANES lanes. In
// k_sc_pass one thread carries a whole item (the folds of NP x G inputs, the message products
// and the round-(r+1) coefficient products, about 10^4 instructions for NP = 2), so a pass
// over a few thousand items costs that chain's latency however few the items are. Here lane
// (i, t) of an item forms input e[i][t], then lane k forms one product task: the message
// term of pair q at point c (k = q (NP + 1) + c), or the round-(r+1) polynomial in ch at
// point c (k = G / 2 (NP + 1) + c). Each lane keeps the sums of its own task; the lanes of
// one task are added across the wave, the workgroup and the grid into the same words as
// k_sc_pass writes.
// ---------------------------------------------------------------------------
static constexpr int LP_LANES = 16;
static constexpr int LP_ITEMS = MAM_BLOCK / LP_LANES;

template <int NP, bool LOOK>
struct LpTasks {
    static constexpr int G = LOOK ? 4 : 2;
    static constexpr int MSG = (G / 2) * (NP + 1);
    static constexpr int ALL = MSG + (LOOK ? NP + 1 : 0);
    static constexpr int WORDS = LOOK ? 2 * (NP + 1) : 2;  // accumulator words per lane
};

template <int NP, bool BASE_IN, int FOLDS, bool LOOK>
__global__ __launch_bounds__(MAM_BLOCK) void k_sc_pass_lp(const u64 *__restrict__ in, size_t in_stride,
                                                         u64 *__restrict__ out, size_t out_stride, size_t items,
                                                         ext2 ca, ext2 cb, const u64 *chp, u64 *partials,
                                                         u64 *res) {
    using T = LpTasks<NP, LOOK>;
    constexpr int G = T::G, AW = T::WORDS, W = PassWords<NP, LOOK>::value;
    static_assert(NP * G <= LP_LANES && T::ALL <= LP_LANES, "an item must fit its lanes");
    static_assert(W <= MAM_BLOCK, "one thread per reduced word");
    __shared__ u64 ev[LP_ITEMS][NP * G][2];
    __shared__ u64 red[MAM_BLOCK / WAVE][LP_LANES][AW];
    const int sub = threadIdx.x % LP_LANES, slot = threadIdx.x / LP_LANES;
    u64 acc[AW];
#pragma unroll
    for (int k = 0; k < AW; k++) acc[k] = 0;
    if (chp) pass_challenges(chp, ca, cb);
    const u64 ca7 = gl_mul7(ca.c1), cb7 = gl_mul7(cb.c1);
    for (size_t base = (size_t)blockIdx.x * LP_ITEMS; base < items; base += (size_t)gridDim.x * LP_ITEMS) {
        const size_t z = base + slot;
        const bool live = z < items;
        if (live && sub < NP * G) {
            const int i = sub / G, t = sub % G;
            const ext2 e = pass_load<BASE_IN, FOLDS>(in + i * in_stride, G * z + t, ca, ca7, cb, cb7);
            if (FOLDS > 0) e2_store(out + i * out_stride, G * z + t, e);
            ev[slot][sub][0] = e.c0;
            ev[slot][sub][1] = e.c1;
        }
        mam_barrier();
        if (live && sub < T::ALL) {
            auto E = [&](int i, int t) { return ext2{ev[slot][i * G + t][0], ev[slot][i * G + t][1]}; };
            if (sub < T::MSG) {
                const int q = sub / (NP + 1), c = sub % (NP + 1);
                ext2 prod{0, 0};
#pragma unroll
                for (int i = 0; i < NP; i++) {
                    const ext2 a = E(i, 2 * q), d = e2_sub(E(i, 2 * q + 1), a);
                    ext2 v = a;
                    for (int k = 0; k < c; k++) v = e2_add(v, d);
                    prod = i == 0 ? v : e2_mul(prod, v);
                }
                acc[0] = gl_add(acc[0], prod.c0);
                acc[1] = gl_add(acc[1], prod.c1);
            } else if (LOOK) {
                // coefficients of ch^k in prod_i (u_i(c) + ch w_i(c)), as in k_sc_pass
                const int c = sub - T::MSG;
                ext2 u[NP], w[NP];
#pragma unroll
                for (int i = 0; i < NP; i++) {
                    const ext2 e0 = E(i, 0), d01 = e2_sub(E(i, 1), e0);
                    const ext2 du = e2_sub(E(i, G - 2), e0), dw = e2_sub(e2_sub(E(i, G - 1), E(i, G - 2)), d01);
                    u[i] = e0;
                    w[i] = d01;
                    for (int k = 0; k < c; k++) {
                        u[i] = e2_add(u[i], du);
                        w[i] = e2_add(w[i], dw);
                    }
                }
                ext2 poly[NP + 1];
                poly[0] = u[0];
                poly[1] = w[0];
#pragma unroll
                for (int i = 1; i < NP; i++) {
                    poly[i + 1] = e2_mul(poly[i], w[i]);
#pragma unroll
                    for (int k = i; k >= 1; k--) poly[k] = e2_add(e2_mul(poly[k], u[i]), e2_mul(poly[k - 1], w[i]));
                    poly[0] = e2_mul(poly[0], u[i]);
                }
#pragma unroll
                for (int k = 0; k <= NP; k++) {
                    acc[2 * k] = gl_add(acc[2 * k], poly[k].c0);
                    acc[2 * k + 1] = gl_add(acc[2 * k + 1], poly[k].c1);
                }
            }
        }
        mam_barrier();
    }
    // lanes of one task: across the wave (lane offsets 16, 32), then across the waves
#pragma unroll
    for (int k = 0; k < AW; k++) {
        acc[k] = gl_add(acc[k], shfl_xor_u64(acc[k], 16));
        acc[k] = gl_add(acc[k], shfl_xor_u64(acc[k], 32));
    }
    const int lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    if (lane < LP_LANES) {
#pragma unroll
        for (int k = 0; k < AW; k++) red[wave][lane][k] = acc[k];
    }
    mam_barrier();
    // thread t < W forms word t of the workgroup's sums: msg[c] (t = 2 c + h), then
    // sym[c][k] (t = 2 (NP + 1) + 2 (c (NP + 1) + k) + h)
    u64 word = 0;
    if ((int)threadIdx.x < W) {
        const int t = threadIdx.x;
        for (int wv = 0; wv < MAM_BLOCK / WAVE; wv++) {
            if (t < 2 * (NP + 1)) {
                for (int q = 0; q < G / 2; q++) word = gl_add(word, red[wv][q * (NP + 1) + t / 2][t % 2]);
            } else {
                const int o = t - 2 * (NP + 1);
                word = gl_add(word, red[wv][T::MSG + o / (2 * (NP + 1))][o % (2 * (NP + 1))]);
            }
        }
    }
    if ((int)threadIdx.x < W) (gridDim.x == 1 ? res : partials + (size_t)blockIdx.x * W)[threadIdx.x] = word;
}

// Workgroups of a pass launch: MAM_SC_GRID (default MAM_RED_GRID), at most what the partials
// scratch holds for W reduced words per workgroup.
static unsigned pass_grid_cap(int w) {
    static const long v = [] {
        const char *e = getenv("MAM_SC_GRID");
        const long x = e ? strtol(e, nullptr, 10) : (long)MAM_RED_GRID;
        return x < 1 ? 1L : x;
    }();
    const long room = (long)MAM_RED_MAX_BLOCKS * 2 * (MAM_MAX_POLYS + 1) / w;
    return (unsigned)(v < room ? v : room);
}

// Largest item count that runs on k_sc_pass_lp (MAM_SC_LP_MAX overrides; 0 disables it).
static size_t lp_max_items() {
    static const long v = [] {
        const char *e = getenv("MAM_SC_LP_MAX");
        const long x = e ? strtol(e, nullptr, 10) : (1L << 14);
        return x < 0 ? 0L : x;
    }();
    return (size_t)v;
}

// The final values: arrays of 2^FOLDS elements folded at the pending challenges (one
// thread per polynomial).
template <bool BASE_IN, int FOLDS>
__global__ void k_sc_pass_final(const u64 *__restrict__ in, size_t in_stride, int np, ext2 ca, ext2 cb,
                                const u64 *chp, u64 *finals) {
    const int i = threadIdx.x;
    if (i >= np) return;
    if (chp) pass_challenges(chp, ca, cb);
    const u64 *p = in + i * in_stride;
    ext2 v;
    if (FOLDS == 2) {
        v = pass_load<BASE_IN, 2>(p, 0, ca, gl_mul7(ca.c1), cb, gl_mul7(cb.c1));
    } else {
        const ext2 a = BASE_IN ? e2_base(gl_canon(p[0])) : e2_load(p, 0);
        const ext2 b = BASE_IN ? e2_base(gl_canon(p[1])) : e2_load(p, 1);
        v = e2_add(a, e2_mul(ca, e2_sub(b, a)));
    }
    e2_store(finals, i, v);
}

template <int NP, bool BASE_IN, int FOLDS, bool LOOK>
static int launch_pass_t(const u64 *in, size_t in_stride, u64 *out, size_t out_stride, size_t items, ext2 ca,
                         ext2 cb, const u64 *chp, const ScScratch &s, hipStream_t st, const MailJob *mj,
                         bool *posted) {
    unsigned g = grid_stride(items, MAM_BLOCK, MAM_RED_MAX_BLOCKS);
    ProfScope prof(PROF_SUMCHECK,
                   (double)items * NP * ((LOOK ? 4 : 2) * (FOLDS ? 4 : 1) * (BASE_IN ? 8.0 : 16.0) +
                                         (FOLDS ? (LOOK ? 4 : 2) * 16.0 : 0.0)),
                   st);
    const unsigned cap = pass_grid_cap(PassWords<NP, LOOK>::value);
    if (g > cap) g = cap;
    if (items <= lp_max_items()) {
        const size_t want = (items + LP_ITEMS - 1) / LP_ITEMS;
        const unsigned glp = (unsigned)(want < cap ? (want ? want : 1) : cap);
        g = glp;
        k_sc_pass_lp<NP, BASE_IN, FOLDS, LOOK><<<g, MAM_BLOCK, 0, st>>>(in, in_stride, out, out_stride, items, ca,
                                                                       cb, chp, s.partials, s.msg);
    } else {
        k_sc_pass<NP, BASE_IN, FOLDS, LOOK><<<g, MAM_BLOCK, 0, st>>>(in, in_stride, out, out_stride, items, ca, cb,
                                                                     chp, s.partials, s.msg);
    }
    const bool p = launch_sum_partials<PassWords<NP, LOOK>::value>(s.partials, g, s.msg, st, mj);
    if (posted) *posted = p;
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NP>
static int launch_pass_np(const u64 *in, size_t in_stride, bool base_in, int folds, bool look, u64 *out,
                          size_t out_stride, size_t items, ext2 ca, ext2 cb, const u64 *chp, const ScScratch &s,
                          hipStream_t st, const MailJob *mj, bool *posted) {
#define MAM_PASS(B, F, L)                                                                            \
    if (base_in == B && folds == F && look == L)                                                     \
        return launch_pass_t<NP, B, F, L>(in, in_stride, out, out_stride, items, ca, cb, chp, s, st, mj, posted);
    MAM_PASS(true, 0, true)
    MAM_PASS(true, 0, false)
    MAM_PASS(true, 2, true)
    MAM_PASS(true, 2, false)
    MAM_PASS(false, 0, true)
    MAM_PASS(false, 0, false)
    MAM_PASS(false, 2, true)
    MAM_PASS(false, 2, false)
#undef MAM_PASS
    return -2;
}

int mam_sc_launch_pass(int np, const u64 *in, size_t in_stride, bool base_in, int folds, bool look, u64 *out,
                       size_t out_stride, size_t items, ext2 ca, ext2 cb, const ScScratch &s, hipStream_t st,
                       const MailJob *mj, bool *posted, const u64 *d_ch) {
    if (posted) *posted = false;
    switch (np) {
        case 1:
            return launch_pass_np<1>(in, in_stride, base_in, folds, look, out, out_stride, items, ca, cb, d_ch, s, st,
                                     mj, posted);
        case 2:
            return launch_pass_np<2>(in, in_stride, base_in, folds, look, out, out_stride, items, ca, cb, d_ch, s, st,
                                     mj, posted);
        case 3:
            return launch_pass_np<3>(in, in_stride, base_in, folds, look, out, out_stride, items, ca, cb, d_ch, s, st,
                                     mj, posted);
        default: return -2;
    }
}

int mam_sc_launch_pass_final(int np, const u64 *in, size_t in_stride, bool base_in, int folds, ext2 ca, ext2 cb,
                             u64 *finals, hipStream_t st, const u64 *d_ch) {
    if (np < 1 || np > MAM_MAX_POLYS || (folds != 1 && folds != 2)) return -2;
    if (base_in && folds == 2) k_sc_pass_final<true, 2><<<1, 64, 0, st>>>(in, in_stride, np, ca, cb, d_ch, finals);
    else if (base_in) k_sc_pass_final<true, 1><<<1, 64, 0, st>>>(in, in_stride, np, ca, cb, d_ch, finals);
    else if (folds == 2) k_sc_pass_final<false, 2><<<1, 64, 0, st>>>(in, in_stride, np, ca, cb, d_ch, finals);
    else k_sc_pass_final<false, 1><<<1, 64, 0, st>>>(in, in_stride, np, ca, cb, d_ch, finals);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// Batched product sumcheck (two polynomials) with the transcript on the device.
//
// k_bsc_pass: workgroup b of the launch finds its job (instance) and runs the same pass body
// as k_sc_pass over the job's workgroups, with the pending challenges read from the
// instance's slot; it leaves its reduced words in partials[b]. k_bsc_round: one wavefront per
// job adds the job's partials, then plays the prover's transcript steps of the pass
// (gpu_prover.rs:21-89: each value of the round message appended under "round_message", the
// challenge drawn under "challenge"; for a two-round pass the second message is evaluated
// from the first challenge, as the whole-proof host prover does), stores the messages and
// challenges, refreshes the challenge slot and, after the instance's last pass, the final
// values. No host round trip anywhere, so the batch is a fixed launch sequence.
// ---------------------------------------------------------------------------
namespace {

__constant__ uint8_t kLRound[13] = {'r', 'o', 'u', 'n', 'd', '_', 'm', 'e', 's', 's', 'a', 'g', 'e'};
__constant__ uint8_t kLChallenge[9] = {'c', 'h', 'a', 'l', 'l', 'e', 'n', 'g', 'e'};

__device__ __forceinline__ int bsc_find(const BscJob *jobs, int njobs, unsigned b) {
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jobs[mid].blk0 <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// lds: PassLds<PassWords<2, true>::value>::value words (the largest of the pass shapes)
template <bool BASE_IN, int FOLDS, bool LOOK>
__device__ __forceinline__ void bsc_body(const BscJob &J, ext2 ca, ext2 cb, unsigned local, u64 *slot, u64 *lds) {
    constexpr int W = PassWords<2, LOOK>::value;
    u64 acc[W];
    sc_pass_body<2, BASE_IN, FOLDS, LOOK>(J.in, J.in_stride, J.out, J.out_stride, J.items, ca, cb, local, J.nblk, acc,
                                          lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < W; k++) slot[k] = acc[k];
    }
}

__global__ __launch_bounds__(MAM_BLOCK) void k_bsc_pass(const BscJob *__restrict__ jobs, int njobs,
                                                       const BscInst *__restrict__ insts, u64 *partials) {
    const int ji = bsc_find(jobs, njobs, blockIdx.x);
    const BscJob J = jobs[ji];
    const unsigned local = blockIdx.x - J.blk0;
    if (local >= J.nblk) return;  // cannot happen for a well-formed table; keeps writes in range
    const u64 *c = insts[J.inst].ch;
    const ext2 ca{gl_canon(c[0]), gl_canon(c[1])}, cb{gl_canon(c[2]), gl_canon(c[3])};
    u64 *slot = partials + (size_t)blockIdx.x * MAM_BSC_WORDS;
    __shared__ u64 lds[PassLds<PassWords<2, true>::value>::value];
    // the pass shapes of a two-rounds-per-pass schedule: the first pass reads the base inputs
    // as they are, the second reads them with two folds applied, later passes read Ext2 arrays
    if (J.base_in && J.folds == 0) {
        if (J.look) bsc_body<true, 0, true>(J, ca, cb, local, slot, lds);
        else bsc_body<true, 0, false>(J, ca, cb, local, slot, lds);
    } else if (J.base_in) {
        if (J.look) bsc_body<true, 2, true>(J, ca, cb, local, slot, lds);
        else bsc_body<true, 2, false>(J, ca, cb, local, slot, lds);
    } else {
        if (J.look) bsc_body<false, 2, true>(J, ca, cb, local, slot, lds);
        else bsc_body<false, 2, false>(J, ca, cb, local, slot, lds);
    }
}

__global__ __launch_bounds__(64) void k_bsc_round(const BscJob *__restrict__ jobs, const BscInst *__restrict__ insts,
                                                  const u64 *__restrict__ partials) {
    constexpr int NP = 2, DP1 = NP + 1;
    const BscJob J = jobs[blockIdx.x];
    const BscInst I = insts[J.inst];
    const int W = J.look ? PassWords<NP, true>::value : PassWords<NP, false>::value;
    // the job's reduced words: word k summed over its workgroups (lanes stride the workgroups)
    u64 words[MAM_BSC_WORDS];
    for (int k = 0; k < MAM_BSC_WORDS; k++) {
        u64 v = 0;
        if (k < W)
            for (unsigned b = threadIdx.x; b < J.nblk; b += 64) v = gl_add(v, partials[(size_t)(J.blk0 + b) * MAM_BSC_WORDS + k]);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v = gl_add(v, shfl_xor_u64(v, off));
        words[k] = v;
    }
    dtr::Wave w;
    w.load(I.tr);
    auto draw = [&](const u64 *m, int r) {
        for (int c = 0; c < DP1; c++) {
            w.label(kLRound, 13);
            w.absorb(m[2 * c]);
            w.absorb(m[2 * c + 1]);
        }
        w.label(kLChallenge, 9);
        const u64 c0 = w.squeeze(), c1 = w.squeeze();
        if (threadIdx.x < 2 * DP1) I.msgs[(size_t)r * 2 * DP1 + threadIdx.x] = m[threadIdx.x];
        if (threadIdx.x == 0) {
            I.chs[2 * r] = c0;
            I.chs[2 * r + 1] = c1;
        }
        return ext2{c0, c1};
    };
    u64 m0[2 * DP1];
    for (int k = 0; k < 2 * DP1; k++) m0[k] = words[k];
    const ext2 ca = draw(m0, J.round);
    ext2 cb = e2_zero();
    if (J.look) {
        // round r + 1: msg(c) = sum_j ca^j sym[c][j] (Horner)
        u64 m1[2 * DP1];
        for (int c = 0; c < DP1; c++) {
            const int o = 2 * DP1 + 2 * (c * DP1);
            ext2 v{words[o + 2 * NP], words[o + 2 * NP + 1]};
            for (int j = NP - 1; j >= 0; j--) v = e2_add(e2_mul(v, ca), ext2{words[o + 2 * j], words[o + 2 * j + 1]});
            m1[2 * c] = v.c0;
            m1[2 * c + 1] = v.c1;
        }
        cb = draw(m1, J.round + 1);
    }
    w.store(I.tr);
    if (threadIdx.x == 0) {
        I.ch[0] = ca.c0;
        I.ch[1] = ca.c1;
        I.ch[2] = cb.c0;
        I.ch[3] = cb.c1;
    }
    if (J.last && threadIdx.x < NP) {
        const u64 *p = J.fin_src + threadIdx.x * J.fin_stride;
        ext2 v;
        if (J.fin_folds == 2) {
            v = J.fin_base ? pass_load<true, 2>(p, 0, ca, gl_mul7(ca.c1), cb, gl_mul7(cb.c1))
                           : pass_load<false, 2>(p, 0, ca, gl_mul7(ca.c1), cb, gl_mul7(cb.c1));
        } else {
            const ext2 a = J.fin_base ? e2_base(gl_canon(p[0])) : e2_load(p, 0);
            const ext2 b = J.fin_base ? e2_base(gl_canon(p[1])) : e2_load(p, 1);
            v = e2_add(a, e2_mul(ca, e2_sub(b, a)));
        }
        e2_store(I.finals, threadIdx.x, v);
    }
}

}  // namespace

int mam_bsc_launch_pass(const BscJob *d_jobs, int njobs, unsigned nblk_total, const BscInst *d_insts, u64 *partials,
                        hipStream_t st) {
    if (njobs <= 0 || nblk_total == 0) return -2;
    k_bsc_pass<<<nblk_total, MAM_BLOCK, 0, st>>>(d_jobs, njobs, d_insts, partials);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mam_bsc_launch_round(const BscJob *d_jobs, int njobs, const BscInst *d_insts, const u64 *partials,
                         hipStream_t st) {
    if (njobs <= 0) return -2;
    k_bsc_round<<<njobs, 64, 0, st>>>(d_jobs, d_insts, partials);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

