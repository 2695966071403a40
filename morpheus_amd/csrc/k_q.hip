// almost-Goldilocks (q = 2^64 - 2^32 - 31) twins of the field, eq, partial-evaluation and
// sumcheck entry points, under the symbol names of almost-goldilocks-cuda-rs/src/ffi.rs.
//
// The kernels are templates over a field policy (fpolicy.h), so the q and Goldilocks
// versions share one body; the q instantiations are exported here. Semantics follow the
// reference's CUDA kernels (cuda_almost_goldilocks/almost_goldilocks_kernels.cu,
// almost_extension_kernels.cu, almost_eq_lagrange.cuh:25-222, almost_partial_eval.cuh,
// almost_sumcheck_prover.cuh:19-368); outputs are canonical residues.
#include "common.h"
#include "eq_table.h"
#include "fpolicy.h"
#include "sumcheck.h"
#include "sync.h"

// almost-goldilocks-cuda-rs/src/ffi.rs:13 (the q crate's init; same device as the p crate)
MAM_API int almost_goldilocks_cuda_init() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return -1;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return -1;
    return hipSetDevice(cur) == hipSuccess ? 0 : -1;
}

namespace {

enum QBinOp { QB_ADD, QB_SUB, QB_MUL, QB_DIV };
enum QUnOp { QU_NEG, QU_SQR, QU_DBL, QU_INV, QU_EXP, QU_SCALE, QU_FROB };

template <class F, int OP>
__global__ void k_base_bin(const u64 *__restrict__ a, const u64 *__restrict__ b, u64 *__restrict__ r, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 x = F::canon(a[i]), y = F::canon(b[i]);
    u64 z;
    if (OP == QB_ADD) z = F::add(x, y);
    else if (OP == QB_SUB) z = F::sub(x, y);
    else if (OP == QB_MUL) z = F::mul(x, y);
    else z = F::mul(x, F::inv(y));
    r[i] = z;
}

template <class F, int OP>
__global__ void k_base_un(const u64 *__restrict__ a, u64 *__restrict__ r, u64 arg, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 x = F::canon(a[i]);
    u64 z;
    if (OP == QU_NEG) z = F::neg(x);
    else if (OP == QU_SQR) z = F::sqr(x);
    else if (OP == QU_DBL) z = F::dbl(x);
    else if (OP == QU_INV) z = F::inv(x);
    else if (OP == QU_EXP) z = F::pow(x, arg);
    else z = F::mul(x, F::canon(arg));
    r[i] = z;
}

template <class F, int OP>
__global__ void k_ext_bin(const u64 *__restrict__ a, const u64 *__restrict__ b, u64 *__restrict__ r, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const typename F::E x = F::e_load(a, i), y = F::e_load(b, i);
    typename F::E z;
    if (OP == QB_ADD) z = F::e_add(x, y);
    else if (OP == QB_SUB) z = F::e_sub(x, y);
    else z = F::e_mul(x, y);
    F::e_store(r, i, z);
}

template <class F, int OP>
__global__ void k_ext_un(const u64 *__restrict__ a, u64 *__restrict__ r, u64 arg, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const typename F::E x = F::e_load(a, i);
    typename F::E z;
    if (OP == QU_NEG) z = F::e_neg(x);
    else if (OP == QU_SQR) z = F::e_sqr(x);
    else if (OP == QU_INV) z = F::e_inv(x);
    else if (OP == QU_EXP) z = F::e_pow(x, arg);
    else if (OP == QU_SCALE) z = F::e_scale(F::canon(arg), x);
    else z = F::e_frob(x);
    F::e_store(r, i, z);
}

// acc[i] += s * src[i] (Ext2)
template <class F>
__global__ void k_ext_scale_acc(typename F::E s, const u64 *__restrict__ src, u64 *__restrict__ acc, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F::e_store(acc, i, F::e_add(F::e_load(acc, i), F::e_mul(s, F::e_load(src, i))));
}

template <class F>
__global__ void k_base_to_ext(const u64 *__restrict__ a, u64 *__restrict__ r, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    F::e_store(r, i, F::e_base(F::canon(a[i])));
}

template <class F>
__global__ void k_ext_to_base(const u64 *__restrict__ a, u64 *__restrict__ r, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    r[i] = F::canon(a[2 * (size_t)i]);
}

template <int OP>
int q_bin(const u64 *a, const u64 *b, u64 *r, int n) {
    if (n <= 0) return 0;
    k_base_bin<FieldQ, OP><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(a, b, r, n);
    return mam_sync_status();
}
template <int OP>
int q_un(const u64 *a, u64 *r, u64 arg, int n) {
    if (n <= 0) return 0;
    k_base_un<FieldQ, OP><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(a, r, arg, n);
    return mam_sync_status();
}
template <int OP>
int q2_bin(const u64 *a, const u64 *b, u64 *r, int n) {
    if (n <= 0) return 0;
    k_ext_bin<FieldQ, OP><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(a, b, r, n);
    return mam_sync_status();
}
template <int OP>
int q2_un(const u64 *a, u64 *r, u64 arg, int n) {
    if (n <= 0) return 0;
    k_ext_un<FieldQ, OP><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(a, r, arg, n);
    return mam_sync_status();
}

}  // namespace

// ---- base field batch ops (almost-goldilocks-cuda-rs/src/ffi.rs:103-112) ----
MAM_API int agl_batch_add_ffi(const u64 *a, const u64 *b, u64 *r, int n) { return q_bin<QB_ADD>(a, b, r, n); }
MAM_API int agl_batch_sub_ffi(const u64 *a, const u64 *b, u64 *r, int n) { return q_bin<QB_SUB>(a, b, r, n); }
MAM_API int agl_batch_mul_ffi(const u64 *a, const u64 *b, u64 *r, int n) { return q_bin<QB_MUL>(a, b, r, n); }
MAM_API int agl_batch_div_ffi(const u64 *a, const u64 *b, u64 *r, int n) { return q_bin<QB_DIV>(a, b, r, n); }
MAM_API int agl_batch_neg_ffi(const u64 *a, u64 *r, int n) { return q_un<QU_NEG>(a, r, 0, n); }
MAM_API int agl_batch_square_ffi(const u64 *a, u64 *r, int n) { return q_un<QU_SQR>(a, r, 0, n); }
MAM_API int agl_batch_double_ffi(const u64 *a, u64 *r, int n) { return q_un<QU_DBL>(a, r, 0, n); }
MAM_API int agl_batch_inverse_ffi(const u64 *a, u64 *r, int n) { return q_un<QU_INV>(a, r, 0, n); }
MAM_API int agl_batch_exp_ffi(const u64 *a, u64 e, u64 *r, int n) { return q_un<QU_EXP>(a, r, e, n); }
MAM_API int agl_batch_mul_scalar_ffi(u64 s, const u64 *a, u64 *r, int n) { return q_un<QU_SCALE>(a, r, s, n); }

// ---- Ext2 batch ops (ffi.rs:122-138) ----
MAM_API int aext2_batch_add_ffi(const u64 *a, const u64 *b, u64 *r, int n) { return q2_bin<QB_ADD>(a, b, r, n); }
MAM_API int aext2_batch_sub_ffi(const u64 *a, const u64 *b, u64 *r, int n) { return q2_bin<QB_SUB>(a, b, r, n); }
MAM_API int aext2_batch_mul_ffi(const u64 *a, const u64 *b, u64 *r, int n) { return q2_bin<QB_MUL>(a, b, r, n); }
MAM_API int aext2_batch_inverse_ffi(const u64 *a, u64 *r, int n) { return q2_un<QU_INV>(a, r, 0, n); }
MAM_API int aext2_batch_neg_ffi(const u64 *a, u64 *r, int n) { return q2_un<QU_NEG>(a, r, 0, n); }
MAM_API int aext2_batch_square_ffi(const u64 *a, u64 *r, int n) { return q2_un<QU_SQR>(a, r, 0, n); }
MAM_API int aext2_batch_frobenius_ffi(const u64 *a, u64 *r, int n) { return q2_un<QU_FROB>(a, r, 0, n); }
MAM_API int aext2_batch_conjugate_ffi(const u64 *a, u64 *r, int n) { return q2_un<QU_FROB>(a, r, 0, n); }
MAM_API int aext2_batch_exp_ffi(const u64 *a, u64 e, u64 *r, int n) { return q2_un<QU_EXP>(a, r, e, n); }
MAM_API int aext2_batch_mul_scalar_ffi(u64 s, const u64 *a, u64 *r, int n) { return q2_un<QU_SCALE>(a, r, s, n); }
MAM_API int aext2_scale_accumulate_ffi(u64 s0, u64 s1, const u64 *src, u64 *acc, int n) {
    if (n <= 0) return 0;
    k_ext_scale_acc<FieldQ><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(qe2{fq::canon(s0), fq::canon(s1)}, src, acc, n);
    return mam_sync_status();
}
// Goldilocks twin (goldilocks-cuda-rs/src/ffi.rs:633-639): acc += s * src over Ext2 X^2 - 7,
// asynchronous like the reference's (wrapper.cu:1807-1815).
MAM_API int ext2_scale_accumulate_ffi(u64 s0, u64 s1, const u64 *src, u64 *acc, int n) {
    if (n <= 0) return 0;
    k_ext_scale_acc<FieldGL><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(ext2{gl_canon(s0), gl_canon(s1)}, src, acc, n);
    return mam_launch_status();
}
MAM_API int agl_to_aext2_batch_ffi(const u64 *in, u64 *out, int n) {
    if (n <= 0) return 0;
    k_base_to_ext<FieldQ><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(in, out, n);
    return mam_sync_status();
}
MAM_API int aext2_to_agl_batch_ffi(const u64 *in, u64 *out, int n) {
    if (n <= 0) return 0;
    k_ext_to_base<FieldQ><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(in, out, n);
    return mam_sync_status();
}

// ---------------------------------------------------------------------------
// eq tables (eq_table.h), one pass; grid.y = table in the batched form, where r and out
// advance by r_stride / out_stride words per table.
// ---------------------------------------------------------------------------
template <class F, bool EXT>
static int eq_launch(const u64 *d_r, size_t r_stride, int nv, u64 *d_out, size_t out_stride, int tables,
                     hipStream_t s) {
    return eqt::launch<F, EXT>(d_r, r_stride, nv, d_out, out_stride, tables, s);
}

// The reference ping-pongs log_n layers between the two buffers, so its result lands in
// buf_a for even log_n and in buf_b for odd log_n; the table is written to that buffer.
static u64 *eq_target(u64 *a, u64 *b, int log_n) { return (log_n & 1) ? b : a; }

// ffi.rs:161-196
MAM_API int agl_eq_dp_all_ffi(const u64 *d_r, u64 *d_buf_a, u64 *d_buf_b, int log_n, u64 **d_result) {
    u64 *dst = eq_target(d_buf_a, d_buf_b, log_n);
    if (eq_launch<FieldQ, false>(d_r, 0, log_n, dst, 0, 1, mam_stream())) return -1;
    if (d_result) *d_result = dst;
    return mam_sync_status();
}
MAM_API int aext2_eq_dp_all_ffi(const u64 *d_r, u64 *d_buf_a, u64 *d_buf_b, int log_n, u64 **d_result) {
    u64 *dst = eq_target(d_buf_a, d_buf_b, log_n);
    if (eq_launch<FieldQ, true>(d_r, 0, log_n, dst, 0, 1, mam_stream())) return -1;
    if (d_result) *d_result = dst;
    return mam_sync_status();
}
// Asynchronous on the caller's stream (null: the calling thread's default stream).
MAM_API int aext2_eq_dp_all_stream_ffi(const u64 *d_r, u64 *d_buf_a, u64 *d_buf_b, int log_n, u64 **d_result,
                                       void *stream) {
    u64 *dst = eq_target(d_buf_a, d_buf_b, log_n);
    if (eq_launch<FieldQ, true>(d_r, 0, log_n, dst, 0, 1, stream ? (hipStream_t)stream : mam_stream())) return -1;
    if (d_result) *d_result = dst;
    return 0;
}
// num_leaves tables: point l at d_r_all + l * log_n Ext2, table l at buffer + l * leaf_stride
// Ext2 elements.
MAM_API int aext2_eq_dp_all_batched_ffi(const u64 *d_r_all, u64 *d_buf_a_all, u64 *d_buf_b_all, int log_n,
                                        int num_leaves, size_t leaf_stride, u64 **d_result, void *stream) {
    u64 *dst = eq_target(d_buf_a_all, d_buf_b_all, log_n);
    if (d_result) *d_result = dst;
    if (num_leaves <= 0) return 0;
    if (leaf_stride < ((size_t)1 << log_n)) return -1;
    return eq_launch<FieldQ, true>(d_r_all, 2 * (size_t)log_n, log_n, dst, 2 * leaf_stride, num_leaves,
                                   stream ? (hipStream_t)stream : mam_stream());
}

// ---------------------------------------------------------------------------
// Partial evaluation (fold the first m variables, LSB first)
// KIND 0: base -> base (base r); 1: base -> Ext2 (Ext2 r); 2: Ext2 -> Ext2
// ---------------------------------------------------------------------------
template <class F, int KIND>
__global__ void k_peval_f(const u64 *__restrict__ in, u64 *__restrict__ out, const u64 *__restrict__ r, int li,
                          size_t pairs) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= pairs) return;
    if (KIND == 0) {
        const u64 ri = F::canon(r[li]);
        const u64 a = F::canon(in[2 * i]), b = F::canon(in[2 * i + 1]);
        out[i] = F::add(a, F::mul(ri, F::sub(b, a)));
    } else if (KIND == 1) {
        const typename F::E ri = F::e_load(r, li);
        const u64 a = F::canon(in[2 * i]), b = F::canon(in[2 * i + 1]);
        F::e_store(out, i, F::e_add(F::e_base(a), F::e_scale(F::sub(b, a), ri)));
    } else {
        const typename F::E ri = F::e_load(r, li);
        const typename F::E a = F::e_load(in, 2 * i), b = F::e_load(in, 2 * i + 1);
        F::e_store(out, i, F::e_add(a, F::e_mul(ri, F::e_sub(b, a))));
    }
}

// ffi.rs:198-212
MAM_API int agl_partial_eval_ffi(u64 *d_data, const u64 *d_r, int log_n, int m) {
    if (m <= 0) return 0;
    if (m > log_n) return -1;
    hipStream_t s = mam_stream();
    size_t cur = (size_t)1 << log_n;
    u64 *scratch = nullptr;
    if (hipMallocAsync((void **)&scratch, (cur / 2) * sizeof(u64), s) != hipSuccess) return -1;
    u64 *src = d_data, *dst = scratch;
    int rc = 0;
    for (int i = 0; i < m && !rc; i++) {
        const size_t pairs = cur / 2;
        k_peval_f<FieldQ, 0><<<grid_1d(pairs), MAM_BLOCK, 0, s>>>(src, dst, d_r, i, pairs);
        if (hipGetLastError() != hipSuccess) rc = -1;
        u64 *t = src;
        src = dst;
        dst = t;
        cur = pairs;
    }
    if (!rc && src != d_data &&
        hipMemcpyAsync(d_data, src, cur * sizeof(u64), hipMemcpyDeviceToDevice, s) != hipSuccess)
        rc = -1;
    (void)hipFreeAsync(scratch, s);
    return rc ? rc : mam_sync_status();
}

MAM_API int agl_partial_eval_ext2_from_base_ffi(const u64 *d_input, u64 *d_output, const u64 *d_r, int log_n, int m) {
    if (m <= 0) return 0;
    if (m > log_n) return -1;
    hipStream_t s = mam_stream();
    size_t cur = (size_t)1 << log_n;
    size_t pairs = cur / 2;
    k_peval_f<FieldQ, 1><<<grid_1d(pairs), MAM_BLOCK, 0, s>>>(d_input, d_output, d_r, 0, pairs);
    if (hipGetLastError() != hipSuccess) return -1;
    cur = pairs;
    if (m == 1) return mam_sync_status();
    u64 *scratch = nullptr;
    if (hipMallocAsync((void **)&scratch, (cur / 2 > 0 ? cur / 2 : 1) * 2 * sizeof(u64), s) != hipSuccess) return -1;
    u64 *src = d_output, *dst = scratch;
    int rc = 0;
    for (int i = 1; i < m && !rc; i++) {
        pairs = cur / 2;
        k_peval_f<FieldQ, 2><<<grid_1d(pairs), MAM_BLOCK, 0, s>>>(src, dst, d_r, i, pairs);
        if (hipGetLastError() != hipSuccess) rc = -1;
        u64 *t = src;
        src = dst;
        dst = t;
        cur = pairs;
    }
    if (!rc && src != d_output &&
        hipMemcpyAsync(d_output, src, cur * 2 * sizeof(u64), hipMemcpyDeviceToDevice, s) != hipSuccess)
        rc = -1;
    (void)hipFreeAsync(scratch, s);
    return rc ? rc : mam_sync_status();
}

// ---------------------------------------------------------------------------
// Sumcheck round message / fold with the reference's per-workgroup partials
// (caller reduces): d polynomials of `original_size` elements back to back; message
// value c = 0..d of sum_y prod_i (a_i + c (b_i - a_i)).
// ---------------------------------------------------------------------------
template <class F, bool EXT>
__global__ __launch_bounds__(MAM_BLOCK) void k_round_msg_f(const u64 *__restrict__ polys, u64 *__restrict__ partial,
                                                           int d, size_t original_size, size_t half) {
    using E = typename F::E;
    __shared__ u64 lds[(MAM_MAX_POLYS + 1) * 2 * MAM_BLOCK / 64];
    E acc[MAM_MAX_POLYS + 1];
    for (int c = 0; c <= d; c++) acc[c] = F::e_zero();
    const size_t stride = EXT ? 2 * original_size : original_size;
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half; y += (size_t)gridDim.x * blockDim.x) {
        E e[MAM_MAX_POLYS], df[MAM_MAX_POLYS];
        for (int i = 0; i < d; i++) {
            const u64 *p = polys + i * stride;
            const E a = EXT ? F::e_load(p, 2 * y) : F::e_base(F::canon(p[2 * y]));
            const E b = EXT ? F::e_load(p, 2 * y + 1) : F::e_base(F::canon(p[2 * y + 1]));
            e[i] = a;
            df[i] = F::e_sub(b, a);
        }
        for (int c = 0; c <= d; c++) {
            E prod = e[0];
            for (int i = 1; i < d; i++) prod = F::e_mul(prod, e[i]);
            acc[c] = F::e_add(acc[c], prod);
            for (int i = 0; i < d; i++) e[i] = F::e_add(e[i], df[i]);
        }
    }
    // wave butterflies, then one LDS slot per wave
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c <= d; c++) {
        for (int off = 32; off > 0; off >>= 1) {
            const u64 o0 = ((u64)__shfl_xor((u32)(acc[c].c0 >> 32), off) << 32) | (u32)__shfl_xor((u32)acc[c].c0, off);
            const u64 o1 = ((u64)__shfl_xor((u32)(acc[c].c1 >> 32), off) << 32) | (u32)__shfl_xor((u32)acc[c].c1, off);
            acc[c] = F::e_add(acc[c], F::e(o0, o1));
        }
        if (lane == 0) {
            lds[(c * (MAM_BLOCK / 64) + wave) * 2] = acc[c].c0;
            lds[(c * (MAM_BLOCK / 64) + wave) * 2 + 1] = acc[c].c1;
        }
    }
    mam_barrier();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x / 64;
        for (int c = 0; c <= d; c++) {
            E s = F::e_zero();
            for (int w = 0; w < nw; w++) s = F::e_add(s, F::e(lds[(c * (MAM_BLOCK / 64) + w) * 2],
                                                                lds[(c * (MAM_BLOCK / 64) + w) * 2 + 1]));
            if (EXT) F::e_store(partial, (size_t)blockIdx.x * (d + 1) + c, s);
            else partial[(size_t)blockIdx.x * (d + 1) + c] = s.c0;
        }
    }
}

template <class F, bool EXT>
__global__ void k_fold_f(const u64 *__restrict__ in, u64 *__restrict__ out, typename F::E ch, int d,
                         size_t original_size, size_t half) {
    const size_t stride = EXT ? 2 * original_size : original_size;
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half; y += (size_t)gridDim.x * blockDim.x) {
        for (int i = 0; i < d; i++) {
            const u64 *p = in + i * stride;
            u64 *q = out + i * stride;
            if (EXT) {
                const typename F::E a = F::e_load(p, 2 * y), b = F::e_load(p, 2 * y + 1);
                F::e_store(q, y, F::e_add(a, F::e_mul(ch, F::e_sub(b, a))));
            } else {
                const u64 a = F::canon(p[2 * y]), b = F::canon(p[2 * y + 1]);
                q[y] = F::add(a, F::mul(ch.c0, F::sub(b, a)));
            }
        }
    }
}

// ffi.rs:228-263
MAM_API int agl_sumcheck_round_message_ffi(const u64 *d_polys, u64 *d_partial, int d, size_t original_size,
                                           size_t half, int num_blocks) {
    if (d < 1 || d > MAM_MAX_POLYS || num_blocks <= 0) return -1;
    k_round_msg_f<FieldQ, false><<<num_blocks, MAM_BLOCK, 0, mam_stream()>>>(d_polys, d_partial, d, original_size,
                                                                           half);
    return mam_sync_status();
}
MAM_API int aext2_sumcheck_round_message_ffi(const u64 *d_polys, u64 *d_partial, int d, size_t original_size,
                                             size_t half, int num_blocks) {
    if (d < 1 || d > MAM_MAX_POLYS || num_blocks <= 0) return -1;
    k_round_msg_f<FieldQ, true><<<num_blocks, MAM_BLOCK, 0, mam_stream()>>>(d_polys, d_partial, d, original_size,
                                                                          half);
    return mam_sync_status();
}
MAM_API int agl_sumcheck_fold_ffi(const u64 *d_input, u64 *d_output, u64 challenge, int d, size_t original_size,
                                  size_t half) {
    if (d < 1 || half == 0) return half == 0 ? 0 : -1;
    k_fold_f<FieldQ, false><<<grid_stride(half, MAM_BLOCK, 8192), MAM_BLOCK, 0, mam_stream()>>>(
        d_input, d_output, q2_base(fq::canon(challenge)), d, original_size, half);
    return mam_sync_status();
}
MAM_API int aext2_sumcheck_fold_ffi(const u64 *d_input, u64 *d_output, u64 c0, u64 c1, int d, size_t original_size,
                                    size_t half) {
    if (d < 1 || half == 0) return half == 0 ? 0 : -1;
    k_fold_f<FieldQ, true><<<grid_stride(half, MAM_BLOCK, 8192), MAM_BLOCK, 0, mam_stream()>>>(
        d_input, d_output, qe2{fq::canon(c0), fq::canon(c1)}, d, original_size, half);
    return mam_sync_status();
}

// ---------------------------------------------------------------------------
// Batched same-size degree-2 sumcheck (GpuBatchedSamePointState, almost-goldilocks-cuda-rs/
// src/sumcheck_prover.rs:468-1027; almost_sumcheck_prover.cuh:219-368): num_leaves
// instances, each two Ext2 polynomials (e, f) of original_size elements back to back;
// grid (num_blocks_x, num_leaves); partial[(block_x * num_leaves + leaf) * 3 + c] for the
// evaluations at c = 0, 1, 2.
// ---------------------------------------------------------------------------
template <class F>
__global__ __launch_bounds__(MAM_BLOCK) void k_batched_msg_f(const u64 *__restrict__ polys, u64 *__restrict__ partial,
                                                             size_t original_size, size_t half, int num_leaves) {
    using E = typename F::E;
    __shared__ u64 lds[3 * 2 * MAM_BLOCK / 64];
    const int leaf = blockIdx.y;
    const size_t ps = 2 * original_size;
    const u64 *e_p = polys + (size_t)leaf * 2 * ps;
    const u64 *f_p = e_p + ps;
    E acc[3] = {F::e_zero(), F::e_zero(), F::e_zero()};
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half; y += (size_t)gridDim.x * blockDim.x) {
        const E e0 = F::e_load(e_p, 2 * y), e1 = F::e_load(e_p, 2 * y + 1);
        const E f0 = F::e_load(f_p, 2 * y), f1 = F::e_load(f_p, 2 * y + 1);
        acc[0] = F::e_add(acc[0], F::e_mul(e0, f0));
        acc[1] = F::e_add(acc[1], F::e_mul(e1, f1));
        const E e2v = F::e_sub(F::e_add(e1, e1), e0), f2v = F::e_sub(F::e_add(f1, f1), f0);
        acc[2] = F::e_add(acc[2], F::e_mul(e2v, f2v));
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < 3; c++) {
        for (int off = 32; off > 0; off >>= 1) {
            const u64 o0 = ((u64)__shfl_xor((u32)(acc[c].c0 >> 32), off) << 32) | (u32)__shfl_xor((u32)acc[c].c0, off);
            const u64 o1 = ((u64)__shfl_xor((u32)(acc[c].c1 >> 32), off) << 32) | (u32)__shfl_xor((u32)acc[c].c1, off);
            acc[c] = F::e_add(acc[c], F::e(o0, o1));
        }
        if (lane == 0) {
            lds[(c * (MAM_BLOCK / 64) + wave) * 2] = acc[c].c0;
            lds[(c * (MAM_BLOCK / 64) + wave) * 2 + 1] = acc[c].c1;
        }
    }
    mam_barrier();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x / 64;
        for (int c = 0; c < 3; c++) {
            E s = F::e_zero();
            for (int w = 0; w < nw; w++)
                s = F::e_add(s, F::e(lds[(c * (MAM_BLOCK / 64) + w) * 2], lds[(c * (MAM_BLOCK / 64) + w) * 2 + 1]));
            F::e_store(partial, ((size_t)blockIdx.x * num_leaves + leaf) * 3 + c, s);
        }
    }
}

template <class F>
__global__ void k_batched_fold_f(const u64 *__restrict__ in, u64 *__restrict__ out, typename F::E ch,
                                 size_t original_size, size_t half) {
    const int leaf = blockIdx.y;
    const size_t ps = 2 * original_size;
    const u64 *src = in + (size_t)leaf * 2 * ps;
    u64 *dst = out + (size_t)leaf * 2 * ps;
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half; y += (size_t)gridDim.x * blockDim.x) {
        for (int i = 0; i < 2; i++) {
            const typename F::E a = F::e_load(src + i * ps, 2 * y), b = F::e_load(src + i * ps, 2 * y + 1);
            F::e_store(dst + i * ps, y, F::e_add(a, F::e_mul(ch, F::e_sub(b, a))));
        }
    }
}

// Round 0 with a binary f (GpuBatchedSamePointState, sumcheck_prover.rs:468-1027): f is still
// a bit vector (`packed_size_u64` words per leaf, bit i of word i/64 = f(i)), so the message
// and the fold read the bits instead of a lifted Ext2 f-slot. The message uses the same
// grid-stride split as k_batched_msg_f (identical per-block partials); the fold writes both
// slots at half size: eq' = e0 + ch (e1 - e0) and f' = f0 + ch (f1 - f0) in {0, ch, 1 - ch, 1}.
// Pairs (2y, 2y+1) never straddle a packed word.
template <class F>
__global__ __launch_bounds__(MAM_BLOCK) void k_batched_r0bin_msg_f(const u64 *__restrict__ polys,
                                                                   const u64 *__restrict__ packed,
                                                                   u64 *__restrict__ partial, size_t original_size,
                                                                   size_t half, int num_leaves, size_t packed_words) {
    using E = typename F::E;
    __shared__ u64 lds[3 * 2 * MAM_BLOCK / 64];
    const int leaf = blockIdx.y;
    const u64 *e_p = polys + (size_t)leaf * 4 * original_size;
    const u64 *bits = packed + (size_t)leaf * packed_words;
    E acc[3] = {F::e_zero(), F::e_zero(), F::e_zero()};
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half; y += (size_t)gridDim.x * blockDim.x) {
        const E e0 = F::e_load(e_p, 2 * y), e1 = F::e_load(e_p, 2 * y + 1);
        const u64 w = bits[y >> 5] >> ((2 * y) & 63);
        const bool f0 = w & 1, f1 = (w >> 1) & 1;
        if (f0) acc[0] = F::e_add(acc[0], e0);
        if (f1) acc[1] = F::e_add(acc[1], e1);
        // (2 e1 - e0) (2 f1 - f0), 2 f1 - f0 in {-1, 0, 1, 2}
        if (f0 != f1 || f1) {
            const E e2v = F::e_sub(F::e_add(e1, e1), e0);
            if (f1 && !f0)
                acc[2] = F::e_add(acc[2], F::e_add(e2v, e2v));
            else if (f1)
                acc[2] = F::e_add(acc[2], e2v);
            else
                acc[2] = F::e_sub(acc[2], e2v);
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < 3; c++) {
        for (int off = 32; off > 0; off >>= 1) {
            const u64 o0 = ((u64)__shfl_xor((u32)(acc[c].c0 >> 32), off) << 32) | (u32)__shfl_xor((u32)acc[c].c0, off);
            const u64 o1 = ((u64)__shfl_xor((u32)(acc[c].c1 >> 32), off) << 32) | (u32)__shfl_xor((u32)acc[c].c1, off);
            acc[c] = F::e_add(acc[c], F::e(o0, o1));
        }
        if (lane == 0) {
            lds[(c * (MAM_BLOCK / 64) + wave) * 2] = acc[c].c0;
            lds[(c * (MAM_BLOCK / 64) + wave) * 2 + 1] = acc[c].c1;
        }
    }
    mam_barrier();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x / 64;
        for (int c = 0; c < 3; c++) {
            E s = F::e_zero();
            for (int w = 0; w < nw; w++)
                s = F::e_add(s, F::e(lds[(c * (MAM_BLOCK / 64) + w) * 2], lds[(c * (MAM_BLOCK / 64) + w) * 2 + 1]));
            F::e_store(partial, ((size_t)blockIdx.x * num_leaves + leaf) * 3 + c, s);
        }
    }
}

template <class F>
__global__ void k_batched_r0bin_fold_f(const u64 *__restrict__ polys, const u64 *__restrict__ packed,
                                       u64 *__restrict__ out, typename F::E ch, size_t original_size, size_t half,
                                       size_t packed_words) {
    using E = typename F::E;
    const int leaf = blockIdx.y;
    const size_t ps = 2 * original_size;
    const u64 *e_p = polys + (size_t)leaf * 2 * ps;
    const u64 *bits = packed + (size_t)leaf * packed_words;
    u64 *dst = out + (size_t)leaf * 2 * ps;
    const E one = F::e_one();
    for (size_t y = blockIdx.x * (size_t)blockDim.x + threadIdx.x; y < half; y += (size_t)gridDim.x * blockDim.x) {
        const E a = F::e_load(e_p, 2 * y), b = F::e_load(e_p, 2 * y + 1);
        F::e_store(dst, y, F::e_add(a, F::e_mul(ch, F::e_sub(b, a))));
        const u64 w = bits[y >> 5] >> ((2 * y) & 63);
        const bool f0 = w & 1, f1 = (w >> 1) & 1;
        const E f = f0 ? (f1 ? one : F::e_sub(one, ch)) : (f1 ? ch : F::e_zero());
        F::e_store(dst + ps, y, f);
    }
}

// Lifts of a packed witness into the f-slot of the batched layout (the eq slot is left
// alone): binary f(i) = bit i; ternary f(i) = pos_i - neg_i in {0, 1, -1} (both set gives 0).
template <class F, bool TERNARY>
__global__ void k_batched_lift_f(const u64 *__restrict__ pos, const u64 *__restrict__ neg, u64 *__restrict__ polys,
                                 size_t original_size, size_t packed_words) {
    const int leaf = blockIdx.y;
    const u64 *pb = pos + (size_t)leaf * packed_words;
    const u64 *nb = TERNARY ? neg + (size_t)leaf * packed_words : nullptr;
    u64 *f = polys + (size_t)leaf * 4 * original_size + 2 * original_size;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < original_size;
         i += (size_t)gridDim.x * blockDim.x) {
        const u64 p = (pb[i >> 6] >> (i & 63)) & 1;
        u64 v = p;
        if (TERNARY) v = F::sub(p, (nb[i >> 6] >> (i & 63)) & 1);
        F::e_store(f, i, F::e_base(v));
    }
}

// The witness bits a launch reads: a leaf's packed block must hold original_size bits.
static bool packed_ok(size_t original_size, size_t packed_words) {
    return packed_words >= (original_size + 63) / 64;
}

// ffi.rs:402-411
MAM_API int aext2_sumcheck_batched_round0_binary_msg_ffi(const u64 *d_polys, const u64 *d_packed, u64 *d_partial,
                                                         size_t original_size, size_t half, int num_leaves,
                                                         int num_blocks_x, size_t packed_size_u64) {
    if (num_leaves <= 0) return 0;
    if (num_blocks_x <= 0 || num_leaves > 65535 || 2 * half > original_size ||
        !packed_ok(original_size, packed_size_u64))
        return -1;
    k_batched_r0bin_msg_f<FieldQ><<<dim3(num_blocks_x, num_leaves), MAM_BLOCK, 0, mam_stream()>>>(
        d_polys, d_packed, d_partial, original_size, half, num_leaves, packed_size_u64);
    return mam_sync_status();
}
// ffi.rs:413-423
MAM_API int aext2_sumcheck_batched_round0_binary_fold_ffi(const u64 *d_polys, const u64 *d_packed, u64 *d_output,
                                                          u64 c0, u64 c1, size_t original_size, size_t half,
                                                          int num_leaves, size_t packed_size_u64) {
    if (num_leaves <= 0 || half == 0) return 0;
    if (num_leaves > 65535 || 2 * half > original_size || !packed_ok(original_size, packed_size_u64)) return -1;
    unsigned gx = grid_stride(half, MAM_BLOCK, 1024);
    k_batched_r0bin_fold_f<FieldQ><<<dim3(gx, num_leaves), MAM_BLOCK, 0, mam_stream()>>>(
        d_polys, d_packed, d_output, qe2{fq::canon(c0), fq::canon(c1)}, original_size, half, packed_size_u64);
    return mam_sync_status();
}
// ffi.rs:293-299
MAM_API int aext2_batched_lift_binary_ffi(const u64 *d_packed, u64 *d_polys, size_t original_size, int num_leaves,
                                          size_t packed_size_u64) {
    if (num_leaves <= 0 || original_size == 0) return 0;
    if (num_leaves > 65535 || !packed_ok(original_size, packed_size_u64)) return -1;
    unsigned gx = grid_stride(original_size, MAM_BLOCK, 1024);
    k_batched_lift_f<FieldQ, false><<<dim3(gx, num_leaves), MAM_BLOCK, 0, mam_stream()>>>(
        d_packed, nullptr, d_polys, original_size, packed_size_u64);
    return mam_sync_status();
}
// ffi.rs:274-281
MAM_API int aext2_batched_lift_ternary_single_ffi(const u64 *d_pos, const u64 *d_neg, u64 *d_polys,
                                                  size_t original_size, int num_leaves, size_t packed_size_u64) {
    if (num_leaves <= 0 || original_size == 0) return 0;
    if (num_leaves > 65535 || !packed_ok(original_size, packed_size_u64)) return -1;
    unsigned gx = grid_stride(original_size, MAM_BLOCK, 1024);
    k_batched_lift_f<FieldQ, true><<<dim3(gx, num_leaves), MAM_BLOCK, 0, mam_stream()>>>(
        d_pos, d_neg, d_polys, original_size, packed_size_u64);
    return mam_sync_status();
}

// Selective add over a shared eq table (almost_sumcheck_prover.cuh:924-966): for each packed
// binary plane p, partial[(bx * n_planes + p) * 2 ..] = sum of the Ext2 eq[i] over the set bits i
// that workgroup bx visits in the grid-stride order of the reference (so every per-block partial,
// which callers add up themselves, holds the same sum). grid (num_blocks_x, n_planes).
__global__ __launch_bounds__(MAM_BLOCK) void k_selective_add_planes(const u64 *__restrict__ eq,
                                                                    const u64 *__restrict__ planes,
                                                                    u64 *__restrict__ partial, size_t total,
                                                                    int n_planes, size_t packed) {
    using F = FieldQ;
    __shared__ u64 lds[2 * MAM_BLOCK / 64];
    const int plane = blockIdx.y;
    const u64 *bits = planes + (size_t)plane * packed;
    F::E acc = F::e_zero();
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
        if ((bits[i >> 6] >> (i & 63)) & 1) acc = F::e_add(acc, F::e_load(eq, i));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        const u64 o0 = ((u64)__shfl_xor((u32)(acc.c0 >> 32), off) << 32) | (u32)__shfl_xor((u32)acc.c0, off);
        const u64 o1 = ((u64)__shfl_xor((u32)(acc.c1 >> 32), off) << 32) | (u32)__shfl_xor((u32)acc.c1, off);
        acc = F::e_add(acc, F::e(o0, o1));
    }
    if (lane == 0) {
        lds[2 * wave] = acc.c0;
        lds[2 * wave + 1] = acc.c1;
    }
    mam_barrier();
    if (threadIdx.x == 0) {
        F::E sum = F::e_zero();
        for (int w = 0; w < (int)blockDim.x / 64; w++) sum = F::e_add(sum, F::e(lds[2 * w], lds[2 * w + 1]));
        F::e_store(partial, (size_t)blockIdx.x * n_planes + plane, sum);
    }
}

// ffi.rs:283-291 / almost wrapper.cu:875-885
MAM_API int aext2_selective_add_batched_planes_ffi(const u64 *d_eq, const u64 *d_packed_planes, u64 *d_partial,
                                                   size_t total, int n_planes, size_t packed_size_u64,
                                                   int num_blocks_x) {
    if (n_planes <= 0) return 0;
    if (num_blocks_x <= 0 || n_planes > 65535 || !d_eq || !d_packed_planes || !d_partial) return -1;
    if (packed_size_u64 * 64 < total) return -1;
    k_selective_add_planes<<<dim3(num_blocks_x, n_planes), MAM_BLOCK, 0, mam_stream()>>>(
        d_eq, d_packed_planes, d_partial, total, n_planes, packed_size_u64);
    return mam_launch_status();
}

// ffi.rs:265-272, :320-328
MAM_API int aext2_sumcheck_batched_round_message_ffi(const u64 *d_polys, u64 *d_partial, size_t original_size,
                                                     size_t half, int num_leaves, int num_blocks_x) {
    if (num_leaves <= 0) return 0;
    if (num_blocks_x <= 0 || num_leaves > 65535) return -1;
    k_batched_msg_f<FieldQ><<<dim3(num_blocks_x, num_leaves), MAM_BLOCK, 0, mam_stream()>>>(d_polys, d_partial,
                                                                                         original_size, half,
                                                                                         num_leaves);
    return mam_sync_status();
}
MAM_API int aext2_sumcheck_batched_fold_ffi(const u64 *d_input, u64 *d_output, u64 c0, u64 c1, size_t original_size,
                                            size_t half, int num_leaves) {
    if (num_leaves <= 0 || half == 0) return 0;
    if (num_leaves > 65535) return -1;
    unsigned gx = grid_stride(half, MAM_BLOCK, 1024);
    k_batched_fold_f<FieldQ><<<dim3(gx, num_leaves), MAM_BLOCK, 0, mam_stream()>>>(
        d_input, d_output, qe2{fq::canon(c0), fq::canon(c1)}, original_size, half);
    return mam_sync_status();
}
