// Basefold commitment and opening kernels, eq tables and partial evaluation.
//
// Reference semantics:
//   bit reversal            cuda/basefold.cuh:57-90
//   BHC interpolation       cuda/basefold.cuh:103-147, goldilocks-cuda-rs/cuda/wrapper.cu:1242-1274
//   encode                  cuda/basefold.cuh:157-205, wrapper.cu:1277-1312
//   codeword folds          cuda/basefold.cuh:397-419, 522-548, 739-761
//   opening sumcheck pieces cuda/basefold.cuh:287-383, 427-735, 966-1006
//   eq tables               cuda/eq_lagrange.cuh:22-158 (bit i of x <-> r_i, LSB first)
//   partial evaluation      cuda/partial_eval.cuh:31-219
//
// MI355X design.
// * The interpolation (Moebius transform: hi -= lo per bit) and the encode (Walsh-Hadamard
//   butterflies lo+hi, lo-hi per bit over the repetition code, then bit reversal) are each
//   a tensor product of one 2x2 matrix per bit, so each runs as ceil(bits/8) passes of a
//   radix-256 transform held in registers with one LDS exchange per pass (k_bfly_reg)
//   instead of one strided launch per bit.
// * Bit reversal commutes with any transform that applies the same 2x2 matrix to every bit.
//   Hence encode(coeffs) = replicate(WHT(bitrev(coeffs))), and the whole commit codeword
//   equals replicate(T(bitrev(evals))) with T = (lo, hi) -> (hi, 2 lo - hi) per bit
//   (the Moebius matrix followed by the Hadamard one). The fused commit path uses that
//   single transform on bh_evals = bitrev(evals), which the commitment keeps anyway.
// * Bit reversal is out of place through 32x32 LDS tiles, so reads and writes are both
//   256-byte coalesced.
#include "basefold.h"
#include "common.h"
#include "eq_table.h"
#include "open_fold.h"
#include "profile.h"
#include "reduce.h"
#include "sync.h"

// ---------------------------------------------------------------------------
// Radix-256 butterfly transform passes
// ---------------------------------------------------------------------------
enum { BF_MOBIUS = 0, BF_WHT = 1, BF_T = 2 };

template <int OP>
__device__ __forceinline__ void bfly(u64 &lo, u64 &hi) {
    if (OP == BF_MOBIUS) {
        hi = gl_sub(hi, lo);
    } else if (OP == BF_WHT) {
        u64 a = gl_add(lo, hi), b = gl_sub(lo, hi);
        lo = a;
        hi = b;
    } else {
        u64 a = hi, b = gl_sub(gl_double(lo), hi);
        lo = a;
        hi = b;
    }
}

static constexpr int BF_MAXKB = 8;  // row bits per pass

// Register-resident form of one pass over bits [b0, b0 + KB) (KB = 4..8): the workgroup's
// tile is 2^KB rows (the pass's bits) x C = 4096 / 2^KB columns (the other bits), column
// gc = (hi, lo) holding element hi * 2^(b0+KB) + row * 2^b0 + lo. Every thread keeps 16 of
// the tile's words in registers twice: first the 16 rows that differ in the top 4 row bits
// (those 4 butterfly stages run in registers), then, after one exchange through LDS, 2^(KB-4)
// rows x 2^(8-KB) columns (the remaining KB - 4 stages). Each word crosses LDS once per pass,
// where the radix-2 form above reads and writes it once per bit, which left the pass bound by
// LDS bandwidth rather than by HBM. Global accesses are coalesced for b0 = 0 (columns are
// contiguous runs of rows) and for b0 >= 4 (rows are runs of at least 16 columns).
// src may alias dst: the workgroup reads its whole tile before it writes.
//
// BREV (first pass of the commit, b0 = 0): the array transformed is bh = bitrev(src) over
// nb bits, written to `side` as well. Column m of the workgroup is bh's column
// gc = bitrev(m) (nb - KB bits), whose row r is src[bitrev_KB(r) * 2^(nb-KB) + m], so the
// workgroup reads runs of consecutive source words, and the bit reversal costs no pass of
// its own.
template <int OP, int KB, bool BREV>
__global__ __launch_bounds__(256) void k_bfly_reg(const u64 *src, u64 *dst, int b0, int reps, size_t rep_stride,
                                                  u64 *side, int nb) {
    constexpr int C = 4096 >> KB;        // columns per workgroup
    constexpr int RA = 1 << (KB - 4);    // phase A: rows per register step
    constexpr int KBB = KB - 4;          // phase B: row bits in registers
    constexpr int CB = 1 << (8 - KB);    // phase B: columns per thread
    __shared__ u64 t[4096 + 4096 / 32];
#define LIDX(e) ((e) + ((e) >> 5))
    const int tid = threadIdx.x;
    const size_t lowmask = ((size_t)1 << b0) - 1;
    const size_t gc0 = (size_t)blockIdx.x * C;
    // address of (row, column c of this workgroup)
    auto addr = [&](int row, int c) -> size_t {
        size_t gc = gc0 + c;
        if (BREV) gc = __brevll(gc) >> (64 - (nb - KB));
        return ((gc >> b0) << (b0 + KB)) + ((size_t)row << b0) + (gc & lowmask);
    };
    // phase A: rows i * RA + rl, i < 16, of column c
    int rl, c;
    if (b0 == 0) {
        rl = tid % RA;
        c = tid / RA;
    } else {
        c = tid % C;
        rl = tid / C;
    }
    u64 v[16];
    if (BREV) {
        // gather with consecutive lanes on consecutive source words (columns), then turn
        // the tile through LDS into the phase-A layout and write bh from it
        const int cc = tid % C, rr = tid / C;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int row = i * RA + rr;
            const size_t sidx = ((size_t)(__brev((u32)row) >> (32 - KB)) << (nb - KB)) + gc0 + cc;
            t[LIDX(row * C + cc)] = gl_canon(src[sidx]);
        }
        mam_barrier();
#pragma unroll
        for (int i = 0; i < 16; i++) {
            v[i] = t[LIDX((i * RA + rl) * C + c)];
            side[addr(i * RA + rl, c)] = v[i];
        }
        mam_barrier();
    } else {
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = gl_canon(src[addr(i * RA + rl, c)]);
    }
#pragma unroll
    for (int st = 0; st < 4; st++) {
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (!(i & (1 << st))) bfly<OP>(v[i], v[i | (1 << st)]);
    }
#pragma unroll
    for (int i = 0; i < 16; i++) t[LIDX((i * RA + rl) * C + c)] = v[i];
    mam_barrier();
    // phase B: rows (rh << KBB) + rlow, rlow < 2^KBB, of columns cb * CB + cp, cp < CB;
    // register k = rlow * CB + cp
    int rh, cb;
    if (b0 == 0) {
        rh = tid % 16;
        cb = tid / 16;
    } else {
        cb = tid % (C / CB);
        rh = tid / (C / CB);
    }
#pragma unroll
    for (int k = 0; k < 16; k++) v[k] = t[LIDX((((rh << KBB) + k / CB)) * C + cb * CB + k % CB)];
#pragma unroll
    for (int st = 0; st < KBB; st++) {
#pragma unroll
        for (int k = 0; k < 16; k++)
            if (!(k & (CB << st))) bfly<OP>(v[k], v[k | (CB << st)]);
    }
    if (b0 == 0) {
        // columns are runs of rows: store from the phase-A layout, where consecutive lanes
        // hold consecutive rows (in the phase-B layout each lane holds a run of its own)
        mam_barrier();
#pragma unroll
        for (int k = 0; k < 16; k++) t[LIDX((((rh << KBB) + k / CB)) * C + cb * CB + k % CB)] = v[k];
        mam_barrier();
#pragma unroll
        for (int i = 0; i < 16; i++) v[i] = t[LIDX((i * RA + rl) * C + c)];
        for (int r = 0; r < reps; r++) {
            u64 *d = dst + r * rep_stride;
#pragma unroll
            for (int i = 0; i < 16; i++) d[addr(i * RA + rl, c)] = v[i];
        }
    } else {
        for (int r = 0; r < reps; r++) {
            u64 *d = dst + r * rep_stride;
#pragma unroll
            for (int k = 0; k < 16; k++) d[addr((rh << KBB) + k / CB, cb * CB + k % CB)] = v[k];
        }
    }
#undef LIDX
}

__global__ void k_bfly_small(const u64 *src, u64 *dst, int nbits, int op, int reps,
                             size_t rep_stride) {
    // whole array (at most 2^12 words) in one workgroup
    __shared__ u64 t[4096];
    const int n = 1 << nbits;
    for (int e = threadIdx.x; e < n; e += blockDim.x) t[e] = gl_canon(src[e]);
    mam_barrier();
    for (int j = 0; j < nbits; j++) {
        for (int p = threadIdx.x; p < n / 2; p += blockDim.x) {
            int lo_i = ((p >> j) << (j + 1)) | (p & ((1 << j) - 1));
            int hi_i = lo_i | (1 << j);
            u64 lo = t[lo_i], hi = t[hi_i];
            if (op == BF_MOBIUS) bfly<BF_MOBIUS>(lo, hi);
            else if (op == BF_WHT) bfly<BF_WHT>(lo, hi);
            else bfly<BF_T>(lo, hi);
            t[lo_i] = lo;
            t[hi_i] = hi;
        }
        mam_barrier();
    }
    for (int r = 0; r < reps; r++)
        for (int e = threadIdx.x; e < n; e += blockDim.x) dst[r * rep_stride + e] = t[e];
}

// Transform src (2^nbits words) into dst; the final result is also replicated `reps`
// times at stride 2^nbits in dst. src may equal dst. brev_side (op BF_T, nbits > 12 only):
// transform bitrev(src) instead and write bitrev(src) to brev_side on the way (the commit).
static int bfly_transform(int op, const u64 *src, u64 *dst, int nbits, int reps, hipStream_t s,
                          u64 *brev_side = nullptr) {
    const size_t n = (size_t)1 << nbits;
    if (brev_side && (op != BF_T || nbits <= 12)) return -2;
    if (nbits <= 12) {
        k_bfly_small<<<1, 256, 0, s>>>(src, dst, nbits, op, reps, n);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    }
    int b0 = 0;
    const u64 *in = src;
    int passes = (nbits + BF_MAXKB - 1) / BF_MAXKB;
    while (b0 < nbits) {
        // split the bits evenly over the remaining passes (each at most 8 bits; with
        // nbits >= 13 every pass has 5 to 8, and each b0 after the first is at least 5)
        int kb = (nbits - b0 + passes - 1) / passes;
        passes--;
        bool last = b0 + kb == nbits;
        int r = last ? reps : 1;
        const unsigned grid = (unsigned)((n >> kb) >> (12 - kb));
#define MAM_BF_REG(K)                                                                                        \
    case K:                                                                                                   \
        if (brev_side && b0 == 0) k_bfly_reg<BF_T, K, true><<<grid, 256, 0, s>>>(in, dst, b0, r, n, brev_side, nbits); \
        else if (op == BF_MOBIUS) k_bfly_reg<BF_MOBIUS, K, false><<<grid, 256, 0, s>>>(in, dst, b0, r, n, nullptr, 0); \
        else if (op == BF_WHT) k_bfly_reg<BF_WHT, K, false><<<grid, 256, 0, s>>>(in, dst, b0, r, n, nullptr, 0);      \
        else k_bfly_reg<BF_T, K, false><<<grid, 256, 0, s>>>(in, dst, b0, r, n, nullptr, 0);                          \
        break;
        switch (kb) {
            MAM_BF_REG(5)
            MAM_BF_REG(6)
            MAM_BF_REG(7)
            MAM_BF_REG(8)
            default: return -2;
        }
#undef MAM_BF_REG
        if (hipGetLastError() != hipSuccess) return -1;
        in = dst;
        b0 += kb;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Bit reversal (out of place), W words per element
// ---------------------------------------------------------------------------
__device__ __forceinline__ u32 brev_bits(u32 x, int bits) { return bits == 0 ? 0 : (__brev(x) >> (32 - bits)); }

template <int W>
__global__ __launch_bounds__(256) void k_bitrev_tiled(const u64 *__restrict__ src, u64 *__restrict__ dst, int nb) {
    // element index = a (top 5 bits) | m (middle nb - 10 bits) | b (low 5 bits)
    __shared__ u64 tile[32][33 * W];
    const u32 m = blockIdx.x;
    const int mb = nb - 10;
    for (int e = threadIdx.x; e < 1024; e += 256) {
        u32 a = e >> 5, b = e & 31;
        size_t i = ((size_t)a << (nb - 5)) | ((size_t)m << 5) | b;
#pragma unroll
        for (int w = 0; w < W; w++) tile[a][b * W + w] = gl_canon(src[i * W + w]);
    }
    mam_barrier();
    const u32 rm = brev_bits(m, mb);
    for (int e = threadIdx.x; e < 1024; e += 256) {
        u32 b = e >> 5, ap = e & 31;
        u32 a = brev_bits(ap, 5);
        size_t o = ((size_t)brev_bits(b, 5) << (nb - 5)) | ((size_t)rm << 5) | ap;
#pragma unroll
        for (int w = 0; w < W; w++) dst[o * W + w] = tile[a][b * W + w];
    }
}

template <int W>
__global__ void k_bitrev_simple(const u64 *__restrict__ src, u64 *__restrict__ dst, int nb) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= ((size_t)1 << nb)) return;
    size_t r = nb == 0 ? 0 : (size_t)(__brevll((unsigned long long)i) >> (64 - nb));
#pragma unroll
    for (int w = 0; w < W; w++) dst[r * W + w] = gl_canon(src[i * W + w]);
}

// Out-of-place bit reversal; src and dst must not overlap.
int mam_bitrev(const u64 *src, u64 *dst, int nb, int words, hipStream_t s) {
    size_t n = (size_t)1 << nb;
    if (nb >= 10) {
        unsigned grid = (unsigned)(n >> 10);
        if (words == 1) k_bitrev_tiled<1><<<grid, 256, 0, s>>>(src, dst, nb);
        else k_bitrev_tiled<2><<<grid, 256, 0, s>>>(src, dst, nb);
    } else {
        if (words == 1) k_bitrev_simple<1><<<grid_1d(n), 256, 0, s>>>(src, dst, nb);
        else k_bitrev_simple<2><<<grid_1d(n), 256, 0, s>>>(src, dst, nb);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// In-place bit reversal through a temporary (reference entry points are in place).
static int bitrev_inplace(u64 *data, int nb, int words, hipStream_t s) {
    size_t bytes = ((size_t)words << nb) * sizeof(u64);
    u64 *tmp = nullptr;
    if (hipMallocAsync((void **)&tmp, bytes, s) != hipSuccess) return -1;
    int rc = mam_bitrev(data, tmp, nb, words, s);
    if (!rc && hipMemcpyAsync(data, tmp, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) rc = -1;
    hipFreeAsync(tmp, s);
    return rc;
}

MAM_API int basefold_bit_reverse_gl_ffi(u64 *d_data, int log_n) {
    if (log_n < 0 || log_n > 40) return -1;
    if (bitrev_inplace(d_data, log_n, 1, mam_stream())) return -1;
    return mam_launch_status();
}
MAM_API int basefold_bit_reverse_ext2_ffi(u64 *d_data, int log_n) {
    if (log_n < 0 || log_n > 40) return -1;
    if (bitrev_inplace(d_data, log_n, 2, mam_stream())) return -1;
    return mam_launch_status();
}

// ---------------------------------------------------------------------------
// Commit: interpolation and encode (reference-named), and the fused commit transform
// ---------------------------------------------------------------------------

// coeffs = Moebius(evals) (natural order); bh_evals = bitrev(evals).
MAM_API int basefold_bhc_interpolate_ffi(const u64 *d_evals, u64 *d_coeffs, u64 *d_bh_evals, int num_vars) {
    if (num_vars < 1 || num_vars > 34) return -1;
    hipStream_t s = mam_stream();
    if (bfly_transform(BF_MOBIUS, d_evals, d_coeffs, num_vars, 1, s)) return -1;
    if (mam_bitrev(d_evals, d_bh_evals, num_vars, 1, s)) return -1;
    return mam_sync_status();
}

// codeword = replicate(WHT(bitrev(coeffs))), 2^log_rate copies of 2^num_vars words.
MAM_API int basefold_encode_ffi(const u64 *d_coeffs, u64 *d_codeword, int num_vars, int log_rate) {
    if (num_vars < 1 || log_rate < 0 || num_vars + log_rate > 36) return -1;
    hipStream_t s = mam_stream();
    if (mam_bitrev(d_coeffs, d_codeword, num_vars, 1, s)) return -1;
    if (bfly_transform(BF_WHT, d_codeword, d_codeword, num_vars, 1 << log_rate, s)) return -1;
    return mam_sync_status();
}

// Fused commit codeword: bh_evals = bitrev(evals); codeword = replicate(T(bh_evals)).
int mam_commit_codeword(const u64 *d_evals, u64 *d_bh_evals, u64 *d_codeword, int nv, int lr, hipStream_t s) {
    // algorithmic bytes: read evals, write bh_evals, write 2^lr codeword copies
    ProfScope prof(PROF_COMMIT_TRANSFORM, 8.0 * (double)((size_t)1 << nv) * (2.0 + (double)(1 << lr)), s);
    if (nv > 12) return bfly_transform(BF_T, d_evals, d_codeword, nv, 1 << lr, s, d_bh_evals);
    if (mam_bitrev(d_evals, d_bh_evals, nv, 1, s)) return -1;
    return bfly_transform(BF_T, d_bh_evals, d_codeword, nv, 1 << lr, s);
}

// ---------------------------------------------------------------------------
// Codeword folds: out[i] = v0 + (alpha - x0_i) * (v1 - v0) * w_i
// ---------------------------------------------------------------------------
template <int KIND>  // 0: base cw, base alpha; 1: base cw, ext2 alpha; 2: ext2 cw, ext2 alpha
__global__ void k_cw_fold(const u64 *__restrict__ cw, const u64 *__restrict__ table, ext2 alpha,
                          u64 *__restrict__ out, size_t pairs, const u64 *alpha_p = nullptr) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= pairs) return;
    if (alpha_p) alpha = ext2{alpha_p[0], alpha_p[1]};  // written on the device after the launch
    const u64 x0 = gl_canon(table[2 * i]), w = gl_canon(table[2 * i + 1]);
    if (KIND == 0) {
        u64 v0 = gl_canon(cw[2 * i]), v1 = gl_canon(cw[2 * i + 1]);
        out[i] = gl_add(v0, gl_mul(gl_mul(gl_sub(alpha.c0, x0), gl_sub(v1, v0)), w));
    } else {
        ext2 v0, v1;
        if (KIND == 1) {
            v0 = e2_base(gl_canon(cw[2 * i]));
            v1 = e2_base(gl_canon(cw[2 * i + 1]));
        } else {
            v0 = e2_load(cw, 2 * i);
            v1 = e2_load(cw, 2 * i + 1);
        }
        ext2 cx = e2_sub(alpha, e2_base(x0));
        e2_store(out, i, mam_fold_value(v0, v1, cx, w));
    }
}

MAM_API int basefold_fold_gl_ffi(const u64 *d_codeword, const u64 *d_table, u64 challenge, u64 *d_output,
                                 int pair_count) {
    if (pair_count <= 0) return 0;
    k_cw_fold<0><<<grid_1d(pair_count), MAM_BLOCK, 0, mam_stream()>>>(d_codeword, d_table,
                                                                       e2_base(gl_canon(challenge)), d_output,
                                                                       (size_t)pair_count);
    return mam_launch_status();
}
MAM_API int basefold_fold_mixed_ffi(const u64 *d_codeword, const u64 *d_table, u64 c0, u64 c1, u64 *d_output,
                                    int pair_count) {
    if (pair_count <= 0) return 0;
    k_cw_fold<1><<<grid_1d(pair_count), MAM_BLOCK, 0, mam_stream()>>>(d_codeword, d_table,
                                                                       e2(gl_canon(c0), gl_canon(c1)), d_output,
                                                                       (size_t)pair_count);
    return mam_launch_status();
}
MAM_API int basefold_fold_ext2_ffi(const u64 *d_codeword, const u64 *d_table, u64 c0, u64 c1, u64 *d_output,
                                   int pair_count) {
    if (pair_count <= 0) return 0;
    k_cw_fold<2><<<grid_1d(pair_count), MAM_BLOCK, 0, mam_stream()>>>(d_codeword, d_table,
                                                                       e2(gl_canon(c0), gl_canon(c1)), d_output,
                                                                       (size_t)pair_count);
    return mam_launch_status();
}

int mam_cw_fold(const u64 *cw, const u64 *table, ext2 alpha, u64 *out, size_t pairs, bool base_cw, hipStream_t s,
                const u64 *alpha_p) {
    if (base_cw)
        k_cw_fold<1><<<grid_1d(pairs), MAM_BLOCK, 0, s>>>(cw, table, alpha, out, pairs, alpha_p);
    else
        k_cw_fold<2><<<grid_1d(pairs), MAM_BLOCK, 0, s>>>(cw, table, alpha, out, pairs, alpha_p);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------------------
// Opening-sumcheck pieces (reference-named; partial sums per workgroup)
// ---------------------------------------------------------------------------
template <bool EXT>
__global__ void k_interp(u64 *__restrict__ d, size_t pairs) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= pairs) return;
    if (EXT) {
        ext2 a = e2_load(d, 2 * i), b = e2_load(d, 2 * i + 1);
        e2_store(d, 2 * i, a);
        e2_store(d, 2 * i + 1, e2_sub(b, a));
    } else {
        u64 a = gl_canon(d[2 * i]), b = gl_canon(d[2 * i + 1]);
        d[2 * i] = a;
        d[2 * i + 1] = gl_sub(b, a);
    }
}

template <int KIND>  // 0: base data, base challenge; 1: base data, ext2 challenge; 2: ext2 data
__global__ void k_eval(const u64 *__restrict__ d, ext2 ch, u64 *__restrict__ out, size_t pairs) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= pairs) return;
    if (KIND == 0) {
        out[i] = gl_add(gl_canon(d[2 * i]), gl_mul(ch.c0, gl_canon(d[2 * i + 1])));
    } else if (KIND == 1) {
        u64 c = gl_canon(d[2 * i]), df = gl_canon(d[2 * i + 1]);
        e2_store(out, i, e2_add(e2_base(c), e2_scale(df, ch)));
    } else {
        e2_store(out, i, e2_add(e2_load(d, 2 * i), e2_mul(ch, e2_load(d, 2 * i + 1))));
    }
}

// eq and bh in coefficient form (value, difference): c0 += e0 b0, c1 += e0 b1 + e1 b0, c2 += e1 b1
template <int KIND>  // 0: both base; 1: eq ext2, bh base; 2: both ext2
__global__ __launch_bounds__(MAM_BLOCK) void k_product(const u64 *__restrict__ eq, const u64 *__restrict__ bh,
                                                       u64 *pc0, u64 *pc1, u64 *pc2, size_t pairs) {
    __shared__ u64 lds[6 * (MAM_BLOCK / WAVE)];
    u64 acc[6] = {0, 0, 0, 0, 0, 0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < pairs; i += (size_t)gridDim.x * blockDim.x) {
        ext2 e0, e1, b0, b1;
        if (KIND == 0) {
            e0 = e2_base(gl_canon(eq[2 * i]));
            e1 = e2_base(gl_canon(eq[2 * i + 1]));
        } else {
            e0 = e2_load(eq, 2 * i);
            e1 = e2_load(eq, 2 * i + 1);
        }
        if (KIND == 2) {
            b0 = e2_load(bh, 2 * i);
            b1 = e2_load(bh, 2 * i + 1);
        } else {
            b0 = e2_base(gl_canon(bh[2 * i]));
            b1 = e2_base(gl_canon(bh[2 * i + 1]));
        }
        mam_oracle_accumulate(acc, e0, e1, b0, b1);
    }
    block_sum<6>(acc, lds);
    if (threadIdx.x == 0) {
        if (KIND == 0) {
            pc0[blockIdx.x] = acc[0];
            pc1[blockIdx.x] = acc[2];
            pc2[blockIdx.x] = acc[4];
        } else {
            pc0[2 * blockIdx.x] = acc[0];
            pc0[2 * blockIdx.x + 1] = acc[1];
            pc1[2 * blockIdx.x] = acc[2];
            pc1[2 * blockIdx.x + 1] = acc[3];
            pc2[2 * blockIdx.x] = acc[4];
            pc2[2 * blockIdx.x + 1] = acc[5];
        }
    }
}

// Fused eval + interp + product over 4*pairs inputs (cuda/basefold.cuh:654-735).
__global__ __launch_bounds__(MAM_BLOCK) void k_fused_round_ref(const u64 *__restrict__ eq_in,
                                                               const u64 *__restrict__ bh_in, ext2 ch,
                                                               u64 *__restrict__ eq_out, u64 *__restrict__ bh_out,
                                                               u64 *pc0, u64 *pc1, u64 *pc2, size_t pairs) {
    __shared__ u64 lds[6 * (MAM_BLOCK / WAVE)];
    u64 acc[6] = {0, 0, 0, 0, 0, 0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < pairs; i += (size_t)gridDim.x * blockDim.x) {
        ext2 ea = e2_add(e2_load(eq_in, 4 * i), e2_mul(ch, e2_load(eq_in, 4 * i + 1)));
        ext2 eb = e2_add(e2_load(eq_in, 4 * i + 2), e2_mul(ch, e2_load(eq_in, 4 * i + 3)));
        ext2 ba = e2_add(e2_load(bh_in, 4 * i), e2_mul(ch, e2_load(bh_in, 4 * i + 1)));
        ext2 bb = e2_add(e2_load(bh_in, 4 * i + 2), e2_mul(ch, e2_load(bh_in, 4 * i + 3)));
        ext2 e1 = e2_sub(eb, ea), b1 = e2_sub(bb, ba);
        mam_oracle_accumulate(acc, ea, e1, ba, b1);
        e2_store(eq_out, 2 * i, ea);
        e2_store(eq_out, 2 * i + 1, e1);
        e2_store(bh_out, 2 * i, ba);
        e2_store(bh_out, 2 * i + 1, b1);
    }
    block_sum<6>(acc, lds);
    if (threadIdx.x == 0) {
        pc0[2 * blockIdx.x] = acc[0];
        pc0[2 * blockIdx.x + 1] = acc[1];
        pc1[2 * blockIdx.x] = acc[2];
        pc1[2 * blockIdx.x + 1] = acc[3];
        pc2[2 * blockIdx.x] = acc[4];
        pc2[2 * blockIdx.x + 1] = acc[5];
    }
}

template <bool MIXED>
__global__ __launch_bounds__(MAM_BLOCK) void k_dot(const u64 *__restrict__ a, const u64 *__restrict__ b,
                                                   u64 *partial, size_t n) {
    __shared__ u64 lds[2 * (MAM_BLOCK / WAVE)];
    u64 acc[2] = {0, 0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        if (MIXED) {
            ext2 p = e2_scale(gl_canon(a[i]), e2_load(b, i));
            acc[0] = gl_add(acc[0], p.c0);
            acc[1] = gl_add(acc[1], p.c1);
        } else {
            acc[0] = gl_add(acc[0], gl_mul(gl_canon(a[i]), gl_canon(b[i])));
        }
    }
    block_sum<2>(acc, lds);
    if (threadIdx.x == 0) {
        if (MIXED) {
            partial[2 * blockIdx.x] = acc[0];
            partial[2 * blockIdx.x + 1] = acc[1];
        } else {
            partial[blockIdx.x] = acc[0];
        }
    }
}

MAM_API int basefold_sumcheck_interp_gl_ffi(u64 *d_data, int pair_count) {
    if (pair_count <= 0) return 0;
    k_interp<false><<<grid_1d(pair_count), MAM_BLOCK, 0, mam_stream()>>>(d_data, (size_t)pair_count);
    return mam_launch_status();
}
MAM_API int basefold_sumcheck_interp_ext2_ffi(u64 *d_data, int pair_count) {
    if (pair_count <= 0) return 0;
    k_interp<true><<<grid_1d(pair_count), MAM_BLOCK, 0, mam_stream()>>>(d_data, (size_t)pair_count);
    return mam_launch_status();
}
MAM_API int basefold_sumcheck_eval_gl_ffi(const u64 *d_data, u64 challenge, u64 *d_output, int pair_count) {
    if (pair_count <= 0) return 0;
    k_eval<0><<<grid_1d(pair_count), MAM_BLOCK, 0, mam_stream()>>>(d_data, e2_base(gl_canon(challenge)), d_output,
                                                                     (size_t)pair_count);
    return mam_launch_status();
}
MAM_API int basefold_sumcheck_eval_mixed_ffi(const u64 *d_data, u64 c0, u64 c1, u64 *d_output, int pair_count) {
    if (pair_count <= 0) return 0;
    k_eval<1><<<grid_1d(pair_count), MAM_BLOCK, 0, mam_stream()>>>(d_data, e2(gl_canon(c0), gl_canon(c1)), d_output,
                                                                     (size_t)pair_count);
    return mam_launch_status();
}
MAM_API int basefold_sumcheck_eval_ext2_ffi(const u64 *d_data, u64 c0, u64 c1, u64 *d_output, int pair_count) {
    if (pair_count <= 0) return 0;
    k_eval<2><<<grid_1d(pair_count), MAM_BLOCK, 0, mam_stream()>>>(d_data, e2(gl_canon(c0), gl_canon(c1)), d_output,
                                                                     (size_t)pair_count);
    return mam_launch_status();
}

#define PRODUCT_FFI(name, kind)                                                                              \
    MAM_API int name(const u64 *d_eq, const u64 *d_bh, u64 *pc0, u64 *pc1, u64 *pc2, int pair_count,        \
                     int num_blocks) {                                                                       \
        if (num_blocks <= 0) return -1;                                                                      \
        k_product<kind><<<num_blocks, MAM_BLOCK, 0, mam_stream()>>>(d_eq, d_bh, pc0, pc1, pc2,              \
                                                                    pair_count > 0 ? (size_t)pair_count : 0); \
        return mam_launch_status();                                                                          \
    }
PRODUCT_FFI(basefold_sumcheck_product_gl_ffi, 0)
PRODUCT_FFI(basefold_sumcheck_product_mixed_ffi, 1)
PRODUCT_FFI(basefold_sumcheck_product_ext2_ffi, 2)

MAM_API int fused_sumcheck_round_ext2_ffi(const u64 *d_eq_in, const u64 *d_bh_in, u64 c0, u64 c1, u64 *d_eq_out,
                                          u64 *d_bh_out, u64 *pc0, u64 *pc1, u64 *pc2, int pair_count,
                                          int num_blocks) {
    if (num_blocks <= 0) return -1;
    k_fused_round_ref<<<num_blocks, MAM_BLOCK, 0, mam_stream()>>>(d_eq_in, d_bh_in, e2(gl_canon(c0), gl_canon(c1)),
                                                                  d_eq_out, d_bh_out, pc0, pc1, pc2,
                                                                  pair_count > 0 ? (size_t)pair_count : 0);
    return mam_launch_status();
}
MAM_API int basefold_dot_product_gl_ffi(const u64 *d_a, const u64 *d_b, u64 *d_partial, int n, int num_blocks) {
    if (num_blocks <= 0) return -1;
    k_dot<false><<<num_blocks, MAM_BLOCK, 0, mam_stream()>>>(d_a, d_b, d_partial, n > 0 ? (size_t)n : 0);
    return mam_launch_status();
}
MAM_API int basefold_dot_product_mixed_ffi(const u64 *d_a, const u64 *d_b, u64 *d_partial, int n, int num_blocks) {
    if (num_blocks <= 0) return -1;
    k_dot<true><<<num_blocks, MAM_BLOCK, 0, mam_stream()>>>(d_a, d_b, d_partial, n > 0 ? (size_t)n : 0);
    return mam_launch_status();
}

// ---------------------------------------------------------------------------
// eq tables (eq_table.h): one pass, one product per entry.
// ---------------------------------------------------------------------------
int mam_eq_table(const u64 *d_r, int nv, u64 *d_out, bool ext, hipStream_t s) {
    if (nv < 0 || nv > 40) return -2;
    return ext ? eqt::launch<FieldGL, true>(d_r, 0, nv, d_out, 0, 1, s)
               : eqt::launch<FieldGL, false>(d_r, 0, nv, d_out, 0, 1, s);
}

MAM_API int eq_dp_all_ffi(const u64 *d_r, u64 *d_buf_a, u64 *d_buf_b, int log_n, u64 **d_result) {
    (void)d_buf_b;
    if (mam_eq_table(d_r, log_n, d_buf_a, false, mam_stream())) return -1;
    if (d_result) *d_result = d_buf_a;
    return mam_launch_status();
}
MAM_API int ext2_eq_dp_all_ffi(const u64 *d_r, u64 *d_buf_a, u64 *d_buf_b, int log_n, u64 **d_result) {
    (void)d_buf_b;
    if (mam_eq_table(d_r, log_n, d_buf_a, true, mam_stream())) return -1;
    if (d_result) *d_result = d_buf_a;
    return mam_launch_status();
}

// ---------------------------------------------------------------------------
// Partial evaluation: fold the first m variables (LSB first)
// ---------------------------------------------------------------------------
template <int KIND>  // 0: base->base (base r); 1: base->ext2 (ext2 r); 2: ext2->ext2
__global__ void k_peval_layer(const u64 *__restrict__ in, u64 *__restrict__ out, const u64 *__restrict__ r, int li,
                              size_t pairs) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= pairs) return;
    if (KIND == 0) {
        u64 ri = gl_canon(r[li]);
        u64 a = gl_canon(in[2 * i]), b = gl_canon(in[2 * i + 1]);
        out[i] = gl_add(a, gl_mul(ri, gl_sub(b, a)));
    } else if (KIND == 1) {
        ext2 ri = e2_load(r, li);
        u64 a = gl_canon(in[2 * i]), b = gl_canon(in[2 * i + 1]);
        e2_store(out, i, e2_add(e2_base(a), e2_scale(gl_sub(b, a), ri)));
    } else {
        ext2 ri = e2_load(r, li);
        ext2 a = e2_load(in, 2 * i), b = e2_load(in, 2 * i + 1);
        e2_store(out, i, e2_add(a, e2_mul(ri, e2_sub(b, a))));
    }
}

MAM_API int partial_eval_gl_ffi(u64 *d_data, const u64 *d_r, int log_n, int m) {
    if (m <= 0) return 0;
    if (m > log_n) return -1;
    hipStream_t s = mam_stream();
    size_t cur = (size_t)1 << log_n;
    u64 *scratch = nullptr;
    if (hipMallocAsync((void **)&scratch, (cur / 2) * sizeof(u64), s) != hipSuccess) return -1;
    u64 *src = d_data, *dst = scratch;
    int rc = 0;
    for (int i = 0; i < m && !rc; i++) {
        size_t pairs = cur / 2;
        k_peval_layer<0><<<grid_1d(pairs), MAM_BLOCK, 0, s>>>(src, dst, d_r, i, pairs);
        if (hipGetLastError() != hipSuccess) rc = -1;
        u64 *t = src;
        src = dst;
        dst = t;
        cur = pairs;
    }
    if (!rc && src != d_data &&
        hipMemcpyAsync(d_data, src, cur * sizeof(u64), hipMemcpyDeviceToDevice, s) != hipSuccess)
        rc = -1;
    hipFreeAsync(scratch, s);
    return rc ? rc : mam_sync_status();
}

MAM_API int partial_eval_ext2_from_gl_ffi(const u64 *d_input, u64 *d_output, const u64 *d_r, int log_n, int m) {
    if (m <= 0) return 0;
    if (m > log_n) return -1;
    hipStream_t s = mam_stream();
    size_t cur = (size_t)1 << log_n;
    size_t pairs = cur / 2;
    k_peval_layer<1><<<grid_1d(pairs), MAM_BLOCK, 0, s>>>(d_input, d_output, d_r, 0, pairs);
    if (hipGetLastError() != hipSuccess) return -1;
    cur = pairs;
    if (m == 1) return mam_sync_status();
    u64 *scratch = nullptr;
    if (hipMallocAsync((void **)&scratch, (cur / 2 > 0 ? cur / 2 : 1) * 2 * sizeof(u64), s) != hipSuccess) return -1;
    u64 *src = d_output, *dst = scratch;
    int rc = 0;
    for (int i = 1; i < m && !rc; i++) {
        pairs = cur / 2;
        k_peval_layer<2><<<grid_1d(pairs), MAM_BLOCK, 0, s>>>(src, dst, d_r, i, pairs);
        if (hipGetLastError() != hipSuccess) rc = -1;
        u64 *t = src;
        src = dst;
        dst = t;
        cur = pairs;
    }
    if (!rc && src != d_output &&
        hipMemcpyAsync(d_output, src, cur * 2 * sizeof(u64), hipMemcpyDeviceToDevice, s) != hipSuccess)
        rc = -1;
    hipFreeAsync(scratch, s);
    return rc ? rc : mam_sync_status();
}

// ---------------------------------------------------------------------------
// Opening path kernels (fused): oracle of round 0 and fold + next oracle
// ---------------------------------------------------------------------------

// Round 0: eq (Ext2) and bh (base or Ext2), natural order; oracle coefficients reduced in-launch.
template <bool BH_BASE>
__global__ __launch_bounds__(MAM_BLOCK) void k_open_oracle0(const u64 *__restrict__ eq, const u64 *__restrict__ bh,
                                                            size_t pairs, u64 *partials, u64 *result) {
    __shared__ u64 lds[6 * (MAM_BLOCK / WAVE)];
    u64 acc[6] = {0, 0, 0, 0, 0, 0};
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < pairs; i += (size_t)gridDim.x * blockDim.x) {
        ext2 e0 = e2_load(eq, 2 * i), ehi = e2_load(eq, 2 * i + 1);
        if (BH_BASE) {
            u64 b0 = gl_canon(bh[2 * i]), bhi = gl_canon(bh[2 * i + 1]);
            mam_oracle_accumulate_ends(acc, e0, ehi, e2_base(b0), e2_base(bhi));
        } else {
            ext2 b0 = e2_load(bh, 2 * i), bhi = e2_load(bh, 2 * i + 1);
            mam_oracle_accumulate_ends(acc, e0, ehi, b0, bhi);
        }
    }
    block_sum<6>(acc, lds);
    grid_partials<6>(acc, partials, result);
}

// Fold eq (Ext2) and bh (base or Ext2) at ch and reduce the next oracle (open_fold.h).
template <bool BH_BASE>
__global__ __launch_bounds__(MAM_BLOCK) void k_open_fold_oracle(OracleJob j) {
    __shared__ u64 lds[6 * (MAM_BLOCK / WAVE)];
    open_fold_oracle_run<BH_BASE, false>(j, blockIdx.x, gridDim.x, lds);
}

template <bool BH_BASE, bool FOLD>
__global__ __launch_bounds__(MAM_BLOCK) void k_open_u(UJob j) {
    open_u_run<BH_BASE, FOLD, false>(j, blockIdx.x, gridDim.x);
}

// Workgroups of the opening's streaming reductions (two-launch form, up to RED_BLOCKS
// partials): MAM_OPEN_GRID, default 512. Measured on the 2^22 proof: 512 -> 3.81 ms,
// 1024 -> 3.81 ms, 2048 -> 3.90 ms per proof; these launches are 10-30 us long and more
// workgroups only lengthen their ramp and the partial sums.
static unsigned open_grid(size_t items) {
    static const unsigned cap = [] {
        const char *e = getenv("MAM_OPEN_GRID");
        long x = e ? strtol(e, nullptr, 10) : 512L;
        if (x < 1) x = 1;
        if (x > RED_BLOCKS) x = RED_BLOCKS;
        return (unsigned)x;
    }();
    return grid_stride(items, MAM_BLOCK, cap);
}

int mam_open_u(const u64 *bh, bool bh_base, bool fold, ext2 ch, u64 *bh_out, size_t pairs, const u64 *r, int nb,
               const RedScratch &rs, hipStream_t s, const u64 *chp, bool rev) {
    if (nb < 0 || nb > 30 || pairs == 0 || pairs > ((size_t)1 << nb)) return -2;
    const unsigned g = open_grid(pairs);
    const UJob j = {bh, bh_out, ch, pairs, r, nb, rs.partials, rs.result, rs.counter, chp, rev ? 1 : 0};
    ProfScope prof(PROF_OPEN_SUMCHECK, (double)pairs * ((bh_base ? 8.0 : 16.0) * (fold ? 4 : 2) + (fold ? 32.0 : 0.0)),
                   s);
    if (fold) {
        if (bh_base) k_open_u<true, true><<<g, MAM_BLOCK, 0, s>>>(j);
        else k_open_u<false, true><<<g, MAM_BLOCK, 0, s>>>(j);
    } else {
        if (bh_base) k_open_u<true, false><<<g, MAM_BLOCK, 0, s>>>(j);
        else k_open_u<false, false><<<g, MAM_BLOCK, 0, s>>>(j);
    }
    launch_sum_partials<4>(rs.partials, g, rs.result, s);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mam_eq_table_scaled(const u64 *d_r, int nv, ext2 s0, ext2 s1, u64 *d_out, hipStream_t s) {
    if (nv < 1 || nv > 40) return -2;
    const ext2 f[2] = {s0, s1};
    return eqt::launch<FieldGL, true>(d_r, 0, nv, d_out, 0, 1, s, f);
}

int mam_open_oracle0(const u64 *eq, const u64 *bh, bool bh_base, size_t pairs, const RedScratch &rs,
                     hipStream_t s) {
    unsigned g = open_grid(pairs);
    if (bh_base)
        k_open_oracle0<true><<<g, MAM_BLOCK, 0, s>>>(eq, bh, pairs, rs.partials, rs.result);
    else
        k_open_oracle0<false><<<g, MAM_BLOCK, 0, s>>>(eq, bh, pairs, rs.partials, rs.result);
    launch_sum_partials<6>(rs.partials, g, rs.result, s);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mam_open_fold_oracle_job(const OracleJob &j, hipStream_t s) {
    const unsigned g = open_grid(j.half_out);
    ProfScope prof(PROF_OPEN_SUMCHECK, (double)j.half_out * (64.0 + (j.bh_base ? 32.0 : 64.0) + 128.0), s);
    if (j.bh_base)
        k_open_fold_oracle<true><<<g, MAM_BLOCK, 0, s>>>(j);
    else
        k_open_fold_oracle<false><<<g, MAM_BLOCK, 0, s>>>(j);
    launch_sum_partials<6>(j.partials, g, j.result, s);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mam_open_fold_oracle(const u64 *eq, const u64 *bh, bool bh_base, ext2 ch, u64 *eq_out, u64 *bh_out,
                         size_t half_out, const RedScratch &rs, hipStream_t s) {
    const OracleJob j = {eq, bh, ch, eq_out, bh_out, half_out, rs.partials, rs.counter, rs.result, 0, bh_base ? 1 : 0};
    return mam_open_fold_oracle_job(j, s);
}
