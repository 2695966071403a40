// Ajtai commitment over R = F_q[X]/(X^64 + 1), q = 2^64 - 2^32 - 31, KAPPA = 42 rows:
//   c_i = sum_j M_{i, j + col_offset} * z_j  in R,
// with M_{i,j} generated on the fly from a 32-byte key by ChaCha8 and rejection sampling.
//
// Reference semantics: almost-goldilocks-cuda-rs/src/ajtai.rs:138-386 (commit_batched,
// commit_batched_at, commit, commit_wide(_device)), the FFI wrappers of
// almost-goldilocks-cuda-rs/cuda/wrapper.cu:1139-1205 and :1741-1760 (B in {1,2,4,8,16}
// else -2, chunk in {64,128,256,1024,4096} else -3), the matrix PRG
// cuda_almost_goldilocks/ajtai_chacha8.cuh:115-178 (coefficient 8 b + k of M_{i,j} is the
// k-th accepted u64 of ChaCha8(key, counter = b (+ 8 per retry), nonce = (i, j_lo, j_hi))),
// and the CPU reference ajtai_cpu_reference.cuh:58-242 (negacyclic products, sparse form).
//
// MI355X layout: a workgroup is one 64-lane wave; lane r owns output coefficient r of the
// 7 rows of its row group (6 groups cover the 42 rows). Per matrix column j the wave runs
// 56 ChaCha8 blocks in parallel (lane = row x block) and writes each row's 64 coefficients
// twice into LDS as L[t] = -M[t], L[64 + t] = M[t]; then (X^l M)[r] = L[64 + r - l], so a set
// bit l of a binary witness word adds one contiguous LDS row to every lane, with no sign
// logic. Sums are kept as separate 64-bit sums of the low and high 32-bit halves (one
// multiply-add each, no carries) and reduced mod q once per workgroup; a second kernel adds
// the per-chunk partials.
#include "chacha.h"
#include "common.h"
#include "fieldq.h"
#include "profile.h"
#include "ring_ajtai.h"
#include "sync.h"

namespace {

// Generate M_{i, col} for the 7 rows of group g into L[ii][0..128) (lanes 0..55 work).
__device__ __forceinline__ void ring_fill(const u32 key[8], int g, u64 col, u64 (*L)[2 * RING_D]) {
    const int lane = threadIdx.x;
    if (lane < RING_RG * 8) {
        const int ii = lane >> 3, blk = lane & 7;
        u64 c[8];
        ring_prg_block(key, (u32)(g * RING_RG + ii), col, (u32)blk, c);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            L[ii][RING_D + 8 * blk + k] = c[k];
            L[ii][8 * blk + k] = fq::neg(c[k]);
        }
    }
}

__device__ __forceinline__ u64 ring_finish(u64 lo_sum, u64 hi_sum) {
    // value = lo_sum + 2^32 hi_sum  (< 2^87)
    const u64 sh = hi_sum << 32;
    const u64 lo = lo_sum + sh;
    const u64 hi = (hi_sum >> 32) + (lo < sh);
    return fq::reduce128(lo, hi);
}

__device__ __forceinline__ u64 uniform_u64(u64 v) {
    const u32 lo = __builtin_amdgcn_readfirstlane((u32)v);
    const u32 hi = __builtin_amdgcn_readfirstlane((u32)(v >> 32));
    return ((u64)hi << 32) | lo;
}

// Dense binary witnesses: z[b * N + j] bit l = coefficient l of z_{b,j}.
// grid (chunks, RING_GROUPS, witness groups of BW); partial[((chunk * B + b) * KAPPA + i) * 64 + r].
// `one` is always 1: as a runtime value it keeps each half-word accumulation a single
// 32x32+64 multiply-add (v_mad_u64_u32) instead of a move plus a 64-bit add.
template <int BW>
__global__ __launch_bounds__(64) void k_ring_dense(const u32 *__restrict__ d_key, const u64 *__restrict__ z, u64 N,
                                                   int B, u64 col_offset, u64 cols, u32 one,
                                                   u64 *__restrict__ partial) {
    __shared__ u64 L[RING_RG][2 * RING_D];
    const int r = threadIdx.x, g = blockIdx.y;
    const int b0 = blockIdx.z * BW;
    u32 key[8];
#pragma unroll
    for (int k = 0; k < 8; k++) key[k] = d_key[k];
    u64 lo[RING_RG][BW], hi[RING_RG][BW];
#pragma unroll
    for (int ii = 0; ii < RING_RG; ii++)
#pragma unroll
        for (int w = 0; w < BW; w++) lo[ii][w] = hi[ii][w] = 0;
    const u64 j0 = (u64)blockIdx.x * cols;
    const u64 j1 = j0 + cols < N ? j0 + cols : N;
    for (u64 j = j0; j < j1; j++) {
        ring_fill(key, g, j + col_offset, L);
        mam_barrier();
#pragma unroll
        for (int w = 0; w < BW; w++) {
            if (b0 + w >= B) break;
            u64 mask = uniform_u64(z[(u64)(b0 + w) * N + j]);
            while (mask) {
                const int l = __builtin_ctzll(mask);
                mask &= mask - 1;
#pragma unroll
                for (int ii = 0; ii < RING_RG; ii++) {
                    const u64 v = L[ii][RING_D + r - l];
                    lo[ii][w] = (u64)(u32)v * one + lo[ii][w];
                    hi[ii][w] = (u64)(u32)(v >> 32) * one + hi[ii][w];
                }
            }
        }
        mam_barrier();
    }
#pragma unroll
    for (int w = 0; w < BW; w++) {
        if (b0 + w >= B) break;
#pragma unroll
        for (int ii = 0; ii < RING_RG; ii++)
            partial[(((u64)blockIdx.x * B + b0 + w) * RING_KAPPA + g * RING_RG + ii) * RING_D + r] =
                ring_finish(lo[ii][w], hi[ii][w]);
    }
}

// Wide witness: z[j * 64 + l] field coefficients (canonicalized on load).
// grid (chunks, RING_GROUPS); partial[(chunk * KAPPA + i) * 64 + r].
__global__ __launch_bounds__(64) void k_ring_wide(const u32 *__restrict__ d_key, const u64 *__restrict__ z, u64 N,
                                                  u64 col_offset, u64 cols, u64 *__restrict__ partial) {
    __shared__ u64 L[RING_RG][2 * RING_D];
    __shared__ u64 zs[RING_D];
    const int r = threadIdx.x, g = blockIdx.y;
    u32 key[8];
#pragma unroll
    for (int k = 0; k < 8; k++) key[k] = d_key[k];
    // 128-bit accumulator per row plus an overflow word: sum of up to 64 * cols products < 2^128 * 2^19
    u64 a0[RING_RG], a1[RING_RG], a2[RING_RG];
#pragma unroll
    for (int ii = 0; ii < RING_RG; ii++) a0[ii] = a1[ii] = a2[ii] = 0;
    const u64 j0 = (u64)blockIdx.x * cols;
    const u64 j1 = j0 + cols < N ? j0 + cols : N;
    for (u64 j = j0; j < j1; j++) {
        ring_fill(key, g, j + col_offset, L);
        zs[r] = fq::canon(z[j * RING_D + r]);
        mam_barrier();
        for (int l = 0; l < RING_D; l++) {
            const u64 zl = zs[l];
#pragma unroll
            for (int ii = 0; ii < RING_RG; ii++) {
                u64 plo, phi;
                mul_wide(L[ii][RING_D + r - l], zl, plo, phi);
                a0[ii] += plo;
                const u64 c0 = a0[ii] < plo;
                const u64 t = phi + c0;  // phi <= 2^64 - 2, no wrap
                a1[ii] += t;
                a2[ii] += a1[ii] < t;
            }
        }
        mam_barrier();
    }
#pragma unroll
    for (int ii = 0; ii < RING_RG; ii++) {
        // a0 + 2^64 a1 + 2^128 a2;  2^128 = C^2 mod q
        const u64 top = fq::mul(fq::canon(a2[ii]), fq::mul(fq::C, fq::C));
        const u64 v = fq::add(fq::reduce128(a0[ii], a1[ii]), top);
        partial[((u64)blockIdx.x * RING_KAPPA + g * RING_RG + ii) * RING_D + r] = v;
    }
}

// Sparse binary witness as positions p (column p >> 6, coefficient p & 63); duplicates add.
// grid (chunks, RING_GROUPS); partial[(chunk * KAPPA + i) * 64 + r].
__global__ __launch_bounds__(64) void k_ring_sparse(const u32 *__restrict__ d_key, const u64 *__restrict__ pos, u64 K,
                                                    u64 per, u32 one, u64 *__restrict__ partial) {
    __shared__ u64 L[RING_RG][2 * RING_D];
    const int r = threadIdx.x, g = blockIdx.y;
    u32 key[8];
#pragma unroll
    for (int k = 0; k < 8; k++) key[k] = d_key[k];
    u64 lo[RING_RG], hi[RING_RG];
#pragma unroll
    for (int ii = 0; ii < RING_RG; ii++) lo[ii] = hi[ii] = 0;
    const u64 k0 = (u64)blockIdx.x * per;
    const u64 k1 = k0 + per < K ? k0 + per : K;
    u64 cur = ~0ull;
    for (u64 k = k0; k < k1; k++) {
        const u64 p = uniform_u64(pos[k]);
        const u64 j = p >> 6;
        const int l = (int)(p & 63);
        if (j != cur) {  // consecutive positions in one column share its matrix entries
            mam_barrier();
            ring_fill(key, g, j, L);
            mam_barrier();
            cur = j;
        }
#pragma unroll
        for (int ii = 0; ii < RING_RG; ii++) {
            const u64 v = L[ii][RING_D + r - l];
            lo[ii] = (u64)(u32)v * one + lo[ii];
            hi[ii] = (u64)(u32)(v >> 32) * one + hi[ii];
        }
    }
#pragma unroll
    for (int ii = 0; ii < RING_RG; ii++)
        partial[((u64)blockIdx.x * RING_KAPPA + g * RING_RG + ii) * RING_D + r] = ring_finish(lo[ii], hi[ii]);
}

// Ternary witnesses of the split (SPLIT_K = 13 chunks): chunk b is z_b = pos_b - neg_b, bit l of
// pos[b * N + j] / neg[b * N + j] being coefficient l of column j. A pos bit adds
// (X^l M)[r] = L[64 + r - l], a neg bit its negation L[(r - l) & 127] (the doubled row holds
// -M then M). The matrix column comes from ChaCha8 (m == nullptr) or from a materialized
// M[j][row][64] (m != nullptr). grid (chunks, RING_GROUPS, chunk groups of BW);
// partial[((chunk * 13 + b) * KAPPA + i) * 64 + r].
static constexpr int RING_SPLIT_K = 13;

template <int BW>
__global__ __launch_bounds__(64) void k_ring_ternary(const u32 *__restrict__ d_key, const u64 *__restrict__ m,
                                                     const u64 *__restrict__ pos, const u64 *__restrict__ neg, u64 N,
                                                     u64 cols, u32 one, u64 *__restrict__ partial) {
    __shared__ u64 L[RING_RG][2 * RING_D];
    const int r = threadIdx.x, g = blockIdx.y;
    const int b0 = blockIdx.z * BW;
    u32 key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (!m) {
#pragma unroll
        for (int k = 0; k < 8; k++) key[k] = d_key[k];
    }
    u64 lo[RING_RG][BW], hi[RING_RG][BW];
#pragma unroll
    for (int ii = 0; ii < RING_RG; ii++)
#pragma unroll
        for (int w = 0; w < BW; w++) lo[ii][w] = hi[ii][w] = 0;
    const u64 j0 = (u64)blockIdx.x * cols;
    const u64 j1 = j0 + cols < N ? j0 + cols : N;
    for (u64 j = j0; j < j1; j++) {
        if (m) {
#pragma unroll
            for (int ii = 0; ii < RING_RG; ii++) {
                const u64 v = m[(j * RING_KAPPA + g * RING_RG + ii) * RING_D + r];
                L[ii][RING_D + r] = v;
                L[ii][r] = fq::neg(v);
            }
        } else {
            ring_fill(key, g, j, L);
        }
        mam_barrier();
#pragma unroll
        for (int w = 0; w < BW; w++) {
            if (b0 + w >= RING_SPLIT_K) break;
            u64 mp = uniform_u64(pos[(u64)(b0 + w) * N + j]);
            while (mp) {
                const int l = __builtin_ctzll(mp);
                mp &= mp - 1;
#pragma unroll
                for (int ii = 0; ii < RING_RG; ii++) {
                    const u64 v = L[ii][RING_D + r - l];
                    lo[ii][w] = (u64)(u32)v * one + lo[ii][w];
                    hi[ii][w] = (u64)(u32)(v >> 32) * one + hi[ii][w];
                }
            }
            u64 mn = uniform_u64(neg[(u64)(b0 + w) * N + j]);
            while (mn) {
                const int l = __builtin_ctzll(mn);
                mn &= mn - 1;
#pragma unroll
                for (int ii = 0; ii < RING_RG; ii++) {
                    const u64 v = L[ii][(r - l) & (2 * RING_D - 1)];
                    lo[ii][w] = (u64)(u32)v * one + lo[ii][w];
                    hi[ii][w] = (u64)(u32)(v >> 32) * one + hi[ii][w];
                }
            }
        }
        mam_barrier();
    }
#pragma unroll
    for (int w = 0; w < BW; w++) {
        if (b0 + w >= RING_SPLIT_K) break;
#pragma unroll
        for (int ii = 0; ii < RING_RG; ii++)
            partial[(((u64)blockIdx.x * RING_SPLIT_K + b0 + w) * RING_KAPPA + g * RING_RG + ii) * RING_D + r] =
                ring_finish(lo[ii][w], hi[ii][w]);
    }
}

// M[j][row][64] for columns j < N: one thread per (column, row, ChaCha8 block of 8 coefficients)
__global__ void k_ring_materialize(const u32 *__restrict__ d_key, u64 N, u64 *__restrict__ m) {
    const u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x;
    if (t >= N * RING_KAPPA * 8) return;
    const u64 j = t / (RING_KAPPA * 8);
    const u32 rem = (u32)(t % (RING_KAPPA * 8)), row = rem >> 3, blk = rem & 7;
    u32 key[8];
#pragma unroll
    for (int k = 0; k < 8; k++) key[k] = d_key[k];
    u64 c[8];
    ring_prg_block(key, row, j, blk, c);
    u64 *o = m + (j * RING_KAPPA + row) * RING_D + 8 * blk;
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = c[k];
}

// out[s * total + t] = sum of partial[c * total + t] over chunks c in slice s (mod q):
// grid (ceil(total / 256), slices); consecutive threads read consecutive words.
__global__ void k_ring_reduce(const u64 *__restrict__ partial, u64 chunks, u64 total, u64 per,
                              u64 *__restrict__ out) {
    const u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const u64 c0 = blockIdx.y * per;
    const u64 c1 = c0 + per < chunks ? c0 + per : chunks;
    u64 lo = 0, hi = 0;
    for (u64 c = c0; c < c1; c++) fq::acc_add(lo, hi, partial[c * total + t]);
    out[blockIdx.y * total + t] = fq::reduce128(lo, hi);
}

bool ring_chunk_ok(int chunk) {
    return chunk == 64 || chunk == 128 || chunk == 256 || chunk == 1024 || chunk == 4096;
}

// Columns per workgroup: at most the caller's chunk, fewer while the grid would hold fewer
// than 16384 single-wave workgroups (several rounds of the ~5 resident waves per SIMD on
// 1024 SIMDs, so the last round's tail is a small share).
u64 ring_cols(u64 n, int chunk, u64 groups) {
    const u64 target = 16384;
    u64 cols = (n * groups + target - 1) / target;
    if (cols < 1) cols = 1;
    if (cols > (u64)chunk) cols = (u64)chunk;
    return cols;
}

// Sum the chunk partials in two passes (slices of >= 32 chunks, then the slices) so that each
// thread adds at most a few dozen words; `partial` must hold room for the slice sums after
// the chunk partials (ring_partial_words).
u64 ring_slice_len(u64 chunks) {
    const u64 per = (chunks + 65534) / 65535;  // grid.y stays within 65535
    return per > 32 ? per : 32;
}
u64 ring_partial_words(u64 chunks, u64 total) {
    const u64 per = ring_slice_len(chunks);
    return (chunks + (chunks + per - 1) / per) * total;
}

int ring_reduce_to(u64 *partial, u64 chunks, u64 total, u64 *d_out, hipStream_t s) {
    const u64 per = ring_slice_len(chunks);
    const u64 slices = (chunks + per - 1) / per;
    u64 *mid = partial + chunks * total;
    const unsigned gx = grid_1d(total);
    if (slices > 1) {
        k_ring_reduce<<<dim3(gx, (unsigned)slices), MAM_BLOCK, 0, s>>>(partial, chunks, total, per, mid);
        k_ring_reduce<<<dim3(gx, 1), MAM_BLOCK, 0, s>>>(mid, slices, total, slices, d_out);
    } else {
        k_ring_reduce<<<dim3(gx, 1), MAM_BLOCK, 0, s>>>(partial, chunks, total, chunks, d_out);
    }
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace

// ffi.rs:500-508 / wrapper.cu:1150-1170
MAM_API int ajtai_commit_dense_batched_at_ffi(const u32 *d_key, const u64 *d_z, u64 N, int B, int chunk,
                                              u64 col_offset, u64 *d_out) {
    if (!ring_chunk_ok(chunk)) return -3;
    if (B != 1 && B != 2 && B != 4 && B != 8 && B != 16) return -2;
    if (N == 0 || !d_key || !d_z || !d_out) return -1;
    hipStream_t s = mam_stream();
    const int BW = B >= 4 ? 4 : B;
    const u64 wgroups = (u64)(B + BW - 1) / BW;
    const u64 cols = ring_cols(N, chunk, RING_GROUPS * wgroups);
    const u64 chunks = (N + cols - 1) / cols;
    const u64 total = (u64)B * RING_KAPPA * RING_D;
    if (chunks > 0x7FFFFFFFull) return -1;
    u64 *partial = nullptr;
    if (hipMallocAsync((void **)&partial, ring_partial_words(chunks, total) * sizeof(u64), s) != hipSuccess)
        return -1;
    const dim3 grid((unsigned)chunks, RING_GROUPS, (unsigned)wgroups);
    {
        ProfScope prof(PROF_AJTAI_RING, (double)N * B, s);
        if (BW == 1) k_ring_dense<1><<<grid, 64, 0, s>>>(d_key, d_z, N, B, col_offset, cols, 1u, partial);
        else if (BW == 2) k_ring_dense<2><<<grid, 64, 0, s>>>(d_key, d_z, N, B, col_offset, cols, 1u, partial);
        else k_ring_dense<4><<<grid, 64, 0, s>>>(d_key, d_z, N, B, col_offset, cols, 1u, partial);
    }
    int rc = hipGetLastError() == hipSuccess ? 0 : -1;
    if (!rc) rc = ring_reduce_to(partial, chunks, total, d_out, s);
    (void)hipFreeAsync(partial, s);
    return rc ? rc : mam_sync_status();
}

// ffi.rs:426-433
MAM_API int ajtai_commit_dense_batched_ffi(const u32 *d_key, const u64 *d_z, u64 N, int B, int chunk, u64 *d_out) {
    return ajtai_commit_dense_batched_at_ffi(d_key, d_z, N, B, chunk, 0, d_out);
}

// ffi.rs:595-602 / wrapper.cu:1188-1205: z_wide is [N][64] field coefficients
MAM_API int ajtai_commit_wide_ffi(const u32 *d_key, const u64 *d_z_wide, u64 N, u64 col_offset, int chunk,
                                  u64 *d_out) {
    if (!ring_chunk_ok(chunk)) return -3;
    if (N == 0 || !d_key || !d_z_wide || !d_out) return -1;
    hipStream_t s = mam_stream();
    const u64 cols = ring_cols(N, chunk, RING_GROUPS);
    const u64 chunks = (N + cols - 1) / cols;
    const u64 total = (u64)RING_KAPPA * RING_D;
    if (chunks > 0x7FFFFFFFull) return -1;
    u64 *partial = nullptr;
    if (hipMallocAsync((void **)&partial, ring_partial_words(chunks, total) * sizeof(u64), s) != hipSuccess)
        return -1;
    k_ring_wide<<<dim3((unsigned)chunks, RING_GROUPS), 64, 0, s>>>(d_key, d_z_wide, N, col_offset, cols, partial);
    int rc = hipGetLastError() == hipSuccess ? 0 : -1;
    if (!rc) rc = ring_reduce_to(partial, chunks, total, d_out, s);
    (void)hipFreeAsync(partial, s);
    return rc ? rc : mam_sync_status();
}

// ffi.rs:649-655 / wrapper.cu:1741-1760: K positions p = 64 j + l (one binary witness)
MAM_API int ajtai_commit_sparse_ffi(const u32 *d_key, const u64 *d_positions, u64 K, int chunk, u64 *d_out) {
    if (!ring_chunk_ok(chunk)) return -3;
    if (!d_key || !d_out || (K > 0 && !d_positions)) return -1;
    hipStream_t s = mam_stream();
    const u64 total = (u64)RING_KAPPA * RING_D;
    if (K == 0) {
        k_ring_reduce<<<dim3(grid_1d(total), 1), MAM_BLOCK, 0, s>>>(nullptr, 0, total, 1, d_out);
        return mam_sync_status();
    }
    const u64 per = ring_cols(K, chunk, RING_GROUPS);
    const u64 chunks = (K + per - 1) / per;
    if (chunks > 0x7FFFFFFFull) return -1;
    u64 *partial = nullptr;
    if (hipMallocAsync((void **)&partial, ring_partial_words(chunks, total) * sizeof(u64), s) != hipSuccess)
        return -1;
    k_ring_sparse<<<dim3((unsigned)chunks, RING_GROUPS), 64, 0, s>>>(d_key, d_positions, K, per, 1u, partial);
    int rc = hipGetLastError() == hipSuccess ? 0 : -1;
    if (!rc) rc = ring_reduce_to(partial, chunks, total, d_out, s);
    (void)hipFreeAsync(partial, s);
    return rc ? rc : mam_sync_status();
}

namespace {
// Shared body of the two ternary commits: out[b][i][r] for the 13 chunks.
int ring_ternary(const u32 *d_key, const u64 *d_m, const u64 *d_pos, const u64 *d_neg, u64 N, int chunk,
                 u64 *d_out) {
    if (!ring_chunk_ok(chunk)) return -3;
    if (N == 0 || (!d_key && !d_m) || !d_pos || !d_neg || !d_out) return -1;
    hipStream_t s = mam_stream();
    constexpr int BW = 4;
    const u64 wgroups = (RING_SPLIT_K + BW - 1) / BW;
    const u64 cols = ring_cols(N, chunk, RING_GROUPS * wgroups);
    const u64 chunks = (N + cols - 1) / cols;
    const u64 total = (u64)RING_SPLIT_K * RING_KAPPA * RING_D;
    if (chunks > 0x7FFFFFFFull) return -1;
    u64 *partial = nullptr;
    if (hipMallocAsync((void **)&partial, ring_partial_words(chunks, total) * sizeof(u64), s) != hipSuccess)
        return -1;
    k_ring_ternary<BW><<<dim3((unsigned)chunks, RING_GROUPS, (unsigned)wgroups), 64, 0, s>>>(d_key, d_m, d_pos, d_neg,
                                                                                           N, cols, 1u, partial);
    int rc = hipGetLastError() == hipSuccess ? 0 : -1;
    if (!rc) rc = ring_reduce_to(partial, chunks, total, d_out, s);
    (void)hipFreeAsync(partial, s);
    return rc;
}
}  // namespace

// ffi.rs:603-610 / wrapper.cu:1642-1661: 13 commitments [13][KAPPA][64] of the ternary chunks
// z_b = pos_b - neg_b (pos, neg: [13][N_ring] bit words); chunk as for the dense commit (else -3).
// Queued on the calling thread's stream.
MAM_API int ajtai_commit_ternary_ffi(const u32 *d_key, const u64 *d_pos, const u64 *d_neg, u64 N_ring, int chunk,
                                     u64 *d_out) {
    if (!d_key) return -1;
    return ring_ternary(d_key, nullptr, d_pos, d_neg, N_ring, chunk, d_out);
}

// ffi.rs:612-616 / wrapper.cu:1663-1671: M[j][row][64] for j < N (N * 42 * 64 words).
MAM_API int ajtai_materialize_m_ffi(const u32 *d_chacha_key, u64 *d_M, u64 N) {
    if (!d_chacha_key || !d_M) return -1;
    if (N == 0) return 0;
    const u64 threads = N * RING_KAPPA * 8;
    k_ring_materialize<<<(unsigned)((threads + MAM_BLOCK - 1) / MAM_BLOCK), MAM_BLOCK, 0, mam_stream()>>>(d_chacha_key,
                                                                                                  N, d_M);
    return mam_launch_status();
}

// ffi.rs:618-625 / wrapper.cu:1682-1700: the ternary commit against a materialized matrix.
MAM_API int ajtai_commit_ternary_premat_ffi(const u64 *d_M, const u64 *d_pos, const u64 *d_neg, u64 N_ring, int chunk,
                                            u64 *d_out) {
    if (!d_M) return -1;
    return ring_ternary(nullptr, d_M, d_pos, d_neg, N_ring, chunk, d_out);
}

// Host copy of matrix entry M_{row, col} (64 coefficients), for checks.
MAM_API int mam_ajtai_ring_matrix_entry(const uint8_t *key32, uint32_t row, uint64_t col, uint64_t *out64) {
    if (!key32 || !out64) return -2;
    u32 key[8];
    for (int k = 0; k < 8; k++)
        key[k] = (u32)key32[4 * k] | ((u32)key32[4 * k + 1] << 8) | ((u32)key32[4 * k + 2] << 16) |
                 ((u32)key32[4 * k + 3] << 24);
    for (u32 b = 0; b < 8; b++) ring_prg_block(key, row, col, b, out64 + 8 * b);
    return 0;
}
