// Workgroup and grid reductions of field sums for 64-wide wavefronts (field policy F,
// Goldilocks unless named).
//
// block_sum: wave-level butterfly with 32-bit lane shuffles (DPP/ds_swizzle on CDNA),
// then one LDS slot per wave.
//
// Grid reductions come in two forms. Two launches (grid_partials + k_sum_partials): each
// workgroup writes its partial and ends, and a one-workgroup launch adds them. One launch
// (grid_finish): each workgroup publishes its partial, takes a ticket from a device-scope
// counter, and the last workgroup to arrive adds all partials and re-arms the counter. The
// ticket needs a device-scope release per workgroup, which on this GPU (one L2 per XCD)
// writes back the XCD's L2; measured on the sumcheck passes, the one-launch form cost
// 10-25 us per launch more than the second launch, so the standalone kernels use two
// launches and grid_finish serves only work that rides along another kernel's launch.
#pragma once

#include "common.h"
#include "fpolicy.h"
#include "sync.h"

static constexpr int WAVE = 64;

__device__ __forceinline__ u64 shfl_xor_u64(u64 v, int mask) {
    u32 lo = (u32)v, hi = (u32)(v >> 32);
    lo = __shfl_xor(lo, mask, WAVE);
    hi = __shfl_xor(hi, mask, WAVE);
    return ((u64)hi << 32) | lo;
}

// Sums M field elements over the workgroup. The result is valid in thread 0 only.
// `lds` must hold M * (blockDim.x / 64) words.
template <int M, class F = FieldGL>
__device__ __forceinline__ void block_sum(u64 (&v)[M], u64 *lds) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < M; k++) v[k] = F::add(v[k], shfl_xor_u64(v[k], off));
    }
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    const int nwaves = (blockDim.x + WAVE - 1) / WAVE;
    mam_barrier();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < M; k++) lds[wave * M + k] = v[k];
    }
    mam_barrier();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nwaves; w++) {
#pragma unroll
            for (int k = 0; k < M; k++) v[k] = F::add(v[k], lds[w * M + k]);
        }
    }
}

// The same for many words (M up to blockDim.x) through LDS instead of lane shuffles: every
// thread stores its M words (word k of all threads in row k, rows blockDim.x + 1 words apart
// against bank conflicts), blockDim.x / M threads per word add a segment of its row each, and
// M threads add the segments. A thread does ~M stores and ~2 M loads instead of 6 shuffle
// steps per word. The result is valid in thread 0 only; `lds` holds
// block_sum_lds_words(M, blockDim.x) words.
__host__ __device__ constexpr int block_sum_lds_words(int M, int B) { return M * (B + 1) + B; }
template <int M, class F = FieldGL>
__device__ __forceinline__ void block_sum_lds(u64 (&v)[M], u64 *lds) {
    const int t = threadIdx.x, B = blockDim.x, stride = B + 1;
    const int S = B / M;  // segments per word
    u64 *part = lds + M * stride;
#pragma unroll
    for (int k = 0; k < M; k++) lds[k * stride + t] = v[k];
    mam_barrier();
    if (t < S * M) {
        const int k = t % M, seg = t / M;
        const int lo = seg * B / S, hi = (seg + 1) * B / S;
        u64 acc = 0;
        for (int i = lo; i < hi; i++) acc = F::add(acc, lds[k * stride + i]);
        part[seg * M + k] = acc;
    }
    mam_barrier();
    if (t < M) {
        u64 acc = 0;
        for (int sg = 0; sg < S; sg++) acc = F::add(acc, part[sg * M + t]);
        lds[t] = acc;  // row 0 is no longer read
    }
    mam_barrier();
    if (t == 0) {
#pragma unroll
        for (int k = 0; k < M; k++) v[k] = lds[k];
    }
}

// Completion of a reduction over `nblocks` workgroups numbered `bid` = 0 .. nblocks - 1
// (grid_finish: every workgroup of the launch; grid_finish_n also serves a subset of a
// launch whose other workgroups do other work). Every thread of each of those workgroups
// calls it after block_sum (thread 0 holds the workgroup's sums). `partials` holds
// nblocks * M words. Returns true in exactly one workgroup (the last to finish), whose
// thread 0 then holds the total in v and has written it to `result`.
template <int M>
__device__ __forceinline__ bool grid_finish_n(u64 (&v)[M], u64 *lds, u64 *partials, u32 *counter, u64 *result,
                                              unsigned bid, unsigned nblocks) {
    if (nblocks == 1) {  // single workgroup: no ticket, no device-scope fences
        if (threadIdx.x == 0) {
#pragma unroll
            for (int k = 0; k < M; k++) result[k] = v[k];
        }
        return true;
    }
    __shared__ u32 s_last;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < M; k++) partials[(size_t)bid * M + k] = v[k];
        __threadfence();
        u32 ticket = atomicAdd(counter, 1u);
        s_last = (ticket == nblocks - 1) ? 1u : 0u;
    }
    mam_barrier();
    if (!s_last) return false;
    __threadfence();
#pragma unroll
    for (int k = 0; k < M; k++) v[k] = 0;
    for (unsigned b = threadIdx.x; b < nblocks; b += blockDim.x) {
#pragma unroll
        for (int k = 0; k < M; k++) v[k] = gl_add(v[k], __hip_atomic_load(&partials[(size_t)b * M + k], __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_AGENT));
    }
    block_sum<M>(v, lds);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < M; k++) result[k] = v[k];
        *counter = 0;
        __threadfence();
    }
    return true;
}

template <int M>
__device__ __forceinline__ bool grid_finish(u64 (&v)[M], u64 *lds, u64 *partials, u32 *counter, u64 *result) {
    return grid_finish_n<M>(v, lds, partials, counter, result, blockIdx.x, gridDim.x);
}

// Two-launch form, first half: a single workgroup writes the total to `result`, otherwise
// each workgroup writes its M words to partials[blockIdx.x * M ..] (thread 0 holds the
// workgroup's sums after block_sum). The caller launches k_sum_partials when gridDim.x > 1.
template <int M>
__device__ __forceinline__ void grid_partials(const u64 (&v)[M], u64 *partials, u64 *result) {
    if (threadIdx.x != 0) return;
    u64 *dst = gridDim.x == 1 ? result : partials + (size_t)blockIdx.x * M;
#pragma unroll
    for (int k = 0; k < M; k++) dst[k] = v[k];
}

// Second half: result[k] = sum of partials[b * M + k] over b < nblocks (one workgroup of
// blockDim.x >= M threads, L = blockDim.x / M of them per word, eight loads in flight each).
// With mj.box set, the M words are also posted to that host mailbox (mj.a is not used).
template <int M, class F = FieldGL>
__global__ __launch_bounds__(256) void k_sum_partials(const u64 *__restrict__ partials, unsigned nblocks,
                                                     u64 *__restrict__ result, MailJob mj) {
    static_assert(M <= 256, "one thread per word");
    __shared__ u64 fin[256];
    const int L = blockDim.x / M;
    u64 v = 0;
    if ((int)threadIdx.x < L * M) {
        const int k = threadIdx.x % M;
        unsigned b = threadIdx.x / M;
        for (; b + 7 * L < nblocks; b += 8 * L) {
            u64 t[8];
#pragma unroll
            for (int u = 0; u < 8; u++) t[u] = partials[(size_t)(b + u * L) * M + k];
#pragma unroll
            for (int u = 0; u < 8; u++) v = F::add(v, t[u]);
        }
        for (; b < nblocks; b += L) v = F::add(v, partials[(size_t)b * M + k]);
    }
    fin[threadIdx.x] = v;
    mam_barrier();
    if ((int)threadIdx.x < M) {
        u64 sum = 0;
        for (int j = 0; j < L; j++) sum = F::add(sum, fin[j * M + threadIdx.x]);
        result[threadIdx.x] = sum;
        if (mj.box) mj.box[threadIdx.x] = sum;
    }
    if (mj.box) mam_mailbox_raise(mj.box, mj.seq);
}

// Host side of the second half (nothing to add for a single workgroup). mj (may be null): a
// mailbox post of the M words made by that launch; returns whether it posted.
template <int M, class F = FieldGL>
static inline bool launch_sum_partials(const u64 *partials, unsigned nblocks, u64 *result, hipStream_t s,
                                       const MailJob *mj = nullptr) {
    if (nblocks <= 1) return false;
    const MailJob none = {};
    k_sum_partials<M, F><<<1, 256, 0, s>>>(partials, nblocks, result, mj ? *mj : none);
    return mj && mj->box;
}
