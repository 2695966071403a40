// Workgroup and grid reductions of Goldilocks sums for 64-wide wavefronts.
//
// block_sum: wave-level butterfly with 32-bit lane shuffles (DPP/ds_swizzle on CDNA),
// then one LDS slot per wave. grid_finish: each workgroup publishes its partial, takes
// a ticket from a device-scope counter, and the last workgroup to arrive adds all
// partials and writes the final value, then re-arms the counter at zero so the next
// launch (or a graph replay) needs no memset.
#pragma once

#include "field.h"

static constexpr int WAVE = 64;

__device__ __forceinline__ u64 shfl_xor_u64(u64 v, int mask) {
    u32 lo = (u32)v, hi = (u32)(v >> 32);
    lo = __shfl_xor(lo, mask, WAVE);
    hi = __shfl_xor(hi, mask, WAVE);
    return ((u64)hi << 32) | lo;
}

// Sums M field elements over the workgroup. The result is valid in thread 0 only.
// `lds` must hold M * (blockDim.x / 64) words.
template <int M>
__device__ __forceinline__ void block_sum(u64 (&v)[M], u64 *lds) {
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int k = 0; k < M; k++) v[k] = gl_add(v[k], shfl_xor_u64(v[k], off));
    }
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    const int nwaves = (blockDim.x + WAVE - 1) / WAVE;
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < M; k++) lds[wave * M + k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nwaves; w++) {
#pragma unroll
            for (int k = 0; k < M; k++) v[k] = gl_add(v[k], lds[w * M + k]);
        }
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
    __syncthreads();
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
