// VALU issue-rate microbenchmark (no reference counterpart): the measured peak against
// which the hash kernels' VALU rooflines are reported. Each thread runs 8 independent
// dependency chains of one instruction each per step, so the loop body is 8 VALU
// instructions plus the loop's own scalar bookkeeping:
//   kind 0: x = (x ^ c) + x          -> v_xad_u32 (32-bit ALU form)
//   kind 1: a = lo(a) * m + a        -> v_mad_u64_u32 (the hash kernels' multiply-add)
//   kind 2: a = a + b (64-bit)       -> v_lshl_add_u64 (the hash kernels' 64-bit add)
//   kind 3: x = x + c (32-bit)       -> v_add_u32 (two-operand ALU form)
// Wave-level VALU instructions per launch = blocks * (256 / 64) * iters * 8 (+ a few).
#include "common.h"

__global__ __launch_bounds__(256) void k_probe_xad(u32 *out, u32 c, int iters) {
    u32 x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = threadIdx.x * 0x9E3779B9u + i;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = (x[i] ^ c) + x[i];
    }
    u32 s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s ^= x[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_probe_mad64(u64 *out, u32 m, int iters) {
    u64 a[8];
#pragma unroll
    for (int i = 0; i < 8; i++) a[i] = threadIdx.x * 0x9E3779B97F4A7C15ull + i;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) a[i] = (u64)(u32)a[i] * m + a[i];
    }
    u64 s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s ^= a[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_probe_add64(u64 *out, u64 b, int iters) {
    u64 a[8];
#pragma unroll
    for (int i = 0; i < 8; i++) a[i] = threadIdx.x * 0x9E3779B97F4A7C15ull + i;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) a[i] = a[i] + b;
    }
    u64 s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s ^= a[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_probe_add32(u32 *out, u32 c, int iters) {
    u32 x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = threadIdx.x * 0x9E3779B9u + i;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 8; i++) x[i] = x[i] + c + (u32)it;
    }
    u32 s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) s ^= x[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// Runs the probe twice (the first launch warms up) and reports the second launch's time.
MAM_API int mam_valu_probe(int kind, int blocks, int iters, double *ms) {
    if (blocks <= 0 || iters <= 0 || !ms || kind < 0 || kind > 3) return -2;
    hipStream_t s = mam_stream();
    u64 *buf = nullptr;
    if (hipMalloc((void **)&buf, (size_t)blocks * 256 * sizeof(u64)) != hipSuccess) return -1;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) ? 0 : -1;
    for (int rep = 0; rep < 2 && !rc; rep++) {
        if (hipEventRecord(a, s) != hipSuccess) rc = -1;
        if (kind == 0) k_probe_xad<<<blocks, 256, 0, s>>>((u32 *)buf, 0x5bd1e995u, iters);
        else if (kind == 1) k_probe_mad64<<<blocks, 256, 0, s>>>(buf, 7u, iters);
        else if (kind == 2) k_probe_add64<<<blocks, 256, 0, s>>>(buf, 0x9E3779B97F4A7C15ull, iters);
        else k_probe_add32<<<blocks, 256, 0, s>>>((u32 *)buf, 0x5bd1e995u, iters);
        if (hipGetLastError() != hipSuccess || hipEventRecord(b, s) != hipSuccess ||
            hipEventSynchronize(b) != hipSuccess)
            rc = -1;
    }
    float t = 0;
    if (!rc && hipEventElapsedTime(&t, a, b) != hipSuccess) rc = -1;
    *ms = t;
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(buf);
    return rc;
}
