// VALU issue-rate microbenchmark (no reference counterpart): the measured peak against
// which the hash kernels' VALU rooflines are reported. Each thread runs 8 independent
// dependency chains of one instruction each per step, so the loop body is 8 VALU
// instructions plus the loop's own scalar bookkeeping:
//   kind 0: x = (x ^ c) + x          -> v_xad_u32 (32-bit ALU form)
//   kind 1: a = lo(a) * m + a        -> v_mad_u64_u32 (the hash kernels' multiply-add)
//   kind 2: a = a + b (64-bit)       -> v_lshl_add_u64 (the hash kernels' 64-bit add)
//   kind 3: x = x + c (32-bit)       -> v_add_u32 (two-operand ALU form)
// Wave-level VALU instructions per launch = blocks * (256 / 64) * iters * 8 (+ a few).
// HBM copy probe (mam_copy_probe): a plain 16-byte-per-lane device copy, the bandwidth a
// streaming kernel of the same size can reach (launch and ramp included).
#include <algorithm>
#include <vector>

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

// One 16-byte word per thread per step over a grid-stride loop.
__global__ __launch_bounds__(256) void k_probe_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n16) {
    for (size_t i = blockIdx.x * (size_t)256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256) dst[i] = src[i];
}

// Copies `bytes` (a multiple of 4096) device to device `reps` times per grid shape (one word
// per thread, and 2048 / 8192 workgroups over a grid-stride loop) after a warm-up launch, and
// reports the median launch time of the fastest shape.
MAM_API int mam_copy_probe(size_t bytes, int reps, double *ms) {
    if (!ms || reps <= 0 || bytes == 0 || bytes % 4096) return -2;
    hipStream_t s = mam_stream();
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return -1;
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return -1;
    }
    const size_t n16 = bytes / 16;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
              hipMemsetAsync(a, 0, bytes, s) == hipSuccess)
                 ? 0
                 : -1;
    const size_t one = (n16 + 255) / 256;
    const unsigned grids[3] = {(unsigned)std::min<size_t>(one, 1u << 30), 2048, 8192};
    double best = 1e30;
    for (int gi = 0; gi < 3 && !rc; gi++) {
        const unsigned g = (unsigned)std::min<size_t>(grids[gi], one);
        std::vector<float> t;
        for (int r = 0; r <= reps && !rc; r++) {
            if (hipEventRecord(e0, s) != hipSuccess) rc = -1;
            k_probe_copy<<<g, 256, 0, s>>>((const uint4 *)a, (uint4 *)b, n16);
            if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess)
                rc = -1;
            float x = 0;
            if (!rc && hipEventElapsedTime(&x, e0, e1) != hipSuccess) rc = -1;
            if (r) t.push_back(x);  // the first launch warms up
        }
        if (!rc) {
            std::sort(t.begin(), t.end());
            best = std::min(best, (double)t[t.size() / 2]);
        }
    }
    *ms = best;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}
