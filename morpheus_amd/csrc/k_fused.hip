// Fused variable permutation + partial evaluation (goldilocks-cuda-rs/src/ffi.rs:657-668,
// cuda/fused_permute_peval.cuh:31-120; q twin almost-goldilocks-cuda-rs/src/ffi.rs:214-226):
//
//   output[j] = sum_{b < 2^m} evals[perm(j 2^m + b)] * eq[b],   j < output_size,
//   perm(x) = lo_lut[x & (2^half - 1)] | hi_lut[x >> half],
//
// evals base-field words, eq an Ext2 table of 2^m entries, output Ext2. The split LUTs map
// the bits of the permuted index back to the stored index (host side:
// goldilocks-cuda-rs/src/partial_eval.rs:266-321).
//
// MI355X layout: thread (j, s) sums the slice s of the b range for output j, with j fastest
// inside a workgroup, so for fixed b consecutive lanes read consecutive outputs' entries
// (coalesced whenever the permutation keeps the output variables in the low stored bits)
// and eq[b] is one broadcast load per wave. Both LUTs live in LDS (dynamic). With S > 1
// slices the Ext2 partials are summed by a second kernel.
#include "common.h"
#include "fpolicy.h"
#include "sync.h"

template <class F>
__global__ __launch_bounds__(256) void k_fused_peval(const u64 *__restrict__ evals, const u64 *__restrict__ eq,
                                                     const u32 *__restrict__ lo_lut, const u32 *__restrict__ hi_lut,
                                                     int n, int m, int half, u64 out_size, u64 slices,
                                                     u64 *__restrict__ out) {
    extern __shared__ u32 luts[];
    const u32 lo_n = 1u << half, hi_n = 1u << (n - half);
    for (u32 i = threadIdx.x; i < lo_n; i += blockDim.x) luts[i] = lo_lut[i];
    for (u32 i = threadIdx.x; i < hi_n; i += blockDim.x) luts[lo_n + i] = hi_lut[i];
    mam_barrier();
    const u64 t = blockIdx.x * (u64)blockDim.x + threadIdx.x;
    if (t >= out_size * slices) return;
    const u64 j = t % out_size, s = t / out_size;
    const u64 inner = (u64)1 << m;
    const u64 per = inner / slices;
    const u64 b0 = s * per, b1 = b0 + per;
    const u32 lo_mask = lo_n - 1;
    const u64 idx_mask = ((u64)1 << n) - 1;
    typename F::E acc = F::e_zero();
    for (u64 b = b0; b < b1; b++) {
        const u64 x = (j << m) + b;
        const u64 old = (luts[x & lo_mask] | luts[lo_n + (x >> half)]) & idx_mask;  // stays inside evals
        const u64 v = F::canon(evals[old]);
        acc = F::e_add(acc, F::e_scale(v, F::e_load(eq, b)));
    }
    F::e_store(out, t, acc);
}

template <class F>
__global__ void k_fused_sum(const u64 *__restrict__ part, u64 out_size, u64 slices, u64 *__restrict__ out) {
    const u64 j = blockIdx.x * (u64)blockDim.x + threadIdx.x;
    if (j >= out_size) return;
    typename F::E acc = F::e_zero();
    for (u64 s = 0; s < slices; s++) acc = F::e_add(acc, F::e_load(part, s * out_size + j));
    F::e_store(out, j, acc);
}

template <class F>
static int fused_peval(const u64 *d_evals, u64 *d_output, const u64 *d_eq, const u32 *d_lo, const u32 *d_hi, int n,
                       int m, int half, int output_size) {
    if (n < 0 || n > 32 || m < 0 || m > n || half < 0 || half > n || output_size <= 0) return -1;
    if ((u64)output_size << m > ((u64)1 << n)) return -1;
    const size_t lds = (((size_t)1 << half) + ((size_t)1 << (n - half))) * sizeof(u32);
    if (lds > 160 * 1024) return -1;
    hipStream_t s = mam_stream();
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute((const void *)k_fused_peval<F>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
            hipSuccess)
        return -1;
    // enough (output, slice) threads to fill the GPU; slices divide 2^m
    const u64 out_n = (u64)output_size, inner = (u64)1 << m;
    u64 slices = 1;
    while (slices * 2 <= inner && out_n * slices < ((u64)1 << 20) && inner / (slices * 2) >= 64) slices *= 2;
    const u64 threads = out_n * slices;
    u64 *part = d_output;
    if (slices > 1 && hipMallocAsync((void **)&part, threads * 2 * sizeof(u64), s) != hipSuccess) return -1;
    k_fused_peval<F><<<grid_1d(threads), 256, lds, s>>>(d_evals, d_eq, d_lo, d_hi, n, m, half, out_n, slices, part);
    int rc = hipGetLastError() == hipSuccess ? 0 : -1;
    if (slices > 1) {
        if (!rc) {
            k_fused_sum<F><<<grid_1d(out_n), 256, 0, s>>>(part, out_n, slices, d_output);
            if (hipGetLastError() != hipSuccess) rc = -1;
        }
        (void)hipFreeAsync(part, s);
    }
    return rc ? rc : mam_sync_status();
}

// smem_bytes is accepted for signature compatibility; the LDS size follows from n and half.
MAM_API int fused_permute_partial_eval_ffi(const u64 *d_evals, u64 *d_output, const u64 *d_eq_table,
                                           const u32 *d_lo_lut, const u32 *d_hi_lut, int n, int m, int half,
                                           int output_size, int smem_bytes) {
    (void)smem_bytes;
    return fused_peval<FieldGL>(d_evals, d_output, d_eq_table, d_lo_lut, d_hi_lut, n, m, half, output_size);
}

MAM_API int agl_fused_permute_partial_eval_ffi(const u64 *d_evals, u64 *d_output, const u64 *d_eq_table,
                                               const u32 *d_lo_lut, const u32 *d_hi_lut, int n, int m, int half,
                                               int output_size, int smem_bytes) {
    (void)smem_bytes;
    return fused_peval<FieldQ>(d_evals, d_output, d_eq_table, d_lo_lut, d_hi_lut, n, m, half, output_size);
}
