// eq tables eq(r, x) = prod_i (x_i ? r_i : 1 - r_i), index bit i <-> r_i (LSB first), for
// both fields (policy F: FieldGL or FieldQ) and both value types (EXT: Ext2 point and table,
// else base point and table). Semantics of cuda/eq_lagrange.cuh:57-158 and
// zk-torch-3/src/poly/mod.rs:136-163; one pass instead of one launch per variable.
//
// A workgroup owns 2^(la + lb) consecutive indices (la, lb <= 6; see split). Its first 64
// threads build A[j] = eq over the lowest la variables, the next 64 build B[j] = eq over the
// next lb variables times the workgroup's factor of all higher variables (the bits of
// blockIdx.x).
// Every output is then one product A[x mod 2^la] * B[x >> la]: about one multiplication and
// one 16-byte store per entry, so the launch runs at the store stream's rate.
#pragma once

#include "fpolicy.h"
#include "sync.h"

namespace eqt {

// Variables of the A table (la) and of the B table (lb); the lh = nv - la - lb above them
// number the workgroups. lb shrinks to keep at least 2^8 workgroups where nv allows it.
__host__ __device__ inline void split(int nv, int &la, int &lb) {
    la = nv < 6 ? nv : 6;
    const int rest = nv - la;
    lb = rest - 8 < 0 ? 0 : (rest - 8 > 6 ? 6 : rest - 8);
}

template <class F, bool EXT>
struct Ops;

template <class F>
struct Ops<F, true> {
    using V = typename F::E;
    static __device__ V one() { return F::e_one(); }
    static __device__ V load(const u64 *r, int i) { return F::e_load(r, i); }
    static __device__ V mul(V a, V b) { return F::e_mul(a, b); }
    static __device__ V one_minus(V a) { return F::e_sub(F::e_one(), a); }
    static __device__ void put(u64 *p, size_t i, V v) { F::e_store(p, i, v); }
    static __device__ V get(const u64 *p, size_t i) { return F::e(p[2 * i], p[2 * i + 1]); }
};

template <class F>
struct Ops<F, false> {
    using V = u64;
    static __device__ V one() { return 1; }
    static __device__ V load(const u64 *r, int i) { return F::canon(r[i]); }
    static __device__ V mul(V a, V b) { return F::mul(a, b); }
    static __device__ V one_minus(V a) { return F::sub(1, a); }
    static __device__ void put(u64 *p, size_t i, V v) { p[i] = v; }
    static __device__ V get(const u64 *p, size_t i) { return p[i]; }
};

// grid (2^(nv - la - lb), tables), 256 threads; table y reads r_all + y * r_stride and
// writes out_all + y * out_stride (words).
// With `scaled`, variable 0 contributes f0 (bit clear) / f1 (bit set) instead of 1 - r_0 / r_0:
// a scaled eq table, s * eq(r, x) with f0 = s (1 - r_0), f1 = s r_0.
template <class F, bool EXT>
__global__ __launch_bounds__(256) void k_eq(const u64 *__restrict__ r_all, size_t r_stride, int nv,
                                            u64 *__restrict__ out_all, size_t out_stride, int scaled,
                                            typename Ops<F, EXT>::V f0, typename Ops<F, EXT>::V f1) {
    using O = Ops<F, EXT>;
    using V = typename O::V;
    constexpr int W = EXT ? 2 : 1;
    __shared__ u64 ta[64 * W], tb[64 * W];
    const u64 *r = r_all + blockIdx.y * r_stride;
    u64 *out = out_all + blockIdx.y * out_stride;
    int la, lb;
    split(nv, la, lb);
    const int lh = nv - la - lb;
    const int t = threadIdx.x;
    if (t < (1 << la)) {
        V v = O::one();
        for (int i = 0; i < la; i++) {
            const V ri = O::load(r, i);
            const bool bit = (t >> i) & 1;
            v = O::mul(v, (i == 0 && scaled) ? (bit ? f1 : f0) : (bit ? ri : O::one_minus(ri)));
        }
        O::put(ta, t, v);
    } else if (t >= 64 && t < 64 + (1 << lb)) {
        const int j = t - 64;
        V v = O::one();
        for (int i = 0; i < lb; i++) {
            const V ri = O::load(r, la + i);
            v = O::mul(v, ((j >> i) & 1) ? ri : O::one_minus(ri));
        }
        const size_t hb = blockIdx.x;
        for (int i = 0; i < lh; i++) {
            const V ri = O::load(r, la + lb + i);
            v = O::mul(v, ((hb >> i) & 1) ? ri : O::one_minus(ri));
        }
        O::put(tb, j, v);
    }
    mam_barrier();
    const int per = 1 << (la + lb), amask = (1 << la) - 1;
    const size_t base = (size_t)blockIdx.x << (la + lb);
    for (int j = t; j < per; j += 256) O::put(out, base + j, O::mul(O::get(ta, j & amask), O::get(tb, j >> la)));
}

template <class F, bool EXT>
static int launch(const u64 *d_r, size_t r_stride, int nv, u64 *d_out, size_t out_stride, int tables,
                  hipStream_t s, const typename Ops<F, EXT>::V *scale01 = nullptr) {
    if (nv < 0 || nv > 40 || tables <= 0 || tables > 65535) return -1;
    if (scale01 && nv < 1) return -1;
    int la, lb;
    split(nv, la, lb);
    const dim3 grid((unsigned)((size_t)1 << (nv - la - lb)), (unsigned)tables);
    using V = typename Ops<F, EXT>::V;
    const V z{};
    k_eq<F, EXT><<<grid, 256, 0, s>>>(d_r, r_stride, nv, d_out, out_stride, scale01 ? 1 : 0,
                                      scale01 ? scale01[0] : z, scale01 ? scale01[1] : z);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace eqt
