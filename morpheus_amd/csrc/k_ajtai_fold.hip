// Folding and splitting of ring-Ajtai witnesses and commitments over R = F_q[X]/(X^64 + 1)
// (almost-goldilocks-cuda-rs/src/ffi.rs:435-493, 635-647; semantics of
// cuda_almost_goldilocks/ajtai.cuh:700-1275 and almost-goldilocks-cuda-rs/cuda/wrapper.cu:1468-1640,
// 1717-1740):
//   fold_witness:          out_j = z1_j + z2_j * r                 (binary z, F_q output)
//   fold_commitment:       out_i = c1_i + r * c2_i                  (R^KAPPA)
//   multifold_witness:     out_j = sum_m z_m,j * r_m               (i16 output)
//   multifold_mixed:       the same over K binary instances and T ternary ones (pos - neg)
//   multifold_commitment:  out_i = sum_m r_m * c_m,i
//   split_witness:         v = sum_{i<13} 2^i (pos_i - neg_i), the bits of |v| on v's sign side
// with challenges r of 64 int8 coefficients (the reference samples them from {-1, 0, 1, 2}). As in
// the reference kernels (ajtai.cuh:808-826, 1120-1140), the commitment folds take only the
// coefficients in [-2, 2] and skip the others; the witness folds take every int8 value.
//
// MI355X layout: one 64-lane wave per ring element, lane k owning coefficient k. A challenge is
// kept in LDS twice, T[d] = ~r[d] = -r[d] - 1 (in int8 range for every r, -128 included) and
// T[64 + d] = r[d], so the coefficient k of X^l r is T[64 + k - l] for l <= k and
// T[64 + k - l] + 1 for l > k: the sum over the set bits l of a witness word (wave-uniform) takes
// one LDS byte per bit and, once per word, the count of its set bits above k. The split reads
// one ring element per wave (coalesced) and forms each of its 26 output words with one ballot.
#include "common.h"
#include "fieldq.h"
#include "sync.h"

namespace {

constexpr int RD = 64;
constexpr int RKAPPA = 42;
constexpr int SPLIT = 13;
constexpr int FOLD_WAVES = 4;  // waves per workgroup

struct Challenge {
    int8_t c[RD];
};

bool fold_chunk_ok(int chunk) {
    return chunk == 64 || chunk == 128 || chunk == 256 || chunk == 1024 || chunk == 4096;
}

__device__ __forceinline__ u64 lift_small(int v) {
    return v >= 0 ? (u64)v : fq::Q - (u64)(-v);
}

// number of set bits of w above bit k (the wrapped rotations for output coefficient k)
__device__ __forceinline__ int above(u64 w, int k) {
    return __builtin_popcountll((w >> k) >> 1);
}

__device__ __forceinline__ u64 wave_uniform(u64 v) {
    const u32 lo = __builtin_amdgcn_readfirstlane((u32)v);
    const u32 hi = __builtin_amdgcn_readfirstlane((u32)(v >> 32));
    return ((u64)hi << 32) | lo;
}

// out[j * 64 + k] = z1 bit k + sum over set bits l of z2[j] of (X^l r)[k], canonical mod q
__global__ __launch_bounds__(FOLD_WAVES * 64) void k_fold_witness(const u64 *__restrict__ z1,
                                                                  const u64 *__restrict__ z2, Challenge r, u64 N,
                                                                  u64 *__restrict__ out) {
    __shared__ int8_t T[2 * RD];
    if (threadIdx.x < RD) {
        T[RD + threadIdx.x] = r.c[threadIdx.x];
        T[threadIdx.x] = (int8_t)~r.c[threadIdx.x];
    }
    mam_barrier();
    const int k = threadIdx.x & 63;
    const u64 stride = (u64)gridDim.x * FOLD_WAVES;
    for (u64 j = (u64)blockIdx.x * FOLD_WAVES + (threadIdx.x >> 6); j < N; j += stride) {
        int acc = (int)((z1[j] >> k) & 1);
        u64 mask = wave_uniform(z2[j]);
        acc += above(mask, k);
        while (mask) {
            const int l = __builtin_ctzll(mask);
            mask &= mask - 1;
            acc += T[RD + k - l];
        }
        out[j * RD + k] = lift_small(acc);
    }
}

// out[i * 64 + k] = c1 + sum_l r[l] (X^l c2)[k]: one workgroup of 64 per row i
__global__ __launch_bounds__(64) void k_fold_commitment(const u64 *__restrict__ c1, const u64 *__restrict__ c2,
                                                        Challenge r, u64 *__restrict__ out) {
    __shared__ u64 L[2 * RD];
    const int i = blockIdx.x, k = threadIdx.x;
    const u64 v = fq::canon(c2[i * RD + k]);
    L[RD + k] = v;
    L[k] = fq::neg(v);
    mam_barrier();
    u64 acc = fq::canon(c1[i * RD + k]);
    for (int l = 0; l < RD; l++) {
        const int rv = r.c[l];
        if (rv == 0 || rv > 2 || rv < -2) continue;  // as the reference: |r| <= 2 only
        const u64 x = L[RD + k - l];
        const u64 t = (rv == 2 || rv == -2) ? fq::add(x, x) : x;
        acc = rv > 0 ? fq::add(acc, t) : fq::sub(acc, t);
    }
    out[i * RD + k] = acc;
}

// Mixed multifold witness: K binary instances (bin[m * N + j]) and T ternary ones
// (pos / neg[t * N + j]) against challenges r_all[(K + T) * 64]; out[j * 64 + k] as i16.
// The challenges sit in LDS as doubled tables T_m[128].
__global__ __launch_bounds__(FOLD_WAVES * 64) void k_multifold_witness(const u64 *__restrict__ bin,
                                                                       const u64 *__restrict__ pos,
                                                                       const u64 *__restrict__ neg,
                                                                       const int8_t *__restrict__ r_all, int K,
                                                                       int Tn, u64 N, int16_t *__restrict__ out) {
    extern __shared__ int8_t Tm[];  // (K + Tn) * 128
    const int M = K + Tn;
    for (int e = threadIdx.x; e < M * RD; e += blockDim.x) {
        const int m = e >> 6, d = e & 63;
        const int8_t v = r_all[e];
        Tm[m * 2 * RD + RD + d] = v;
        Tm[m * 2 * RD + d] = (int8_t)~v;
    }
    mam_barrier();
    const int k = threadIdx.x & 63;
    const u64 stride = (u64)gridDim.x * FOLD_WAVES;
    for (u64 j = (u64)blockIdx.x * FOLD_WAVES + (threadIdx.x >> 6); j < N; j += stride) {
        int acc = 0;
        for (int m = 0; m < K; m++) {
            const int8_t *t = Tm + m * 2 * RD + RD + k;
            u64 mask = wave_uniform(bin[(u64)m * N + j]);
            acc += above(mask, k);
            while (mask) {
                const int l = __builtin_ctzll(mask);
                mask &= mask - 1;
                acc += t[-l];
            }
        }
        for (int m = 0; m < Tn; m++) {
            const int8_t *t = Tm + (K + m) * 2 * RD + RD + k;
            u64 mp = wave_uniform(pos[(u64)m * N + j]);
            acc += above(mp, k);
            while (mp) {
                const int l = __builtin_ctzll(mp);
                mp &= mp - 1;
                acc += t[-l];
            }
            u64 mn = wave_uniform(neg[(u64)m * N + j]);
            acc -= above(mn, k);
            while (mn) {
                const int l = __builtin_ctzll(mn);
                mn &= mn - 1;
                acc -= t[-l];
            }
        }
        out[j * RD + k] = (int16_t)acc;
    }
}

// out[i * 64 + k] = sum_m sum_l r_m[l] (X^l c_m,i)[k]: one workgroup of 64 per row i
__global__ __launch_bounds__(64) void k_multifold_commitment(const u64 *__restrict__ c_all,
                                                             const int8_t *__restrict__ r_all, int M,
                                                             u64 *__restrict__ out) {
    __shared__ u64 L[2 * RD];
    __shared__ int8_t rs[RD];
    const int i = blockIdx.x, k = threadIdx.x;
    u64 acc = 0;
    for (int m = 0; m < M; m++) {
        const u64 v = fq::canon(c_all[((u64)m * RKAPPA + i) * RD + k]);
        L[RD + k] = v;
        L[k] = fq::neg(v);
        rs[k] = r_all[m * RD + k];
        mam_barrier();
        for (int l = 0; l < RD; l++) {
            const int rv = rs[l];
            if (rv == 0 || rv > 2 || rv < -2) continue;
            const u64 x = L[RD + k - l];
            const u64 t = (rv == 2 || rv == -2) ? fq::add(x, x) : x;
            acc = rv > 0 ? fq::add(acc, t) : fq::sub(acc, t);
        }
        mam_barrier();
    }
    out[i * RD + k] = acc;
}

// One wave per ring element j: lane k reads coefficient k; pos[i * N + j] and neg[i * N + j]
// are ballots of (bit i of |v| set) on the positive and the negative side.
__global__ __launch_bounds__(FOLD_WAVES * 64) void k_split_witness(const int16_t *__restrict__ z, u64 N,
                                                                   u64 *__restrict__ pos, u64 *__restrict__ neg) {
    const int k = threadIdx.x & 63;
    const u64 stride = (u64)gridDim.x * FOLD_WAVES;
    for (u64 j = (u64)blockIdx.x * FOLD_WAVES + (threadIdx.x >> 6); j < N; j += stride) {
        const int v = z[j * RD + k];
        const int a = v < 0 ? -v : v;
        u64 p_mine = 0, n_mine = 0;
#pragma unroll
        for (int i = 0; i < SPLIT; i++) {
            const bool bit = (a >> i) & 1;
            const u64 pm = __ballot(bit && v > 0);
            const u64 nm = __ballot(bit && v < 0);
            if (k == i) {
                p_mine = pm;
                n_mine = nm;
            }
        }
        if (k < SPLIT) {
            pos[(u64)k * N + j] = p_mine;
            neg[(u64)k * N + j] = n_mine;
        }
    }
}

unsigned fold_grid(u64 n) {
    const u64 g = (n + FOLD_WAVES - 1) / FOLD_WAVES;
    return (unsigned)(g > 16384 ? 16384 : (g ? g : 1));
}

int multifold(const u64 *bin, const u64 *pos, const u64 *neg, const int8_t *r_all, int16_t *out, int K, int T,
              u64 N) {
    if (K < 0 || T < 0 || K + T < 1 || K + T > 511 || !r_all || !out || (K && !bin) || (T && (!pos || !neg)))
        return -1;
    if (N == 0) return 0;
    const size_t lds = (size_t)(K + T) * 2 * RD;
    k_multifold_witness<<<fold_grid(N), FOLD_WAVES * 64, lds, mam_stream()>>>(bin, pos, neg, r_all, K, T, N, out);
    return mam_launch_status();
}

}  // namespace

// ffi.rs:435-442 / wrapper.cu:1468-1490: out[N_ring * 64] canonical; r_coeffs is host memory
// (64 int8); chunk in {64, 128, 256, 1024, 4096} else -3 (it does not change the result).
MAM_API int ajtai_fold_witness_ffi(const u64 *d_z1, const u64 *d_z2, const int8_t *r_coeffs, u64 N_ring, int chunk,
                                   u64 *d_out) {
    if (!fold_chunk_ok(chunk)) return -3;
    if (!d_z1 || !d_z2 || !r_coeffs || !d_out) return -1;
    if (N_ring == 0) return 0;
    Challenge r;
    for (int k = 0; k < RD; k++) r.c[k] = r_coeffs[k];
    k_fold_witness<<<fold_grid(N_ring), FOLD_WAVES * 64, 0, mam_stream()>>>(d_z1, d_z2, r, N_ring, d_out);
    return mam_launch_status();
}

// ffi.rs:444-451 / wrapper.cu:1495-1508: M = num_instances binary witnesses [M][N_ring],
// challenges [M][64] on the device, out [N_ring][64] i16; chunk_size only shapes the reference's
// grid (0 is rejected).
MAM_API int ajtai_multifold_witness_ffi(const u64 *d_z_packed, const int8_t *d_r_all, int16_t *d_out,
                                        int num_instances, u64 N_ring, u64 chunk_size) {
    if (chunk_size == 0) return -1;
    return multifold(d_z_packed, nullptr, nullptr, d_r_all, d_out, num_instances, 0, N_ring);
}

// ffi.rs:453-463 / wrapper.cu:1512-1530: K binary instances and T ternary chunks (pos, neg),
// challenges [(K + T)][64] (binary first).
MAM_API int ajtai_multifold_mixed_witness_ffi(const u64 *d_z_bin_packed, const u64 *d_pos_packed,
                                              const u64 *d_neg_packed, const int8_t *d_r_all, int16_t *d_out,
                                              int num_binary, int num_ternary, u64 N_ring, u64 chunk_size) {
    if (chunk_size == 0) return -1;
    return multifold(d_z_bin_packed, d_pos_packed, d_neg_packed, d_r_all, d_out, num_binary, num_ternary, N_ring);
}

// ffi.rs:465-485 / wrapper.cu:1534-1615: the reference's tensor-core forms of the mixed multifold;
// same result, computed here by the same LDS-rotation kernel.
MAM_API int ajtai_multifold_mixed_witness_tc_ffi(const u64 *d_z_bin_packed, const u64 *d_pos_packed,
                                                 const u64 *d_neg_packed, const int8_t *d_r_all, int16_t *d_out,
                                                 int num_binary, int num_ternary, u64 N_ring) {
    return multifold(d_z_bin_packed, d_pos_packed, d_neg_packed, d_r_all, d_out, num_binary, num_ternary, N_ring);
}
MAM_API int ajtai_multifold_mixed_witness_tc_fused_ffi(const u64 *d_z_bin_packed, const u64 *d_pos_packed,
                                                       const u64 *d_neg_packed, const int8_t *d_r_all,
                                                       int16_t *d_out, int num_binary, int num_ternary, u64 N_ring) {
    return multifold(d_z_bin_packed, d_pos_packed, d_neg_packed, d_r_all, d_out, num_binary, num_ternary, N_ring);
}

// ffi.rs:487-492 / wrapper.cu:1619-1630: z_wide [N_ring][64] i16 -> pos, neg [13][N_ring]
MAM_API int ajtai_split_witness_ffi(const int16_t *d_z_wide, u64 *d_pos_chunks, u64 *d_neg_chunks, u64 N_ring) {
    if (!d_z_wide || !d_pos_chunks || !d_neg_chunks) return -1;
    if (N_ring == 0) return 0;
    k_split_witness<<<fold_grid(N_ring), FOLD_WAVES * 64, 0, mam_stream()>>>(d_z_wide, N_ring, d_pos_chunks,
                                                                             d_neg_chunks);
    return mam_launch_status();
}

// ffi.rs:635-640 / wrapper.cu:1717-1727: commitments [M][KAPPA][64], challenges [M][64] on the
// device, out [KAPPA][64]
MAM_API int ajtai_multifold_commitment_ffi(const u64 *d_c_packed, const int8_t *d_r_all, int num_instances,
                                           u64 *d_out) {
    if (!d_c_packed || !d_r_all || !d_out || num_instances < 1) return -1;
    k_multifold_commitment<<<RKAPPA, 64, 0, mam_stream()>>>(d_c_packed, d_r_all, num_instances, d_out);
    return mam_launch_status();
}

// ffi.rs:642-647 / wrapper.cu:1729-1739: r_coeffs is host memory (64 int8)
MAM_API int ajtai_fold_commitment_ffi(const u64 *d_c1, const u64 *d_c2, const int8_t *r_coeffs, u64 *d_out) {
    if (!d_c1 || !d_c2 || !r_coeffs || !d_out) return -1;
    Challenge r;
    for (int k = 0; k < RD; k++) r.c[k] = r_coeffs[k];
    k_fold_commitment<<<RKAPPA, 64, 0, mam_stream()>>>(d_c1, d_c2, r, d_out);
    return mam_launch_status();
}
