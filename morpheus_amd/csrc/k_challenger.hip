// Device Fiat-Shamir challengers (goldilocks-cuda-rs/src/ffi.rs:315-371, semantics of
// cuda/challenger.cuh:43-259): n independent Poseidon2-w8 duplex sponges (rate 4, capacity 4)
// in device memory, one per thread of each batch launch.
//
//   observe(v):  output buffer invalidated; v appended to the input buffer; a full input
//                buffer (4 values) duplexes.
//   duplex:      rate cells <- input buffer (missing cells zero), permute, output buffer
//                <- rate cells reversed so that pops return state[0], state[1], ...
//   sample():    an empty output buffer duplexes first; then pop.
//
// This is the duplex of zk-torch-3's Transcript without labels (transcript.rs:272-317):
// an observe is an append with an empty label, a sample a challenge with an empty label.
// Values are absorbed as residues (the sponge output depends only on residues), so the state
// copied to the host holds canonical words.
//
// State layout (136 bytes, challenger_state_size()): sponge[8], in[4], out[4] (u64),
// in_count, out_count (int32). The buffers are kept in registers inside a launch: every
// access to in[] / out[] with a runtime count is an unrolled select, so nothing spills.
#include "common.h"
#include "poseidon2.h"
#include "sync.h"

namespace {

struct ChState {
    u64 sponge[8];
    u64 in[4];
    u64 out[4];
    int in_count;
    int out_count;
};
static_assert(sizeof(ChState) == 136, "challenger state is 136 bytes");

struct ChReg {
    u64 s[8], in[4], out[4];
    int ni, no;

    __device__ void load(const ChState *p) {
#pragma unroll
        for (int k = 0; k < 8; k++) s[k] = p->sponge[k];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            in[k] = p->in[k];
            out[k] = p->out[k];
        }
        ni = p->in_count;
        no = p->out_count;
    }
    __device__ void store(ChState *p) const {
#pragma unroll
        for (int k = 0; k < 8; k++) p->sponge[k] = s[k];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            p->in[k] = in[k];
            p->out[k] = out[k];
        }
        p->in_count = ni;
        p->out_count = no;
    }
    __device__ void duplex() {
#pragma unroll
        for (int k = 0; k < 4; k++) s[k] = k < ni ? in[k] : 0;
        ni = 0;
        p2::permute(s);
#pragma unroll
        for (int k = 0; k < 4; k++) out[3 - k] = s[k];
        no = 4;
    }
    __device__ void observe(u64 v) {
        no = 0;
        v = gl_canon(v);
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k == ni) in[k] = v;
        ni++;
        if (ni == 4) duplex();
    }
    __device__ u64 sample() {
        if (no == 0) duplex();
        no--;
        u64 r = 0;
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (k == no) r = out[k];
        return r;
    }
};

__global__ void k_ch_init(ChState *st, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ChReg r;
#pragma unroll
    for (int k = 0; k < 8; k++) r.s[k] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) r.in[k] = r.out[k] = 0;
    r.ni = r.no = 0;
    r.store(st + i);
}

// Challenger i observes values[i * count .. i * count + count).
__global__ void k_ch_observe(ChState *st, const u64 *__restrict__ values, int count, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ChReg r;
    r.load(st + i);
    // a state written by the host may carry counts outside 0..4; clamp so indexing stays bounded
    r.ni = min(max(r.ni, 0), 3);
    r.no = min(max(r.no, 0), 4);
    for (int k = 0; k < count; k++) r.observe(values[(size_t)i * count + k]);
    r.store(st + i);
}

// Challenger i samples count values into out[i * count ..].
__global__ void k_ch_sample(ChState *st, u64 *__restrict__ out, int count, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ChReg r;
    r.load(st + i);
    r.ni = min(max(r.ni, 0), 3);
    r.no = min(max(r.no, 0), 4);
    for (int k = 0; k < count; k++) out[(size_t)i * count + k] = r.sample();
    r.store(st + i);
}

int ch_observe(void *d_states, const u64 *d_values, int count, int n) {
    if (n <= 0 || count <= 0) return 0;
    if (!d_states || !d_values) return -1;
    k_ch_observe<<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>((ChState *)d_states, d_values, count, n);
    return mam_launch_status();
}

int ch_sample(void *d_states, u64 *d_out, int count, int n) {
    if (n <= 0 || count <= 0) return 0;
    if (!d_states || !d_out) return -1;
    k_ch_sample<<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>((ChState *)d_states, d_out, count, n);
    return mam_launch_status();
}

}  // namespace

MAM_API int challenger_state_size() { return (int)sizeof(ChState); }

MAM_API int challenger_alloc_states(void **d_states, int n) {
    if (!d_states || n <= 0) return -1;
    return hipMalloc(d_states, (size_t)n * sizeof(ChState)) == hipSuccess ? 0 : -1;
}

MAM_API int challenger_init_states(void *d_states, int n) {
    if (n <= 0) return 0;
    if (!d_states) return -1;
    k_ch_init<<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>((ChState *)d_states, n);
    return mam_launch_status();
}

MAM_API int challenger_observe_ffi(void *d_states, const u64 *d_values, int n) {
    return ch_observe(d_states, d_values, 1, n);
}
MAM_API int challenger_observe_slice_ffi(void *d_states, const u64 *d_values, int count, int n) {
    return ch_observe(d_states, d_values, count, n);
}
MAM_API int challenger_observe_ext2_ffi(void *d_states, const u64 *d_values, int n) {
    return ch_observe(d_states, d_values, 2, n);
}
MAM_API int challenger_sample_ffi(void *d_states, u64 *d_outputs, int n) { return ch_sample(d_states, d_outputs, 1, n); }
MAM_API int challenger_sample_array_ffi(void *d_states, u64 *d_outputs, int count, int n) {
    return ch_sample(d_states, d_outputs, count, n);
}
MAM_API int challenger_sample_ext2_ffi(void *d_states, u64 *d_outputs, int n) {
    return ch_sample(d_states, d_outputs, 2, n);
}
// Ext5 challenges are five consecutive samples (cuda/challenger.cuh:230-236); no Ext5 arithmetic.
MAM_API int challenger_sample_ext5_ffi(void *d_states, u64 *d_outputs, int n) {
    return ch_sample(d_states, d_outputs, 5, n);
}

// One state (136 bytes), ordered after the calling thread's queued work.
MAM_API int challenger_copy_state_to_host(void *h_state, const void *d_state) {
    if (!h_state || !d_state) return -1;
    hipStream_t s = mam_stream();
    if (hipMemcpyAsync(h_state, d_state, sizeof(ChState), hipMemcpyDeviceToHost, s) != hipSuccess) return -1;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}
MAM_API int challenger_copy_state_to_device(void *d_state, const void *h_state) {
    if (!h_state || !d_state) return -1;
    hipStream_t s = mam_stream();
    if (hipMemcpyAsync(d_state, h_state, sizeof(ChState), hipMemcpyHostToDevice, s) != hipSuccess) return -1;
    return hipStreamSynchronize(s) == hipSuccess ? 0 : -1;
}

// ---- bit permutation of a base-field table (goldilocks-cuda-rs/src/ffi.rs:645-651) ----
// out[x] = in[y], bit b of x = bit inv_map[b] of y (the caller passes the inverse map, as
// goldilocks-cuda-rs/src/field.rs:205-214 does). The map is checked on the device side by
// masking: indices stay below 2^n_bits.
namespace {
__global__ void k_bit_permute(const u64 *__restrict__ in, u64 *__restrict__ out, const int *__restrict__ inv_map,
                              int n_bits, int total) {
    __shared__ int map[32];
    if (threadIdx.x < 32) map[threadIdx.x] = threadIdx.x < n_bits ? (inv_map[threadIdx.x] & 31) : 0;
    mam_barrier();
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= total) return;
    unsigned y = 0;
    for (int b = 0; b < n_bits; b++) y |= ((unsigned)(x >> b) & 1u) << map[b];
    out[x] = in[y & (unsigned)(total - 1)];
}
}  // namespace

MAM_API int bit_permute_gl_ffi(const u64 *d_input, u64 *d_output, const int *d_perm_map, int n_bits, int total) {
    if (total <= 0) return 0;
    if (n_bits < 0 || n_bits > 30 || total != (1 << n_bits) || !d_input || !d_output || !d_perm_map) return -1;
    k_bit_permute<<<grid_1d(total), MAM_BLOCK, 0, mam_stream()>>>(d_input, d_output, d_perm_map, n_bits, total);
    return mam_launch_status();
}
// The q crate's twin (almost-goldilocks-cuda-rs/src/ffi.rs:113-119): the gather moves words and
// does no field arithmetic, so it is the same launch.
MAM_API int agl_bit_permute_ffi(const u64 *d_input, u64 *d_output, const int *d_perm_map, int n_bits, int total) {
    return bit_permute_gl_ffi(d_input, d_output, d_perm_map, n_bits, total);
}
