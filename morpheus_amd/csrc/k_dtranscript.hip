// Device Fiat-Shamir transcript with labels: a program of labelled appends and challenges
// run by one wavefront against a transcript state in device memory (dtranscript.h), queued
// on the calling thread's stream like any kernel, so the host never waits on it.
//
// Program (u64 words): per operation [kind, label_len, ceil(label_len / 8) words of label
// bytes (little-endian), n]; kind 0 appends the next n words of `src` (append_u64 /
// append_scalars / append_ext2 with n = 1 / k / 2), kind 1 squeezes n challenges into the
// next n words of `dst` (challenge_scalar / challenge_ext2 with n = 1 / 2). A program that
// runs past its own end, or past n_src / n_dst, stops there and reports it in the state's
// error word.
#include "common.h"
#include "dtranscript.h"

namespace {

constexpr int MAX_LABEL = 4096;

__global__ __launch_bounds__(64) void k_dtr_run(u64 *st, const u64 *__restrict__ prog, long nwords,
                                                const u64 *__restrict__ src, long n_src, u64 *__restrict__ dst,
                                                long n_dst, int *err) {
    dtr::Wave w;
    w.load(st);
    long pc = 0, si = 0, di = 0;
    int bad = 0;
    while (pc < nwords) {
        if (pc + 2 > nwords) {
            bad = 1;
            break;
        }
        const u64 kind = prog[pc], len = prog[pc + 1];
        const long lw = (long)((len + 7) / 8);
        if (kind > 1 || len > MAX_LABEL || pc + 2 + lw + 1 > nwords) {
            bad = 1;
            break;
        }
        const uint8_t *lb = (const uint8_t *)(prog + pc + 2);
        w.label(lb, (int)len);
        const long n = (long)prog[pc + 2 + lw];
        if (kind == 0) {
            if (n < 0 || si + n > n_src) {
                bad = 1;
                break;
            }
            for (long i = 0; i < n; i++) w.absorb(src[si + i]);
            si += n;
        } else {
            if (n < 0 || di + n > n_dst) {
                bad = 1;
                break;
            }
            for (long i = 0; i < n; i++) {
                const u64 v = w.squeeze();
                if (threadIdx.x == 0) dst[di + i] = v;
            }
            di += n;
        }
        pc += 3 + lw;
    }
    w.store(st);
    if (threadIdx.x == 0 && err) *err = bad;
}

}  // namespace

// d_state: a mam_transcript (18 words) in device memory; d_err (may be null): set to 1 when
// the program was malformed or overran src / dst, else 0. Queued, not synchronized.
MAM_API int mam_dtranscript_run(u64 *d_state, const u64 *d_prog, long prog_words, const u64 *d_src, long n_src,
                                u64 *d_dst, long n_dst, int *d_err) {
    if (!d_state || prog_words < 0 || (prog_words && !d_prog) || n_src < 0 || n_dst < 0 || (n_src && !d_src) ||
        (n_dst && !d_dst))
        return -2;
    k_dtr_run<<<1, 64, 0, mam_stream()>>>(d_state, d_prog, prog_words, d_src, n_src, d_dst, n_dst, d_err);
    return mam_launch_status();
}

// One Poseidon2-w8 permutation of each of n states (8 words each) with the lane-parallel
// form, one 16-lane row per state (4 per wavefront); for checking it against the scalar form.
namespace {
__global__ __launch_bounds__(256) void k_p2w_permute(u64 *states, long n) {
    const long g = ((long)blockIdx.x * blockDim.x + threadIdx.x) / p2w::GROUP;
    const int j = threadIdx.x & 7;
    const bool live = g < n;
    p2w::Lane L;
    p2w::lane_consts(j, L);
    const u64 x = live ? states[g * 8 + j] : 0;
    const u64 y = p2w::permute(x, j, L);
    if (live && (threadIdx.x & 8) == 0) states[g * 8 + j] = y;
}
}  // namespace

MAM_API int mam_poseidon2_wide_permute(u64 *d_states, long n) {
    if (n <= 0) return 0;
    if (!d_states) return -2;
    const long threads = n * p2w::GROUP;
    k_p2w_permute<<<(unsigned)((threads + 255) / 256), 256, 0, mam_stream()>>>(d_states, n);
    return mam_sync_status();
}
