// The zk-torch-3 Fiat-Shamir transcript (transcript.rs:167-330; host twin in
// host_transcript.cpp) held by one wavefront on the device, so a proof's challenges can be
// drawn without a host round trip.
//
// Layout in a wave: lane i holds sponge word i & 7 (every 16-lane row holds the state twice,
// as the lane-parallel permutation needs, and the rows compute identical copies), lanes 0..3 hold the input buffer (lane k: in[k]) and the output buffer
// (lane k: out[k]); the counts are wave-uniform. In memory the state is the host struct
// mam_transcript (18 words: state[8], in_buf[4], out_buf[4], in_count, out_count), so a
// transcript moves between host and device by copying it.
#pragma once

#include "field.h"
#include "poseidon2_wide.h"

namespace dtr {

struct Wave {
    u64 x;       // sponge word (lane & 7)
    u64 in;      // in_buf[lane] for lanes 0..3
    u64 out;     // out_buf[lane] for lanes 0..3
    int ni, no;  // wave-uniform counts
    int j;
    p2w::Lane L;

    __device__ void load(const u64 *st) {
        j = threadIdx.x & 7;
        p2w::lane_consts(j, L);
        x = st[j];
        in = j < 4 ? st[8 + j] : 0;
        out = j < 4 ? st[12 + j] : 0;
        // counts outside 0..3 / 0..4 cannot come from a transcript; clamp so reads stay in range
        const long long a = (long long)st[16], b = (long long)st[17];
        ni = (int)(a < 0 ? 0 : a > 3 ? 3 : a);
        no = (int)(b < 0 ? 0 : b > 4 ? 4 : b);
    }
    // every lane of the wave may call; lanes 0..7 write
    __device__ void store(u64 *st) const {
        const int lane = threadIdx.x;
        if (lane < 8) st[lane] = x;
        if (lane < 4) {
            st[8 + lane] = in;
            st[12 + lane] = out;
        }
        if (lane == 0) {
            st[16] = (u64)(long long)ni;
            st[17] = (u64)(long long)no;
        }
    }
    __device__ void duplex() {
        if (j < 4) x = j < ni ? in : 0;
        ni = 0;
        x = p2w::permute(x, j, L);
        // out[3 - k] = state[k]: lane k takes the word of lane 3 - k
        const u64 rev = p2w::rev4(x);
        if (j < 4) out = rev;
        no = 4;
    }
    __device__ void absorb(u64 v) {
        no = 0;
        v = gl_canon(v);
        if (j == ni) in = v;
        ni++;
        if (ni == 4) duplex();
    }
    __device__ u64 squeeze() {
        if (no == 0) duplex();
        no--;
        const u32 lo = (u32)__builtin_amdgcn_readlane((int)(u32)out, no);
        const u32 hi = (u32)__builtin_amdgcn_readlane((int)(u32)(out >> 32), no);
        return ((u64)hi << 32) | lo;
    }
    // len(label) then the label bytes; nothing for an empty label (transcript.rs:201-210)
    __device__ void label(const uint8_t *s, int len) {
        if (len <= 0) return;
        absorb((u64)len);
        for (int i = 0; i < len; i++) absorb(s[i]);
    }
};

}  // namespace dtr
