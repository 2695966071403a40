// Shared launch helpers for the HIP kernels of libmorpheus_amd.
//
// Streams: every entry point enqueues on the calling host thread's per-thread stream
// (hipStreamPerThread), the HIP counterpart of the reference's `--default-stream
// per-thread` build (goldilocks-cuda-rs/build.rs:125). Entry points that the reference
// documents as synchronizing (encode, interpolation, Merkle builds, anything returning
// a host value) synchronize that stream before returning.
//
// Return convention (goldilocks-cuda-rs/cuda/wrapper.cu:27-97): 0 on success, -1 on a
// runtime error; -2 for an argument the entry point rejects.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "field.h"
#include "sync.h"

#define MAM_API extern "C" __attribute__((visibility("default")))

static constexpr int MAM_BLOCK = 256;

static inline hipStream_t mam_stream() { return hipStreamPerThread; }

// Workgroups of an in-launch (ticketed) grid reduction: 2 per CU keeps HBM streams full
// while bounding the contention on the single ticket counter.
static constexpr unsigned MAM_RED_GRID = 512;

static inline unsigned grid_1d(size_t n, int block = MAM_BLOCK) {
    size_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    if (g > 0x7FFFFFFFull) g = 0x7FFFFFFFull;
    return (unsigned)g;
}

// Grid for grid-stride reduction kernels: enough workgroups to fill 256 CUs several
// times over, never more than the work needs.
static inline unsigned grid_stride(size_t n, int block = MAM_BLOCK, unsigned cap = 2048) {
    size_t g = (n + block - 1) / block;
    if (g == 0) g = 1;
    if (g > cap) g = cap;
    return (unsigned)g;
}

// Grid for the ticketed reduction kernels: one item per thread up to MAM_RED_GRID
// workgroups; work that fits one workgroup runs as a single workgroup, which skips the
// ticket and its device-scope fences. (A single workgroup looping over more items is
// latency-bound: 16 items per thread cost more than the ticket.)
static inline unsigned grid_reduce(size_t n, int block = MAM_BLOCK) {
    return grid_stride(n, block, MAM_RED_GRID);
}

// Keeps the device's default stream-ordered pool from returning freed hipMallocAsync blocks
// to the driver at each synchronization (mem_pool.hip; once per device).
void mam_stream_pool_setup();

static inline int mam_launch_status() { return hipGetLastError() == hipSuccess ? 0 : -1; }

static inline int mam_sync_status() {
    if (hipGetLastError() != hipSuccess) return -1;
    return hipStreamSynchronize(mam_stream()) == hipSuccess ? 0 : -1;
}

#define MAM_CHECK(x)                         \
    do {                                     \
        if ((x) != hipSuccess) return -1;    \
    } while (0)

// Per-thread reusable device and pinned-host scratch (abi_core.hip). `slot` selects one of
// a few independent arenas; a request larger than the arena's size reallocates it (the
// old contents are not kept). Returns nullptr on allocation failure. The memory stays
// owned by the library until the thread's arenas are released with mam_release_scratch.
u64 *mam_dev_arena(size_t words, int slot);
u64 *mam_host_arena(size_t words);

// Copies na device words from a, then nb from b, to the host array `out` through the
// calling thread's pinned mailbox, ordered after all work already queued on stream s.
// Costs one tiny kernel and a host spin on a flag instead of a copy and a stream
// synchronization. Returns 0, -1 (stream error) or -2 (na + nb > MAM_MAILBOX_WORDS).
static constexpr int MAM_MAILBOX_WORDS = 63;
int mam_publish_wait(const u64 *a, int na, const u64 *b, int nb, u64 *out, hipStream_t s);

// The same in three steps, for a kernel that posts its own result: arm reserves the next
// sequence value of the calling thread's mailbox (nullptr on failure); post queues the
// copy kernel; a kernel may instead write the words itself and call mam_mailbox_raise;
// wait spins until the value arrived and copies n words.
u64 *mam_mailbox_arm(u64 *seq);
struct GateJob;
// g (may be null): a gate the posting kernel waits for after the post.
int mam_mailbox_post(const u64 *a, int na, const u64 *b, int nb, u64 *box, u64 seq, hipStream_t s,
                     const GateJob *g = nullptr);
int mam_mailbox_wait(const u64 *box, u64 seq, int n, u64 *out, hipStream_t s);

// Device side of a post by one workgroup whose threads have written the mailbox words:
// make them visible to the host, then release the sequence flag.
__device__ __forceinline__ void mam_mailbox_raise(u64 *box, u64 seq) {
    __threadfence_system();
    mam_barrier();
    if (threadIdx.x == 0) __hip_atomic_store(&box[MAM_MAILBOX_WORDS], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Pre-queued rounds. A host-driven proving round needs values the host computes from the
// previous round's result (Fiat-Shamir challenges). Launching the round's kernels only after
// that costs an idle-queue launch (~10-15 us until the first kernel runs on the box); queuing
// them beforehand behind a gate costs only the host's turnaround. A gate releases up to 8
// scalars, which the host writes to a pinned slot, into device memory where the queued
// kernels read them. Two forms:
//  - mam_gate_queue: a stream wait on the pinned release flag (hipStreamWaitValue64) and an
//    asynchronous copy of the slot. The sumcheck's passes use it (measured faster there).
//  - mam_gate_arm + a spin: one wave polls the flag (bounded) and copies the slot, either as
//    a launch of its own (mam_gate_launch) or as the tail of the kernel that posts the
//    previous round's result (mam_mailbox_post and the lane-parallel tree's post take a
//    GateJob), which saves the gate's own dispatches. The Basefold opening uses it.
// mam_gate_open writes the scalars and releases. Gates are released in the order they were
// armed or queued, at most two outstanding. Every gate must be opened (mam_gate_open_all on
// error paths); a spin that has not seen its release after MAM_GATE_LIMIT_S seconds gives up
// (the stream then runs on with stale scalars) and records it: mam_gate_take_failure()
// reports and clears that.
// A stream waiting at a gate also holds up the other streams that share its hardware queue,
// so threads that prove side by side with others (the worker pools) do not pre-queue:
// mam_prequeue_enabled() is false on them (mam_prequeue_set), and MAM_PREQUEUE=0 turns it
// off everywhere.
static constexpr double MAM_GATE_LIMIT_S = 2.0;
struct GateJob {
    const u64 *flag;  // pinned: released number (nullptr: no gate)
    u64 seq;
    const u64 *slot;  // pinned scalars
    u64 *dst;         // device copy read by the gated kernels
    int n;
    u64 *err;    // pinned: set to 1 when the spin gave up
    u64 limit;   // spin limit in wall-clock ticks
};
bool mam_prequeue_enabled();
void mam_prequeue_set(bool on);
bool mam_gate_ready();  // the calling thread's gate memory is set up (false: cannot pre-queue)
const u64 *mam_gate_arm(int nwords, GateJob *g, u64 *seq);
int mam_gate_launch(const GateJob &g, hipStream_t s);
const u64 *mam_gate_queue(hipStream_t s, int nwords, u64 *seq);
void mam_gate_open(u64 seq, const u64 *vals, int nwords);
void mam_gate_open_all();
bool mam_gate_take_failure();

// Device side of a gate, run by the first wave of the workgroup (the other waves return):
// lane 0 spins on the release flag, then the lanes copy the n scalars (one host read for up
// to 64; a gate armed with mam_gate_arm may be pointed at a longer pinned source).
__device__ __forceinline__ void mam_gate_wait(const GateJob &g) {
    if (!g.flag || threadIdx.x >= 64) return;
    const int t = threadIdx.x;
    if (t == 0) {
        const u64 t0 = (u64)wall_clock64();
        while (__hip_atomic_load(g.flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) < g.seq) {
            if ((u64)wall_clock64() - t0 > g.limit) {
                __hip_atomic_store(g.err, (u64)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __atomic_thread_fence(__ATOMIC_ACQUIRE);  // the slot reads follow the flag
    for (int i = t; i < g.n; i += 64) g.dst[i] = __hip_atomic_load(&g.slot[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// A mailbox post handed to a kernel: na words from a, then the kernel's own words.
struct MailJob {
    const u64 *a;
    int na;
    u64 *box;  // nullptr: no post
    u64 seq;
    GateJob gate;  // the next round's gate, waited for after the post (gate.flag nullptr: none)
};
