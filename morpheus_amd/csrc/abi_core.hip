// Initialization, memory management, device selection and elementwise field batches.
//
// Entry points replace goldilocks-cuda-rs/src/ffi.rs:11-26, 30-53, 306-317, 159-296
// (the `gl_batch_*`, `ext2_batch_*` and conversion symbols) with the same names,
// arguments and return codes; implementations follow goldilocks-cuda-rs/cuda/wrapper.cu:27-97.
#include <stdlib.h>
#include <string.h>

#include "common.h"

// ---------------------------------------------------------------------------
// Initialization (wrapper.cu:27-45). Constants live in code, so the hash inits only
// check that a device is usable.
// ---------------------------------------------------------------------------
MAM_API int goldilocks_cuda_init() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return -1;
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return -1;
    return hipSetDevice(cur) == hipSuccess ? 0 : -1;
}

MAM_API int poseidon2_cuda_init() { return goldilocks_cuda_init(); }
MAM_API int monolith_cuda_init() { return goldilocks_cuda_init(); }

// ---------------------------------------------------------------------------
// Memory management (wrapper.cu:51-97); copies are ordered on the per-thread stream and
// complete before returning, as the reference's synchronous cudaMemcpy does.
// ---------------------------------------------------------------------------
// cuda_malloc / cuda_free and the rest of the memory entry points: mem_pool.hip.

static int copy_sync(void *dst, const void *src, size_t size, hipMemcpyKind kind) {
    if (size == 0) return 0;
    if (hipMemcpyAsync(dst, src, size, kind, mam_stream()) != hipSuccess) return -1;
    return hipStreamSynchronize(mam_stream()) == hipSuccess ? 0 : -1;
}

MAM_API int cuda_memcpy_htod(void *dst, const void *src, size_t size) {
    return copy_sync(dst, src, size, hipMemcpyHostToDevice);
}
MAM_API int cuda_memcpy_dtoh(void *dst, const void *src, size_t size) {
    return copy_sync(dst, src, size, hipMemcpyDeviceToHost);
}
MAM_API int cuda_memcpy_dtod(void *dst, const void *src, size_t size) {
    return copy_sync(dst, src, size, hipMemcpyDeviceToDevice);
}

MAM_API int cuda_device_synchronize() { return hipDeviceSynchronize() == hipSuccess ? 0 : -1; }
MAM_API int cuda_get_last_error() { return (int)hipGetLastError(); }
MAM_API int cuda_peek_at_last_error() { return (int)hipPeekAtLastError(); }
MAM_API int cuda_mem_get_info(size_t *free_b, size_t *total_b) {
    return hipMemGetInfo(free_b, total_b) == hipSuccess ? 0 : -1;
}

MAM_API int cuda_set_device(int device) { return hipSetDevice(device) == hipSuccess ? 0 : -1; }
MAM_API int cuda_get_device(int *device) { return hipGetDevice(device) == hipSuccess ? 0 : -1; }
MAM_API int cuda_get_device_count(int *count) {
    return hipGetDeviceCount(count) == hipSuccess ? 0 : -1;
}
MAM_API int cuda_get_device_name(int device, char *name, int max_len) {
    if (!name || max_len <= 0) return -1;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return -1;
    strncpy(name, prop.name, (size_t)max_len - 1);
    name[max_len - 1] = 0;
    return 0;
}

// ---------------------------------------------------------------------------
// Elementwise batches. Inputs may be non-canonical; outputs are canonical.
// ---------------------------------------------------------------------------
__global__ void k_fill_zero(u64 *__restrict__ d, size_t n) {
    size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) d[i] = 0;
}

MAM_API int gl_memset_zero(u64 *d_buf, size_t n_u64) {
    if (n_u64 == 0) return 0;
    k_fill_zero<<<grid_1d(n_u64), MAM_BLOCK, 0, mam_stream()>>>(d_buf, n_u64);
    return mam_launch_status();
}

MAM_API int gl_zero_buffer(u64 *d_buf, int n) { return n <= 0 ? 0 : gl_memset_zero(d_buf, (size_t)n); }

// almost-goldilocks-cuda-rs/src/ffi.rs:687: n zero words, complete on return; returns the
// runtime's error code (0 on success).
MAM_API int agl_zero_buffer_ffi(u64 *d_buf, size_t n) {
    if (n == 0) return 0;
    k_fill_zero<<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_buf, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(mam_stream());
    return (int)e;
}

enum BinOp { OP_ADD, OP_SUB, OP_MUL, OP_DIV };
enum UnOp { OP_NEG, OP_SQR, OP_DBL, OP_INV, OP_EXP, OP_SCALE, OP_FROB };

template <int OP>
__global__ void k_gl_bin(const u64 *__restrict__ a, const u64 *__restrict__ b, u64 *__restrict__ r, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 x = gl_canon(a[i]), y = gl_canon(b[i]);
    u64 z;
    if (OP == OP_ADD) z = gl_add(x, y);
    else if (OP == OP_SUB) z = gl_sub(x, y);
    else if (OP == OP_MUL) z = gl_mul(x, y);
    else z = gl_mul(x, gl_inv(y));
    r[i] = z;
}

template <int OP>
__global__ void k_gl_un(const u64 *__restrict__ a, u64 *__restrict__ r, u64 arg, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 x = gl_canon(a[i]);
    u64 z;
    if (OP == OP_NEG) z = gl_neg(x);
    else if (OP == OP_SQR) z = gl_sqr(x);
    else if (OP == OP_DBL) z = gl_double(x);
    else if (OP == OP_INV) z = gl_inv(x);
    else if (OP == OP_EXP) z = gl_pow(x, arg);
    else z = gl_mul(x, gl_canon(arg));
    r[i] = z;
}

template <int OP>
__global__ void k_e2_bin(const u64 *__restrict__ a, const u64 *__restrict__ b, u64 *__restrict__ r, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ext2 x = e2_load(a, i), y = e2_load(b, i);
    ext2 z;
    if (OP == OP_ADD) z = e2_add(x, y);
    else if (OP == OP_SUB) z = e2_sub(x, y);
    else z = e2_mul(x, y);
    e2_store(r, i, z);
}

template <int OP>
__global__ void k_e2_un(const u64 *__restrict__ a, u64 *__restrict__ r, u64 arg, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ext2 x = e2_load(a, i);
    ext2 z;
    if (OP == OP_NEG) z = e2_neg(x);
    else if (OP == OP_SQR) z = e2_sqr(x);
    else if (OP == OP_INV) z = e2_inv(x);
    else if (OP == OP_EXP) z = e2_pow(x, arg);
    else if (OP == OP_SCALE) z = e2_scale(gl_canon(arg), x);
    else z = e2_frobenius(x);
    e2_store(r, i, z);
}

__global__ void k_gl_to_e2(const u64 *__restrict__ a, u64 *__restrict__ r, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    e2_store(r, i, e2_base(gl_canon(a[i])));
}

__global__ void k_e2_to_gl(const u64 *__restrict__ a, u64 *__restrict__ r, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    r[i] = gl_canon(a[2 * (size_t)i]);
}

#define GL_BIN(name, op)                                                                   \
    MAM_API int name(const u64 *d_a, const u64 *d_b, u64 *d_result, int n) {               \
        if (n <= 0) return 0;                                                              \
        k_gl_bin<op><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_a, d_b, d_result, n);   \
        return mam_launch_status();                                                        \
    }
#define GL_UN(name, op)                                                                    \
    MAM_API int name(const u64 *d_a, u64 *d_result, int n) {                               \
        if (n <= 0) return 0;                                                              \
        k_gl_un<op><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_a, d_result, 0, n);      \
        return mam_launch_status();                                                        \
    }
#define E2_BIN(name, op)                                                                   \
    MAM_API int name(const u64 *d_a, const u64 *d_b, u64 *d_result, int n) {               \
        if (n <= 0) return 0;                                                              \
        k_e2_bin<op><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_a, d_b, d_result, n);   \
        return mam_launch_status();                                                        \
    }
#define E2_UN(name, op)                                                                    \
    MAM_API int name(const u64 *d_a, u64 *d_result, int n) {                               \
        if (n <= 0) return 0;                                                              \
        k_e2_un<op><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_a, d_result, 0, n);      \
        return mam_launch_status();                                                        \
    }

GL_BIN(gl_batch_add, OP_ADD)
GL_BIN(gl_batch_sub, OP_SUB)
GL_BIN(gl_batch_mul, OP_MUL)
GL_BIN(gl_batch_div, OP_DIV)
GL_UN(gl_batch_inverse, OP_INV)
GL_UN(gl_batch_neg, OP_NEG)
GL_UN(gl_batch_square, OP_SQR)
GL_UN(gl_batch_double, OP_DBL)

MAM_API int gl_batch_exp(const u64 *d_a, u64 exp, u64 *d_result, int n) {
    if (n <= 0) return 0;
    k_gl_un<OP_EXP><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_a, d_result, exp, n);
    return mam_launch_status();
}
MAM_API int gl_batch_mul_scalar(u64 scalar, const u64 *d_a, u64 *d_result, int n) {
    if (n <= 0) return 0;
    k_gl_un<OP_SCALE><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_a, d_result, scalar, n);
    return mam_launch_status();
}

E2_BIN(ext2_batch_add_ffi, OP_ADD)
E2_BIN(ext2_batch_sub_ffi, OP_SUB)
E2_BIN(ext2_batch_mul_ffi, OP_MUL)
E2_UN(ext2_batch_inverse_ffi, OP_INV)
E2_UN(ext2_batch_neg_ffi, OP_NEG)
E2_UN(ext2_batch_square_ffi, OP_SQR)
E2_UN(ext2_batch_frobenius_ffi, OP_FROB)
E2_UN(ext2_batch_conjugate_ffi, OP_FROB)

MAM_API int ext2_batch_exp_ffi(const u64 *d_a, u64 exp, u64 *d_result, int n) {
    if (n <= 0) return 0;
    k_e2_un<OP_EXP><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_a, d_result, exp, n);
    return mam_launch_status();
}
MAM_API int ext2_batch_mul_scalar_ffi(u64 scalar, const u64 *d_a, u64 *d_result, int n) {
    if (n <= 0) return 0;
    k_e2_un<OP_SCALE><<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_a, d_result, scalar, n);
    return mam_launch_status();
}
MAM_API int gl_to_ext2_batch_ffi(const u64 *d_input, u64 *d_output, int n) {
    if (n <= 0) return 0;
    k_gl_to_e2<<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_input, d_output, n);
    return mam_launch_status();
}
MAM_API int ext2_to_gl_batch_ffi(const u64 *d_input, u64 *d_output, int n) {
    if (n <= 0) return 0;
    k_e2_to_gl<<<grid_1d(n), MAM_BLOCK, 0, mam_stream()>>>(d_input, d_output, n);
    return mam_launch_status();
}

// ---------------------------------------------------------------------------
// Opt-in kernel-group timing (profile.h)
// ---------------------------------------------------------------------------
#include <mutex>
#include <vector>

#include "profile.h"
#include "sync.h"

namespace {
struct ProfRec {
    int cat;
    double work;
    hipEvent_t a, b;
};
std::mutex g_prof_mu;
std::vector<ProfRec> g_prof_recs;
unsigned g_prof_mask = 0;  // bit c enables category c
}  // namespace

thread_local bool t_prof_suspended = false;
bool mam_prof_enabled(int cat) { return !t_prof_suspended && cat >= 0 && cat < 32 && ((g_prof_mask >> cat) & 1u); }
void mam_prof_suspend(bool on) { t_prof_suspended = on; }

void mam_prof_push(int cat, double work, hipEvent_t a, hipEvent_t b) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    g_prof_recs.push_back({cat, work, a, b});
}

static void prof_clear() {
    for (auto &r : g_prof_recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    g_prof_recs.clear();
}

// Enable (1) or disable (0) timing; either call discards what was recorded.
MAM_API int mam_profile_enable(int on) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    prof_clear();
    g_prof_mask = on ? ~0u : 0u;
    return 0;
}

MAM_API int mam_profile_enable_mask(unsigned mask) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    prof_clear();
    g_prof_mask = mask;
    return 0;
}

// Totals per category (PROF_NCAT entries each): milliseconds, scopes, work units.
// Categories: 0 Merkle builds (work = permutations), 1 opening fold+oracle (bytes),
// 2 commit transform (bytes), 3 sumcheck fold+message (bytes), 4 single leaf-hashing
// launches of base-field codewords (permutations), 5 int8 Ajtai GEMM launches (bytes).
MAM_API int mam_profile_collect(double *ms, long *count, double *work) {
    std::lock_guard<std::mutex> g(g_prof_mu);
    for (int c = 0; c < PROF_NCAT; c++) {
        ms[c] = 0;
        count[c] = 0;
        work[c] = 0;
    }
    for (auto &r : g_prof_recs) {
        if (hipEventSynchronize(r.b) != hipSuccess) return -1;
        float t = 0;
        if (hipEventElapsedTime(&t, r.a, r.b) != hipSuccess) return -1;
        ms[r.cat] += t;
        count[r.cat] += 1;
        work[r.cat] += r.work;
    }
    prof_clear();
    return 0;
}

// ---------------------------------------------------------------------------
// Per-thread scratch arenas (common.h)
// ---------------------------------------------------------------------------
namespace {
struct Arena {
    void *p = nullptr;
    size_t bytes = 0;
};
struct ThreadArenas {
    Arena dev[4];
    Arena host;
    u64 *mailbox = nullptr;  // fine-grained pinned host words, MAM_MAILBOX_WORDS + 1 (flag last)
    u64 seq = 0;
    ~ThreadArenas() {
        for (auto &a : dev)
            if (a.p) (void)hipFree(a.p);
        if (host.p) (void)hipHostFree(host.p);
        if (mailbox) (void)hipHostFree(mailbox);
    }
};
thread_local ThreadArenas t_arenas;
}  // namespace

u64 *mam_dev_arena(size_t words, int slot) {
    if (slot < 0 || slot >= 4) return nullptr;
    Arena &a = t_arenas.dev[slot];
    size_t bytes = (words ? words : 1) * sizeof(u64);
    if (a.bytes < bytes) {
        if (a.p) {
            (void)hipStreamSynchronize(mam_stream());
            (void)hipFree(a.p);
            a.p = nullptr;
            a.bytes = 0;
        }
        if (hipMalloc(&a.p, bytes) != hipSuccess) return nullptr;
        a.bytes = bytes;
    }
    return (u64 *)a.p;
}

u64 *mam_host_arena(size_t words) {
    Arena &a = t_arenas.host;
    size_t bytes = (words ? words : 1) * sizeof(u64);
    if (a.bytes < bytes) {
        if (a.p) {
            (void)hipStreamSynchronize(mam_stream());
            (void)hipHostFree(a.p);
            a.p = nullptr;
            a.bytes = 0;
        }
        if (hipHostMalloc(&a.p, bytes, hipHostMallocDefault) != hipSuccess) return nullptr;
        a.bytes = bytes;
    }
    return (u64 *)a.p;
}

// Page-locked host memory for caller-owned outputs (no counterpart in the reference).
MAM_API int mam_host_alloc(size_t bytes, void **p) {
    if (!p) return -2;
    *p = nullptr;
    return hipHostMalloc(p, bytes ? bytes : 8, hipHostMallocDefault) == hipSuccess ? 0 : -1;
}
MAM_API int mam_host_free(void *p) {
    if (!p) return 0;
    return hipHostFree(p) == hipSuccess ? 0 : -1;
}

// Free the calling thread's scratch arenas (they are re-created on demand).
MAM_API int mam_release_scratch() {
    (void)hipStreamSynchronize(mam_stream());
    for (auto &a : t_arenas.dev) {
        if (a.p) (void)hipFree(a.p);
        a = Arena();
    }
    if (t_arenas.host.p) (void)hipHostFree(t_arenas.host.p);
    t_arenas.host = Arena();
    if (t_arenas.mailbox) (void)hipHostFree(t_arenas.mailbox);
    t_arenas.mailbox = nullptr;
    return 0;
}

// ---------------------------------------------------------------------------
// Host mailbox (common.h): a one-thread-per-word kernel copies device words into
// fine-grained pinned host memory and then sets a sequence flag with a system-scope
// release store; the host spins on the flag instead of a copy + stream synchronization.
// ---------------------------------------------------------------------------
__global__ void k_publish(const u64 *a, int na, const u64 *b, int nb, u64 *box, u64 seq, GateJob g) {
    const int t = threadIdx.x;
    if (t < na)
        box[t] = a[t];
    else if (t < na + nb)
        box[t] = b[t - na];
    __threadfence_system();
    mam_barrier();
    if (t == 0) __hip_atomic_store(&box[MAM_MAILBOX_WORDS], seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    mam_gate_wait(g);
}

__global__ void k_gate(GateJob g) { mam_gate_wait(g); }

// ---------------------------------------------------------------------------
// Gates for pre-queued rounds (common.h). Per thread: pinned words [flag | err | 6 unused |
// 2 slots of 8] and a device copy of each slot; slot seq & 1 serves release seq, so with at
// most two gates armed ahead of the release the host never rewrites a slot that a waiting
// gate may still read.
// ---------------------------------------------------------------------------
namespace {
struct Gate {
    u64 *pin = nullptr;  // [0]: released number; [1]: spin gave up; [8 + 8 k]: slot k
    u64 *dev = nullptr;  // 2 x 8 words
    u64 armed = 0;       // last number armed
    u64 limit = 0;       // spin limit in wall-clock ticks
    bool ok = false;
    ~Gate() {
        if (pin) (void)hipHostFree(pin);
        if (dev) (void)hipFree(dev);
    }
};
thread_local Gate t_gate;
thread_local int t_prequeue = -1;  // -1: not set on this thread (the environment decides)

bool gate_ready() {
    if (t_gate.ok) return true;
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess ||
        khz <= 0)
        return false;
    if (hipHostMalloc((void **)&t_gate.pin, 24 * sizeof(u64), hipHostMallocCoherent) != hipSuccess) {
        t_gate.pin = nullptr;
        return false;
    }
    if (hipMalloc((void **)&t_gate.dev, 16 * sizeof(u64)) != hipSuccess) {
        (void)hipHostFree(t_gate.pin);
        t_gate.pin = nullptr;
        t_gate.dev = nullptr;
        return false;
    }
    memset(t_gate.pin, 0, 24 * sizeof(u64));
    t_gate.armed = 0;
    t_gate.limit = (u64)(MAM_GATE_LIMIT_S * 1000.0 * khz);
    t_gate.ok = true;
    return true;
}
}  // namespace

bool mam_prequeue_enabled() {
    static const bool env_on = [] {
        const char *e = getenv("MAM_PREQUEUE");
        return !(e && e[0] == '0');
    }();
    if (t_prequeue >= 0) return t_prequeue == 1;
    return env_on;
}

void mam_prequeue_set(bool on) { t_prequeue = on ? 1 : 0; }

bool mam_gate_ready() { return gate_ready(); }

const u64 *mam_gate_arm(int nwords, GateJob *g, u64 *seq) {
    if (nwords < 1 || nwords > 8 || !g || !gate_ready()) return nullptr;
    const u64 n = ++t_gate.armed;
    const int k = (int)(n & 1);
    *g = GateJob{t_gate.pin, n, t_gate.pin + 8 + 8 * k, t_gate.dev + 8 * k, nwords, t_gate.pin + 1, t_gate.limit};
    *seq = n;
    return g->dst;
}

int mam_gate_launch(const GateJob &g, hipStream_t s) {
    k_gate<<<1, 64, 0, s>>>(g);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

const u64 *mam_gate_queue(hipStream_t s, int nwords, u64 *seq) {
    if (nwords < 1 || nwords > 8 || !gate_ready()) return nullptr;
    const u64 n = t_gate.armed + 1;
    const int k = (int)(n & 1);
    if (hipStreamWaitValue64(s, t_gate.pin, n, hipStreamWaitValueGte, ~(u64)0) != hipSuccess) return nullptr;
    t_gate.armed = n;
    if (hipMemcpyAsync(t_gate.dev + 8 * k, t_gate.pin + 8 + 8 * k, nwords * sizeof(u64), hipMemcpyHostToDevice, s) !=
        hipSuccess) {
        mam_gate_open_all();
        return nullptr;
    }
    *seq = n;
    return t_gate.dev + 8 * k;
}

void mam_gate_open(u64 seq, const u64 *vals, int nwords) {
    if (!t_gate.ok) return;
    u64 *slot = t_gate.pin + 8 + 8 * (seq & 1);
    for (int i = 0; i < nwords && i < 8; i++) slot[i] = vals[i];
    __atomic_store_n(&t_gate.pin[0], seq, __ATOMIC_RELEASE);
}

void mam_gate_open_all() {
    if (t_gate.ok) __atomic_store_n(&t_gate.pin[0], t_gate.armed, __ATOMIC_RELEASE);
}

bool mam_gate_take_failure() {
    if (!t_gate.ok) return false;
    return __atomic_exchange_n(&t_gate.pin[1], (u64)0, __ATOMIC_ACQ_REL) != 0;
}

u64 *mam_mailbox_arm(u64 *seq) {
    if (!t_arenas.mailbox) {
        if (hipHostMalloc((void **)&t_arenas.mailbox, (MAM_MAILBOX_WORDS + 1) * sizeof(u64),
                          hipHostMallocCoherent) != hipSuccess) {
            t_arenas.mailbox = nullptr;
            return nullptr;
        }
        t_arenas.mailbox[MAM_MAILBOX_WORDS] = 0;
        t_arenas.seq = 0;
    }
    *seq = ++t_arenas.seq;
    return t_arenas.mailbox;
}

int mam_mailbox_post(const u64 *a, int na, const u64 *b, int nb, u64 *box, u64 seq, hipStream_t s,
                     const GateJob *g) {
    if (na < 0 || nb < 0 || na + nb > MAM_MAILBOX_WORDS || !box) return -2;
    k_publish<<<1, 64, 0, s>>>(a, na, b, nb, box, seq, g ? *g : GateJob{});
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

int mam_mailbox_wait(const u64 *box, u64 seq, int n, u64 *out, hipStream_t s) {
    if (!box || n < 0 || n > MAM_MAILBOX_WORDS) return -2;
    // Spin on the flag; every few thousand polls ask the stream whether it failed or drained.
    for (unsigned it = 1;; it++) {
        if (__atomic_load_n(&box[MAM_MAILBOX_WORDS], __ATOMIC_ACQUIRE) == seq) break;
        if ((it & 4095) == 0) {
            const hipError_t e = hipStreamQuery(s);
            if (e == hipSuccess) {
                if (__atomic_load_n(&box[MAM_MAILBOX_WORDS], __ATOMIC_ACQUIRE) == seq) break;
                return -1;
            }
            if (e != hipErrorNotReady) return -1;
        }
        __builtin_ia32_pause();
    }
    for (int i = 0; i < n; i++) out[i] = box[i];
    return 0;
}

int mam_publish_wait(const u64 *a, int na, const u64 *b, int nb, u64 *out, hipStream_t s) {
    if (na < 0 || nb < 0 || na + nb > MAM_MAILBOX_WORDS) return -2;
    u64 seq = 0;
    u64 *box = mam_mailbox_arm(&seq);
    if (!box) return -1;
    if (mam_mailbox_post(a, na, b, nb, box, seq, s)) return -1;
    return mam_mailbox_wait(box, seq, na + nb, out, s);
}
