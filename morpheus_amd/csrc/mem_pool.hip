// Device-memory entry points of both ABIs and the caching allocator behind cuda_malloc /
// cuda_free (goldilocks-cuda-rs/src/ffi.rs:30-53; almost-goldilocks-cuda-rs/src/ffi.rs:61-100,
// 342-350, 687). Host code only.
//
// Contract kept from the reference crate (its callers create and drop device buffers at a
// high rate): cuda_free does not return memory to the driver; a freed block is never handed
// out while work queued before its free can still touch it; when the driver refuses an
// allocation the cache is released and the request retried; AGL_ALLOC_CACHE=0 selects plain
// hipMalloc / hipFree; agl_cached_bytes reports the cache and cuda_pool_trim releases it.
//
// Design. Requests map to quarter-octave size classes (512 B, then 5/8, 6/8, 7/8 and 8/8
// of each power of two: at most 25% slack, no search). Every device keeps one FIFO of freed
// blocks per class. A block owns one event for its whole life; cuda_free records it on the
// freeing thread's per-thread stream and queues the block at the back of its class.
// cuda_malloc looks in its own class and the next one up and takes the first block whose
// event has already completed (hipEventQuery, no wait); only when every cached block of the
// class is still in flight does it wait, on the oldest one. A miss allocates a fresh block of
// the class size.
#include <stdlib.h>

#include <atomic>
#include <deque>
#include <mutex>
#include <unordered_map>

#include "common.h"

namespace {

constexpr int MAX_DEV = 16;
constexpr int NCLASS = 4 * 56 + 1;  // class 0 = 512 B; then four per octave up to 2^64
constexpr int SCAN = 8;             // cached blocks probed per class before waiting

struct Block {
    void *p = nullptr;
    size_t bytes = 0;  // class size
    int dev = 0;
    int cls = 0;
    hipEvent_t done = nullptr;  // recorded at each free; created at the first one
};

struct Device {
    std::deque<Block *> bins[NCLASS];
    size_t cached = 0;
};

std::mutex g_mu;
Device g_devs[MAX_DEV];
std::unordered_map<void *, Block *> g_out;  // blocks handed out, by address

bool cache_enabled() {
    static const bool on = [] {
        const char *e = getenv("AGL_ALLOC_CACHE");
        return !(e && e[0] == '0');
    }();
    return on;
}

// Size class of an n-byte request. For n > 512 with 2^(e-1) < n <= 2^e the class sizes are
// k * 2^(e-3), k = 5..8.
int class_of(size_t n) {
    if (n <= 512) return 0;
    const int e = 64 - __builtin_clzll((unsigned long long)(n - 1));
    const size_t step = (size_t)1 << (e - 3);
    const int k = (int)((n + step - 1) / step);  // 5..8
    return 1 + (e - 10) * 4 + (k - 5);
}

size_t class_bytes(int cls) {
    if (cls == 0) return 512;
    const int e = 10 + (cls - 1) / 4, k = 5 + (cls - 1) % 4;
    return (size_t)k << (e - 3);
}

int current_device() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    return dev;
}

void destroy(Block *b) {
    if (b->done) {
        (void)hipEventSynchronize(b->done);
        (void)hipEventDestroy(b->done);
    }
    (void)hipFree(b->p);
    delete b;
}

// Takes every cached block of `dev` out of its bins (the caller holds g_mu) and returns them,
// so the caller can wait on and free them after unlocking.
std::deque<Block *> drain_locked(int dev) {
    std::deque<Block *> out;
    Device &d = g_devs[dev];
    for (auto &bin : d.bins) {
        for (Block *b : bin) out.push_back(b);
        bin.clear();
    }
    d.cached = 0;
    return out;
}

void release_cache(int dev) {
    std::deque<Block *> blocks;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        blocks = drain_locked(dev);
    }
    for (Block *b : blocks) destroy(b);
}

// A cached block of class `cls` on `dev` whose free has completed, or, when `may_wait`, the
// oldest one (the caller then waits on its event); nullptr when the class has none. The
// caller holds g_mu.
Block *take_locked(int dev, int cls, bool may_wait, bool *must_wait) {
    auto &bin = g_devs[dev].bins[cls];
    *must_wait = false;
    int probed = 0;
    for (auto it = bin.begin(); it != bin.end() && probed < SCAN; ++it, ++probed) {
        if (hipEventQuery((*it)->done) == hipSuccess) {
            Block *b = *it;
            bin.erase(it);
            g_devs[dev].cached -= b->bytes;
            return b;
        }
    }
    (void)hipGetLastError();  // hipErrorNotReady from the probes
    if (!may_wait || bin.empty()) return nullptr;
    Block *b = bin.front();
    bin.pop_front();
    g_devs[dev].cached -= b->bytes;
    *must_wait = true;
    return b;
}

}  // namespace

// The stream-ordered allocations some entry points make for their own scratch
// (hipMallocAsync on the calling thread's stream) stay cached in the device's default pool
// between calls instead of going back to the driver at each synchronization. Once per device.
void mam_stream_pool_setup() {
    static std::atomic<unsigned long long> done{0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return;
    const unsigned long long bit = 1ull << dev;
    if (done.load(std::memory_order_acquire) & bit) return;
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t thr = ~0ull;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
    }
    done.fetch_or(bit, std::memory_order_release);
}

MAM_API int cuda_malloc(void **ptr, size_t size) {
    if (!ptr) return -1;
    const int dev = current_device();
    if (!cache_enabled() || dev < 0 || dev >= MAX_DEV)
        return hipMalloc(ptr, size ? size : 1) == hipSuccess ? 0 : -1;
    const int cls = class_of(size);
    Block *b = nullptr;
    bool wait = false;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        b = take_locked(dev, cls, false, &wait);
        if (!b && cls + 1 < NCLASS) b = take_locked(dev, cls + 1, false, &wait);
        if (!b) b = take_locked(dev, cls, true, &wait);
        if (b) g_out[b->p] = b;
    }
    if (b) {
        if (wait && hipEventSynchronize(b->done) != hipSuccess) {
            std::lock_guard<std::mutex> lk(g_mu);
            g_out.erase(b->p);
            destroy(b);
            return -1;
        }
        *ptr = b->p;
        return 0;
    }
    b = new Block;
    b->bytes = class_bytes(cls);
    b->dev = dev;
    b->cls = cls;
    if (hipMalloc(&b->p, b->bytes) != hipSuccess) {
        (void)hipGetLastError();
        release_cache(dev);
        if (hipMalloc(&b->p, b->bytes) != hipSuccess) {
            (void)hipGetLastError();
            delete b;
            return -1;
        }
    }
    {
        std::lock_guard<std::mutex> lk(g_mu);
        g_out[b->p] = b;
    }
    *ptr = b->p;
    return 0;
}

MAM_API int cuda_free(void *ptr) {
    if (!ptr) return 0;
    if (!cache_enabled()) return hipFree(ptr) == hipSuccess ? 0 : -1;
    Block *b = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_out.find(ptr);
        if (it != g_out.end()) {
            b = it->second;
            g_out.erase(it);
        }
    }
    if (!b) return hipFree(ptr) == hipSuccess ? 0 : -1;  // not from cuda_malloc
    // The event and the per-thread stream belong to the block's device.
    const int cur = current_device();
    if (cur != b->dev) (void)hipSetDevice(b->dev);
    bool ok = b->done || hipEventCreateWithFlags(&b->done, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventRecord(b->done, mam_stream()) == hipSuccess;
    if (cur >= 0 && cur != b->dev) (void)hipSetDevice(cur);
    if (!ok) {
        destroy(b);
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    Device &d = g_devs[b->dev];
    d.bins[b->cls].push_back(b);
    d.cached += b->bytes;
    return 0;
}

// almost-goldilocks-cuda-rs/src/ffi.rs:61: bytes held in the current device's free lists.
MAM_API size_t agl_cached_bytes() {
    const int dev = current_device();
    if (dev < 0 || dev >= MAX_DEV) return 0;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_devs[dev].cached;
}

// ffi.rs:85: release the allocator's cache and trim the default stream-ordered pool.
MAM_API int cuda_pool_trim(size_t min_bytes_to_keep) {
    const int dev = current_device();
    if (dev < 0) return -1;
    if (dev < MAX_DEV) release_cache(dev);
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) return -1;
    return hipMemPoolTrimTo(pool, min_bytes_to_keep) == hipSuccess ? 0 : -1;
}

// ffi.rs:84: byte fill, complete on return.
MAM_API int cuda_memset(void *dst, int value, size_t size) {
    if (size == 0) return 0;
    if (hipMemsetAsync(dst, value, size, mam_stream()) != hipSuccess) return -1;
    return hipStreamSynchronize(mam_stream()) == hipSuccess ? 0 : -1;
}

// ffi.rs:87: wait for the calling thread's queued work (AGL_DEVICE_SYNC=1: the whole device).
MAM_API int cuda_sync_mine() {
    static const bool dev_sync = [] {
        const char *e = getenv("AGL_DEVICE_SYNC");
        return e && e[0] == '1';
    }();
    const hipError_t r = dev_sync ? hipDeviceSynchronize() : hipStreamSynchronize(mam_stream());
    return r == hipSuccess ? 0 : -1;
}

// ffi.rs:96-100: caller-owned streams (default priority).
MAM_API int cuda_stream_create(void **stream) {
    if (!stream) return -1;
    hipStream_t s;
    if (hipStreamCreate(&s) != hipSuccess) return -1;
    *stream = (void *)s;
    return 0;
}
MAM_API int cuda_stream_destroy(void *stream) {
    return hipStreamDestroy((hipStream_t)stream) == hipSuccess ? 0 : -1;
}
MAM_API int cuda_stream_synchronize(void *stream) {
    return hipStreamSynchronize((hipStream_t)stream) == hipSuccess ? 0 : -1;
}
MAM_API int cuda_memcpy_dtod_async(void *dst, const void *src, size_t size, void *stream) {
    if (size == 0) return 0;
    return hipMemcpyAsync(dst, src, size, hipMemcpyDeviceToDevice, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}
MAM_API int cuda_memcpy_htod_async(void *dst, const void *src, size_t size, void *stream) {
    if (size == 0) return 0;
    return hipMemcpyAsync(dst, src, size, hipMemcpyHostToDevice, (hipStream_t)stream) == hipSuccess ? 0 : -1;
}

// ffi.rs:342-350: copies between devices (staged by the runtime when there is no peer path)
// and a best-effort peer enable (already enabled counts as success).
MAM_API int cuda_memcpy_peer(void *dst, int dst_dev, const void *src, int src_dev, size_t size) {
    if (size == 0) return 0;
    return hipMemcpyPeer(dst, dst_dev, src, src_dev, size) == hipSuccess ? 0 : -1;
}
MAM_API int cuda_enable_peer_access(int peer) {
    const hipError_t r = hipDeviceEnablePeerAccess(peer, 0);
    if (r == hipSuccess) return 0;
    (void)hipGetLastError();
    return r == hipErrorPeerAccessAlreadyEnabled ? 0 : -1;
}

// ffi.rs:80: text of a runtime error code (the codes cuda_get_last_error returns).
MAM_API const char *agl_cuda_error_string(int code) { return hipGetErrorString((hipError_t)code); }

// ffi.rs:63-64: host <-> device copies of large buffers after the calling thread's queued
// work, complete on return (the runtime stages pageable host memory itself).
MAM_API int agl_d2h_staged(void *h_dst, const void *d_src, size_t bytes) {
    if (bytes == 0) return 0;
    if (hipStreamSynchronize(mam_stream()) != hipSuccess) return -1;
    return hipMemcpy(h_dst, d_src, bytes, hipMemcpyDeviceToHost) == hipSuccess ? 0 : -1;
}
MAM_API int agl_h2d_staged(void *d_dst, const void *h_src, size_t bytes) {
    if (bytes == 0) return 0;
    if (hipStreamSynchronize(mam_stream()) != hipSuccess) return -1;
    return hipMemcpy(d_dst, h_src, bytes, hipMemcpyHostToDevice) == hipSuccess ? 0 : -1;
}
