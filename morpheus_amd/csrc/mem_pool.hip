// Device-memory entry points of both ABIs and the caching allocator behind cuda_malloc /
// cuda_free (goldilocks-cuda-rs/src/ffi.rs:30-53; almost-goldilocks-cuda-rs/src/ffi.rs:61-100,
// 342-350, 687). Host code only.
//
// Allocator (the behaviour of almost-goldilocks-cuda-rs/cuda/wrapper.cu:118-248, whose callers
// create and drop DeviceBuffers at a high rate): cuda_free puts a block on a per-device free
// list instead of returning it to the driver, and cuda_malloc hands out the smallest cached
// block that is at least the rounded request and at most 1.25 x + 2 MiB of it. Requests round
// up to 512-byte granules below 1 MiB and to 2 MiB pages above. cuda_free records an event on
// the freeing thread's per-thread stream and cuda_malloc waits for it before a reuse, so a
// block is never handed out while work queued before its free can still touch it. When the
// driver refuses an allocation, the device's cache is released and the request retried.
// Cached blocks count as used in cuda_mem_get_info; cuda_pool_trim releases them.
// AGL_ALLOC_CACHE=0 (the reference's switch) selects plain hipMalloc / hipFree.
#include <stdlib.h>

#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.h"

namespace {

constexpr int MAX_DEV = 16;

struct CachedBlock {
    void *p;
    size_t bytes;
    hipEvent_t freed;  // recorded on the freeing thread's stream
};

struct DevCache {
    std::multimap<size_t, CachedBlock> free_list;  // by size
    size_t cached = 0;
    std::vector<hipEvent_t> spare_events;
    bool pool_threshold_set = false;
};

struct LiveBlock {
    int dev;
    size_t bytes;
};

std::mutex g_mu;
DevCache g_dev[MAX_DEV];
std::unordered_map<void *, LiveBlock> g_live;

bool cache_enabled() {
    static const bool on = [] {
        const char *e = getenv("AGL_ALLOC_CACHE");
        return !(e && e[0] == '0');
    }();
    return on;
}

size_t round_request(size_t n) {
    if (n == 0) n = 1;
    const size_t small = (size_t)1 << 20, page = (size_t)2 << 20;
    if (n <= small) return (n + 511) & ~(size_t)511;
    return (n + page - 1) / page * page;
}

// Frees every cached block of `dev`; the caller holds g_mu and has `dev` current.
void flush_locked(int dev) {
    DevCache &c = g_dev[dev];
    for (auto &kv : c.free_list) {
        (void)hipEventSynchronize(kv.second.freed);
        (void)hipFree(kv.second.p);
        c.spare_events.push_back(kv.second.freed);
    }
    c.free_list.clear();
    c.cached = 0;
}

int current_device() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return -1;
    return dev;
}

}  // namespace

MAM_API int cuda_malloc(void **ptr, size_t size) {
    if (!ptr) return -1;
    const int dev = current_device();
    if (!cache_enabled() || dev < 0 || dev >= MAX_DEV)
        return hipMalloc(ptr, size ? size : 1) == hipSuccess ? 0 : -1;
    const size_t want = round_request(size);
    {
        std::unique_lock<std::mutex> lk(g_mu);
        DevCache &c = g_dev[dev];
        if (!c.pool_threshold_set) {
            // stream-ordered allocations inside the library keep their pool between calls
            c.pool_threshold_set = true;
            hipMemPool_t pool;
            if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
                uint64_t thr = ~0ull;
                (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
            }
        }
        auto it = c.free_list.lower_bound(want);
        if (it != c.free_list.end() && it->first <= want + want / 4 + ((size_t)2 << 20)) {
            CachedBlock b = it->second;
            c.free_list.erase(it);
            c.cached -= b.bytes;
            g_live[b.p] = LiveBlock{dev, b.bytes};
            lk.unlock();
            const bool done = hipEventSynchronize(b.freed) == hipSuccess;
            lk.lock();
            c.spare_events.push_back(b.freed);
            if (!done) {
                g_live.erase(b.p);
                (void)hipFree(b.p);
                return -1;
            }
            *ptr = b.p;
            return 0;
        }
    }
    void *p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> lk(g_mu);
        flush_locked(dev);
        if (hipMalloc(&p, want) != hipSuccess) {
            (void)hipGetLastError();
            return -1;
        }
    }
    std::lock_guard<std::mutex> lk(g_mu);
    g_live[p] = LiveBlock{dev, want};
    *ptr = p;
    return 0;
}

MAM_API int cuda_free(void *ptr) {
    if (!ptr) return 0;
    if (!cache_enabled()) return hipFree(ptr) == hipSuccess ? 0 : -1;
    LiveBlock info;
    hipEvent_t ev = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_live.find(ptr);
        if (it == g_live.end()) return hipFree(ptr) == hipSuccess ? 0 : -1;  // not from cuda_malloc
        info = it->second;
        g_live.erase(it);
        auto &spare = g_dev[info.dev].spare_events;
        if (!spare.empty()) {
            ev = spare.back();
            spare.pop_back();
        }
    }
    const int cur = current_device();
    if (cur != info.dev) (void)hipSetDevice(info.dev);
    bool ok = ev || hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipEventRecord(ev, mam_stream()) == hipSuccess;
    if (cur >= 0 && cur != info.dev) (void)hipSetDevice(cur);
    if (!ok) {
        if (ev) (void)hipEventDestroy(ev);
        return hipFree(ptr) == hipSuccess ? 0 : -1;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    DevCache &c = g_dev[info.dev];
    c.free_list.emplace(info.bytes, CachedBlock{ptr, info.bytes, ev});
    c.cached += info.bytes;
    return 0;
}

// almost-goldilocks-cuda-rs/src/ffi.rs:61: bytes held in the current device's free list.
MAM_API size_t agl_cached_bytes() {
    const int dev = current_device();
    if (dev < 0 || dev >= MAX_DEV) return 0;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_dev[dev].cached;
}

// ffi.rs:85: release the allocator's cache and trim the default stream-ordered pool.
MAM_API int cuda_pool_trim(size_t min_bytes_to_keep) {
    const int dev = current_device();
    if (dev < 0) return -1;
    if (dev < MAX_DEV) {
        std::lock_guard<std::mutex> lk(g_mu);
        flush_locked(dev);
    }
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
