"""The caching allocator behind cuda_malloc / cuda_free (both crates' memory entry points,
almost-goldilocks-cuda-rs/src/ffi.rs:61-100): size classes, reuse, the next class up, the
wait on a block whose free is still in flight on another thread's stream, and trimming."""
import ctypes
import threading

import numpy as np
import pytest

import morpheus_amd as M

pytestmark = pytest.mark.gpu
lib = M.lib
MiB = 1 << 20


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert lib.almost_goldilocks_cuda_init() == 0
    yield
    assert lib.cuda_pool_trim(0) == 0


def malloc(n):
    p = ctypes.c_void_p()
    assert lib.cuda_malloc(ctypes.byref(p), n) == 0
    return p.value


def test_size_classes_and_reuse():
    print("allocator: classes", flush=True)
    assert lib.cuda_pool_trim(0) == 0
    assert lib.agl_cached_bytes() == 0
    a = malloc(3 * MiB)                  # class 6/8 of 4 MiB = 3 MiB
    assert lib.cuda_free(a) == 0
    assert lib.agl_cached_bytes() == 3 * MiB
    b = malloc(3 * MiB - 300 * 1024)     # same class
    assert b == a and lib.agl_cached_bytes() == 0
    assert lib.cuda_free(b) == 0
    c = malloc(2 * MiB)                  # a class two below: a fresh block
    assert c != a and lib.agl_cached_bytes() == 3 * MiB
    d = malloc(2 * MiB + 100 * 1024)     # class 5/8 of 4 MiB is empty: takes the 3 MiB block
    assert d == a and lib.agl_cached_bytes() == 0
    e = malloc(700)                      # sub-KiB classes
    assert lib.cuda_free(e) == 0
    assert lib.agl_cached_bytes() == 768
    for p in (c, d):
        assert lib.cuda_free(p) == 0
    assert lib.agl_cached_bytes() == 768 + 2 * MiB + 3 * MiB
    assert lib.cuda_pool_trim(0) == 0
    assert lib.agl_cached_bytes() == 0


def test_data_survives_reuse_across_threads():
    """A block freed while kernels that write it are still queued on the freeing thread's
    stream is handed to another thread only after those kernels end: that thread's data is
    not overwritten afterwards."""
    print("allocator: cross-thread reuse", flush=True)
    assert lib.cuda_pool_trim(0) == 0
    n = 1 << 21                                   # Ext2 elements: 32 MiB
    rng = np.random.default_rng(7)
    src = M.DeviceBuffer.from_numpy(rng.integers(0, 0xFFFFFFFEFFFFFFE1, 2 * n, dtype=np.uint64))
    freed = {}

    def writer():
        x = malloc(16 * n)
        for _ in range(6):                        # queued, not waited for
            assert lib.aext2_batch_exp_ffi(src.ptr, (1 << 64) - 1, x, n) == 0
        assert lib.cuda_free(x) == 0
        freed["x"] = x

    t = threading.Thread(target=writer)
    t.start()
    t.join()
    y = malloc(16 * n)
    assert y == freed["x"]
    pattern = np.arange(2 * n, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    assert lib.cuda_memcpy_htod(y, pattern.ctypes.data, pattern.nbytes) == 0
    assert lib.cuda_device_synchronize() == 0
    back = np.empty_like(pattern)
    assert lib.cuda_memcpy_dtoh(back.ctypes.data, y, back.nbytes) == 0
    assert np.array_equal(back, pattern)
    assert lib.cuda_free(y) == 0
    src.free()
    assert lib.cuda_pool_trim(0) == 0


def test_many_buffers_round_trip():
    """Buffers of mixed sizes allocated and dropped at a high rate keep their contents."""
    print("allocator: churn", flush=True)
    rng = np.random.default_rng(8)
    live = []
    for it in range(200):
        n = int(rng.integers(1, 1 << 16))
        a = rng.integers(0, 1 << 63, n, dtype=np.uint64)
        live.append((M.DeviceBuffer.from_numpy(a), a))
        if len(live) > 12 or rng.random() < 0.3:
            buf, want = live.pop(int(rng.integers(0, len(live))))
            assert np.array_equal(buf.to_numpy(), want), it
            buf.free()
    for buf, want in live:
        assert np.array_equal(buf.to_numpy(), want)
        buf.free()
