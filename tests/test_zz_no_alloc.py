"""The allocation-free int8 Ajtai commit (mam_ajtai_i8_commit_noalloc) against the allocating
entry point (mam_ajtai_i8_commit_prefix) through the C-ABI: two calls in a row, the second
with a smaller input, then two calls queued back to back on one stream, and the size limit
of its fixed scratch."""
import ctypes

import numpy as np
import pytest

import morpheus_amd as M
from morpheus_amd.ajtai import AjtaiI8, key_from_seed

pytestmark = pytest.mark.gpu
lib = M.lib
P = M.P


@pytest.fixture(scope="module")
def ajtai():
    assert lib.goldilocks_cuda_init() == 0
    a = AjtaiI8(key_from_seed(77), 256, 1 << 22)
    yield a
    a.free()


def reference(a, dz, batch, n_used):
    out = M.DeviceBuffer(batch * a.rows)
    assert lib.mam_ajtai_i8_commit_prefix(a.h, dz.ptr, batch, n_used, out.ptr) == 0
    return out.to_numpy()


def noalloc(a, dz, batch, n_used, out=None):
    out = out or M.DeviceBuffer(batch * a.rows)
    assert lib.mam_ajtai_i8_commit_noalloc(a.h, dz.ptr, batch, n_used, out.ptr) == 0
    return out


def test_noalloc_equals_allocating_commit(ajtai):
    rng = np.random.default_rng(2026)
    n = ajtai.n
    print("noalloc: batch 8 x 2^22, then batch 1 x 2^20 prefix", flush=True)
    z1 = M.DeviceBuffer.from_numpy(rng.integers(0, P, 8 * n, dtype=np.uint64))
    o1 = noalloc(ajtai, z1, 8, n)
    assert lib.cuda_sync_mine() == 0
    want1 = reference(ajtai, z1, 8, n)
    assert np.array_equal(o1.to_numpy(), want1)
    n2 = 1 << 20
    z2 = M.DeviceBuffer.from_numpy(rng.integers(0, P, n2, dtype=np.uint64))
    o2 = noalloc(ajtai, z2, 1, n2)
    assert lib.cuda_sync_mine() == 0
    assert np.array_equal(o2.to_numpy(), reference(ajtai, z2, 1, n2))
    # queued back to back on the calling thread's stream, one synchronize
    print("noalloc: two calls queued on one stream", flush=True)
    z3 = M.DeviceBuffer.from_numpy(rng.integers(0, P, 2 * n, dtype=np.uint64))
    o3 = noalloc(ajtai, z3, 2, n)
    o4 = noalloc(ajtai, z2, 1, n2)
    assert lib.cuda_sync_mine() == 0
    assert np.array_equal(o3.to_numpy(), reference(ajtai, z3, 2, n))
    assert np.array_equal(o4.to_numpy(), o2.to_numpy())
    # the first result is unchanged by the later calls (outputs are the caller's)
    assert np.array_equal(o1.to_numpy(), want1)


def test_noalloc_rejects_what_its_scratch_cannot_hold(ajtai):
    out = M.DeviceBuffer(16 * ajtai.rows)
    z = M.DeviceBuffer(16 * ajtai.n)
    # batch 16 over 2^22 columns needs 32 MiB of chunk partials
    assert lib.mam_ajtai_i8_commit_noalloc(ajtai.h, z.ptr, 16, ajtai.n, out.ptr) == -2
    # the allocating twin's argument checks hold too
    assert lib.mam_ajtai_i8_commit_noalloc(ajtai.h, z.ptr, 1, 100, out.ptr) == -2
    assert lib.mam_ajtai_i8_commit_noalloc(None, z.ptr, 1, 64, out.ptr) == -2
    with pytest.raises(ValueError):
        ajtai.commit_noalloc_device(z, 16, ajtai.n, out)
