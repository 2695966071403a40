"""The tensor-core commit probe (ajtai_tc_commit_probe_ffi) on INT8 MFMA against an exact
integer GEMM in numpy: full chunks, a partial last chunk (K not a multiple of 64) and an
empty chunk past K."""
import numpy as np
import pytest

import morpheus_amd as M

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.almost_goldilocks_cuda_init() == 0


@pytest.mark.parametrize("K,chunks,seed", [(2048, 2, 1), (4112, 6, 2), (1024 * 9, 9, 3)])
def test_tc_probe_matches_integer_gemm(K, chunks, seed):
    print("tc_probe", K, chunks, flush=True)
    rng = np.random.default_rng(seed)
    z = rng.integers(-128, 128, (16, K), dtype=np.int8)
    Mc = rng.integers(-128, 128, (960, K), dtype=np.int8)  # column-major M: column n is Mc[n]
    z[0, :64] = -128
    Mc[0, :64] = -128  # extreme products
    d_z = M.DeviceBuffer.from_numpy(np.frombuffer(z.tobytes(), dtype=np.uint64))
    d_m = M.DeviceBuffer.from_numpy(np.frombuffer(Mc.tobytes(), dtype=np.uint64))
    d_p = M.DeviceBuffer(chunks * 60 * 256 // 2)
    assert M.lib.ajtai_tc_commit_probe_ffi(d_z.ptr, d_m.ptr, d_p.ptr, K, chunks) == 0
    got = d_p.to_numpy().view(np.int32).reshape(chunks, 60, 16, 16)
    z64, m64 = z.astype(np.int64), Mc.astype(np.int64)
    for kb in range(chunks):
        lo, hi = min(1024 * kb, K), min(1024 * (kb + 1), K)
        want = z64[:, lo:hi] @ m64[:, lo:hi].T  # [16, 960]
        assert np.array_equal(got[kb].transpose(1, 0, 2).reshape(16, 960), want), kb
