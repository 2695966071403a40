"""GPU parity of the int8-matrix Ajtai commitment (MFMA) against the oracle's exact matvec."""
import numpy as np
import pytest

import pyoracle as O

import morpheus_amd as M
from morpheus_amd.ajtai import AjtaiI8, key_from_seed

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.goldilocks_cuda_init() == 0


@pytest.mark.parametrize("rows,n,batch", [(256, 4096, 1), (256, 4096, 3), (16, 64, 1), (32, 1 << 15, 2),
                                          (256, (1 << 14) + 320, 8), (64, 8192, 32), (48, 4096 + 64, 5),
                                          (256, 8192, 12), (80, (1 << 13) + 192, 17)])
def test_ajtai_i8_matches_oracle(rows, n, batch):
    print("ajtai", rows, n, batch, flush=True)
    key = key_from_seed(7 + rows)
    rng = np.random.default_rng(rows + n + batch)
    z = rng.integers(0, 1 << 64, batch * n, dtype=np.uint64)  # includes non-canonical words
    A = AjtaiI8(key, rows, n)
    got = A.commit(z, batch)
    want = O.ajtai_i8_commit(key, rows, n, z, batch).reshape(batch, rows)
    assert np.array_equal(got, want)
    for r in (0, rows // 2 + 1, rows - 1):
        want_row = O.ajtai_i8_row(key, r, n)
        assert np.array_equal(A.matrix_row(r, 0, n), want_row)
        assert np.array_equal(A.matrix_row(r, 13, min(n - 13, 77)), want_row[13:13 + min(n - 13, 77)])
    A.free()


def test_ajtai_i8_extremes_and_linearity():
    key = key_from_seed(99)
    rows, n = 256, 4096
    A = AjtaiI8(key, rows, n)
    P = M.P
    for fill in (0, P - 1, (1 << 64) - 1):
        z = np.full(n, fill, dtype=np.uint64)
        assert np.array_equal(A.commit(z)[0], O.ajtai_i8_commit(key, rows, n, z, 1))
    rng = np.random.default_rng(5)
    a = rng.integers(0, P, n, dtype=np.uint64)
    b = rng.integers(0, P, n, dtype=np.uint64)
    s = np.array([(int(x) + int(y)) % P for x, y in zip(a, b)], dtype=np.uint64)
    ca, cb, cs = A.commit(a)[0], A.commit(b)[0], A.commit(s)[0]
    assert [int(v) for v in cs] == [(int(x) + int(y)) % P for x, y in zip(ca, cb)]
    A.free()


@pytest.mark.parametrize("rows,n,n_used,batch", [(256, 1 << 14, 1 << 12, 1), (64, 8192, 64, 3), (32, 4096, 4032, 16),
                                                 (256, 1 << 14, 1 << 14, 2)])
def test_ajtai_i8_prefix_commit(rows, n, n_used, batch):
    print("ajtai prefix", rows, n, n_used, batch, flush=True)
    key = key_from_seed(3 + rows)
    rng = np.random.default_rng(n_used + batch)
    z = rng.integers(0, 1 << 64, batch * n_used, dtype=np.uint64)
    A = AjtaiI8(key, rows, n)
    got = A.commit_prefix(z, n_used, batch)
    want = O.ajtai_i8_commit(key, rows, n_used, z, batch).reshape(batch, rows)
    assert np.array_equal(got, want)
    A.free()
