"""GPU parity of the fused permute + partial evaluation and of the matmul proof C = A B
against the oracle (every transcript step restated with oracle functions)."""
import numpy as np
import pytest

import pyoracle as O

import morpheus_amd as M
from morpheus_amd.ajtai import AjtaiI8, key_from_seed
from morpheus_amd.matmul import MatmulJob, prove_matmul, serve

pytestmark = pytest.mark.gpu
P = O.P


@pytest.fixture(scope="module", autouse=True)
def gpu_init():
    assert M.lib.goldilocks_cuda_init() == 0


def up(a, dtype=np.uint64):
    a = np.ascontiguousarray(a, dtype=dtype)
    return M.DeviceBuffer.from_numpy(a.view(np.uint64) if dtype != np.uint64 else a)


@pytest.mark.parametrize("field", ["p", "q"])
@pytest.mark.parametrize("n,m,half,perm_seed", [(6, 2, 3, 1), (10, 4, 5, 2), (12, 7, 6, 3), (13, 0, 6, 4),
                                                (16, 8, 8, 5)])
def test_fused_permute_peval(field, n, m, half, perm_seed):
    print("fused", field, n, m, half, flush=True)
    rng = np.random.default_rng(perm_seed)
    perm = list(rng.permutation(n))
    evals = rng.integers(0, 1 << 64, 1 << n, dtype=np.uint64)
    eq = rng.integers(0, 1 << 64, 2 << m, dtype=np.uint64)
    lo, hi = O.split_luts(n, perm, half)
    out_size = 1 << (n - m)
    mod = P if field == "p" else O.Q
    want = O.fused_permute_peval(mod, evals, eq, lo, hi, n, m, half, out_size)
    lut = np.concatenate([lo, hi]).astype(np.uint32)
    if lut.size % 2:
        lut = np.concatenate([lut, np.zeros(1, dtype=np.uint32)])
    d_lut = M.DeviceBuffer.from_numpy(lut.view(np.uint64))
    d_out = M.DeviceBuffer(2 * out_size)
    fn = M.lib.fused_permute_partial_eval_ffi if field == "p" else M.lib.agl_fused_permute_partial_eval_ffi
    assert fn(up(evals).ptr, d_out.ptr, up(eq).ptr, d_lut.ptr, d_lut.ptr + 4 * lo.size, n, m, half, out_size, 0) == 0
    assert np.array_equal(d_out.to_numpy(), want)


def oracle_matmul(a, b, label, table, lr, nq, key=None, rows=0):
    return O.matmul_prove(a, b, label, table.entries, table.num_vars, lr, nq, key, rows)


@pytest.mark.parametrize("dims,ajtai", [((3, 4, 3), True), ((4, 2, 5), False), ((2, 6, 4), True)])
def test_matmul_proof_matches_oracle(dims, ajtai):
    print("matmul", dims, ajtai, flush=True)
    lm, lk, ln = dims
    rng = np.random.default_rng(sum(dims))
    a = rng.integers(-128, 128, (1 << lm, 1 << lk))
    b = rng.integers(-128, 128, (1 << lk, 1 << ln))
    lr, nq = 1, 9
    tnv = max(lm + lk, lk + ln, lm + ln)
    table = M.BasefoldTable.generate(tnv, lr, tnv, 11)
    key = key_from_seed(77)
    aj = AjtaiI8(key, 64, 1 << 10) if ajtai else None
    job = MatmulJob.from_int8(a, b, b"mm-test")
    got = prove_matmul(job, table, lr, 0, nq, aj)
    want = oracle_matmul(a, b, b"mm-test", table, lr, nq, key if ajtai else None, 64)
    for k in ("roots", "point", "claim", "round_msgs", "challenges", "final_evals", "proof_a", "proof_b", "proof_c"):
        assert np.array_equal(got[k], want[k]), k
    if ajtai:
        assert np.array_equal(got["ajtai"], want["ajtai"])
    # the first round message sums to the claim; the openings' evaluations are the claims
    g = got["round_msgs"][0]
    assert [(int(g[0][c]) + int(g[1][c])) % P for c in range(2)] == [int(x) for x in got["claim"]]
    assert np.array_equal(got["proof_a"][:2], got["final_evals"][0])
    assert np.array_equal(got["proof_b"][:2], got["final_evals"][1])
    assert np.array_equal(got["proof_c"][:2], got["claim"])
    if aj:
        aj.free()


def test_service_results_equal_alone_and_in_the_mix():
    print("serve", flush=True)
    rng = np.random.default_rng(21)
    dims = [(3, 4, 3), (5, 3, 4), (3, 2, 3), (4, 5, 3), (3, 3, 6), (6, 2, 3), (4, 4, 4)]  # |C| >= 64 for Ajtai
    jobs = []
    for idx, (lm, lk, ln) in enumerate(dims):
        a = rng.integers(-128, 128, (1 << lm, 1 << lk))
        b = rng.integers(-128, 128, (1 << lk, 1 << ln))
        jobs.append(MatmulJob.from_int8(a, b, b"serve-%d" % idx))
    tnv = max(max(lm + lk, lk + ln, lm + ln) for lm, lk, ln in dims)
    table = M.BasefoldTable.generate(tnv, 1, tnv, 5)
    aj = AjtaiI8(key_from_seed(9), 32, 1 << 10)
    alone = [prove_matmul(j, table, 1, 0, 7, aj) for j in jobs]
    arrivals = np.cumsum(rng.exponential(2.0, len(jobs)))
    for workers, budget in ((3, 64 << 30), (2, 1 << 20)):  # the small budget forces one-at-a-time admission
        outs, admit, done = serve(jobs, arrivals, table, 1, 0, 7, aj, workers, budget)
        for i in range(len(jobs)):
            for k in alone[i]:
                assert np.array_equal(outs[i][k], alone[i][k]), (workers, i, k)
        assert np.all(admit >= arrivals - 1e-3) and np.all(done >= admit)
    aj.free()
