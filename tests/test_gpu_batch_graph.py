"""The graph form of batched proving (mam_prove_batch_graph: sumchecks with device
transcripts and commitments in one hipGraph launch, openings on worker threads) against the
thread form (mam_prove_batch), word for word, over sizes that exercise every pass shape
(one-round last passes, two-round passes, tiny arrays), both hashes, every branch count and
small and large workgroup caps; the graph is replayed with new input values."""
import numpy as np
import pytest

import morpheus_amd as M
from morpheus_amd.batch import Instance, prove_batch, prove_batch_graph

pytestmark = pytest.mark.gpu
P = M.P


def make_instances(sizes, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i, nv in enumerate(sizes):
        f = rng.integers(0, P, 1 << nv, dtype=np.uint64)
        g = rng.integers(0, P, 1 << nv, dtype=np.uint64)
        out.append(Instance(f, g, f"graph-{i}".encode()))
    return out


def same(a, b):
    for x, y in zip(a, b):
        for k in ("round_msgs", "challenges", "final_evals", "root", "proof"):
            assert np.array_equal(x[k], y[k]), k


@pytest.mark.parametrize("hash_kind,branches,max_wg", [(0, 2, 256), (1, 0, 1), (0, 4, 4096), (0, 1, 3)])
def test_graph_batch_matches_thread_batch(hash_kind, branches, max_wg):
    print(f"graph batch hash={hash_kind} branches={branches} max_wg={max_wg}", flush=True)
    sizes = [2, 3, 4, 5, 7, 9, 12, 14, 11, 6, 13, 2]
    inst = make_instances(sizes, 11)
    table = M.BasefoldTable.generate(14, 1, 14, 42)
    want = prove_batch(inst, table, 1, hash_kind, 12, workers=4)
    got = prove_batch_graph(inst, table, 1, hash_kind, 12, workers=4, branches=branches, max_wg=max_wg)
    same(got, want)
    # the same graph replayed over new values in the same buffers
    rng = np.random.default_rng(99)
    for it in inst:
        vals = rng.integers(0, P, 2 << it.num_vars, dtype=np.uint64)
        M.lib.cuda_memcpy_htod(it.d_fg.ptr, vals.ctypes.data, vals.size * 8)
    want2 = prove_batch(inst, table, 1, hash_kind, 12, workers=2)
    got2 = prove_batch_graph(inst, table, 1, hash_kind, 12, workers=4, branches=branches, max_wg=max_wg)
    same(got2, want2)
    assert M.lib.mam_batch_graph_release() == 0


def test_graph_batch_rejects_bad_settings():
    inst = make_instances([3], 1)
    table = M.BasefoldTable.generate(3, 1, 3, 42)
    with pytest.raises(M.GpuError):
        prove_batch_graph(inst, table, 1, 0, 4, workers=1, branches=5)
    with pytest.raises(M.GpuError):
        prove_batch_graph(inst, table, 1, 0, 4, workers=1, max_wg=0)
