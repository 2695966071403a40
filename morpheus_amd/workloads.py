"""Seeded synthetic inputs of the benchmark workloads, shared by bench.py and the tests so
that the inputs a test checks are the ones the bench times.

- Batched proving: `count` sumcheck + Basefold instances of size 2^u, u ~ U[12, 22]
  (continuous, seeded), each f and g uniform in [0, p) and zero-padded to the next power of
  two.
- Matmul proving service: requests C = A B with m, k, n uniform over {256, ..., 4096}, A and
  B uniform int8.
"""
import numpy as np

P = 0xFFFFFFFF00000001
SERVE_DIMS = (256, 512, 1024, 2048, 4096)


def batch_sizes(count, seed=4):
    """Real size and num_vars of each instance: log2 size ~ U[12, 22] (seeded); the
    instance is padded with zeros to 2^num_vars."""
    u = np.random.default_rng(seed).uniform(12.0, 22.0, count)
    n = np.floor(2.0 ** u).astype(np.int64)
    nv = np.ceil(np.log2(n)).astype(np.int64)
    return n, nv


def batch_instance_data(count, seed=5, size_seed=4):
    """[(f, g, label)] of the batched-proving workload: f and g are 2^num_vars words each,
    uniform over the first `size` entries and zero after."""
    sizes, nvs = batch_sizes(count, size_seed)
    rng = np.random.default_rng(seed)
    out = []
    for i, (n, nv) in enumerate(zip(sizes, nvs)):
        n, nv = int(n), int(nv)
        f = np.zeros(1 << nv, dtype=np.uint64)
        g = np.zeros(1 << nv, dtype=np.uint64)
        f[:n] = rng.integers(0, P, n, dtype=np.uint64)
        g[:n] = rng.integers(0, P, n, dtype=np.uint64)
        out.append((f, g, b"c4-%d" % i))
    return out


def serve_trace(count, seed=8):
    """m, k, n of each request, uniform over SERVE_DIMS (seeded)."""
    return np.random.default_rng(seed).choice(SERVE_DIMS, (count, 3))


def matmul_operands(dims, seed=9):
    """int8 A (m x k) and B (k x n) of each request, uniform over [-128, 127]."""
    rng = np.random.default_rng(seed)
    out = []
    for m, k, n in dims:
        a = rng.integers(-128, 128, (int(m), int(k)), dtype=np.int8)
        b = rng.integers(-128, 128, (int(k), int(n)), dtype=np.int8)
        out.append((a, b))
    return out
