"""Batched proving of independent sumcheck + Basefold instances of uneven size
(include/morpheus_prover.h, mam_prove_batch): each instance proves sum f*g with the product
sumcheck and opens a Basefold commitment of f at the sumcheck challenges, on one transcript."""
import ctypes
import time

import numpy as np

from ._lib import BatchInstance_t, lib
from .device import DeviceBuffer, check
from .prover import proof_len

_u64p = ctypes.POINTER(ctypes.c_uint64)


class Instance:
    """One instance: f and g (2^num_vars base words each) resident on the device."""

    def __init__(self, f, g, label, d_fg=None):
        f = np.ascontiguousarray(f, dtype=np.uint64)
        g = np.ascontiguousarray(g, dtype=np.uint64)
        assert f.size == g.size and f.size & (f.size - 1) == 0
        self.num_vars = f.size.bit_length() - 1
        self.label = bytes(label)
        self.d_fg = d_fg if d_fg is not None else DeviceBuffer.from_numpy(np.concatenate([f, g]))


def _outputs(instances, arr, log_rate, num_queries, proof_bufs):
    """Output arrays per instance, pointed to by arr. proof_bufs (optional): one array per
    instance, at least proof_len words, that receives the proof (e.g. PinnedArray.a, which the
    device copies into directly; fresh pageable arrays cost a staging copy and page faults)."""
    outs = []
    for i, inst in enumerate(instances):
        nv = inst.num_vars
        plen = proof_len(nv, log_rate, num_queries)
        if proof_bufs is not None:
            pb = proof_bufs[i]
            assert pb.dtype == np.uint64 and pb.flags.c_contiguous and pb.size >= plen
            proof = pb[:plen]
        else:
            proof = np.zeros(plen, dtype=np.uint64)
        o = {"round_msgs": np.zeros((nv, 3, 2), dtype=np.uint64), "challenges": np.zeros((nv, 2), dtype=np.uint64),
             "final_evals": np.zeros((2, 2), dtype=np.uint64), "root": np.zeros(4, dtype=np.uint64), "proof": proof}
        outs.append(o)
        a = arr[i]
        a.d_fg, a.num_vars = inst.d_fg.ptr, nv
        a.label, a.label_len = inst.label, len(inst.label)
        for k in ("round_msgs", "challenges", "final_evals", "root", "proof"):
            setattr(a, k, o[k].ctypes.data_as(_u64p))
    return outs


def prove_batch(instances, table, log_rate=1, hash_kind=0, num_queries=100, workers=1, proof_bufs=None):
    """Returns one dict per instance (round_msgs, challenges, final_evals, root, proof,
    start_ms, end_ms) in input order. proof_bufs: see _outputs."""
    d_table = table.upload()
    n = len(instances)
    arr = (BatchInstance_t * max(n, 1))()
    outs = _outputs(instances, arr, log_rate, num_queries, proof_bufs)
    check(lib.mam_prove_batch(arr, n, d_table.ptr, table.num_vars, log_rate, hash_kind, num_queries, workers),
          "mam_prove_batch")
    for i, o in enumerate(outs):
        o["start_ms"], o["end_ms"] = arr[i].start_ms, arr[i].end_ms
    return outs


def prove_batch_graph(instances, table, log_rate=1, hash_kind=0, num_queries=100, workers=8, branches=2,
                      max_wg=256, proof_bufs=None):
    """The graph form (mam_prove_batch_graph): every instance's sumcheck and commitment in one
    hipGraph launch with the transcripts on the device, then the openings on `workers`
    threads. Same outputs as prove_batch, word for word. The graph is kept for the next call
    with the same instance buffers and settings."""
    d_table = table.upload()
    n = len(instances)
    arr = (BatchInstance_t * max(n, 1))()
    outs = _outputs(instances, arr, log_rate, num_queries, proof_bufs)
    check(lib.mam_prove_batch_graph(arr, n, d_table.ptr, table.num_vars, log_rate, hash_kind, num_queries, workers,
                                    branches, max_wg), "mam_prove_batch_graph")
    for i, o in enumerate(outs):
        o["start_ms"], o["end_ms"] = arr[i].start_ms, arr[i].end_ms
    return outs


def graph_phase_ms():
    """(all instances ready, call end) in ms from the start of the last prove_batch_graph call:
    the openings start per instance inside the graph's run."""
    v = (ctypes.c_double * 2)()
    check(lib.mam_batch_graph_phase_ms(v), "mam_batch_graph_phase_ms")
    return float(v[0]), float(v[1])


def autotune_graph(instances, table, log_rate=1, hash_kind=0, num_queries=100, branches=(0, 2, 4),
                   max_wgs=(64, 256, 1024), workers=(4, 8, 16), repeats=2):
    """Settings of the graph form that prove this batch fastest: the commitment branches and
    the workgroup cap per instance and pass (the sumcheck tile) with 8 opening workers, then
    the opening workers. Each setting is timed after one untimed call that captures its
    graph. Returns ({"branches", "max_wg", "workers"}, {setting: seconds})."""
    def timed(b, m, w):
        prove_batch_graph(instances, table, log_rate, hash_kind, num_queries, w, b, m)
        best = None
        for _ in range(repeats):
            t0 = time.perf_counter()
            prove_batch_graph(instances, table, log_rate, hash_kind, num_queries, w, b, m)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best
    times = {}
    for b in branches:
        for m in max_wgs:
            times[(b, m, 8)] = timed(b, m, 8)
    b, m, _ = min(times, key=times.get)
    for w in workers:
        if (b, m, w) not in times:
            times[(b, m, w)] = timed(b, m, w)
    best = min(times, key=times.get)
    return {"branches": best[0], "max_wg": best[1], "workers": best[2]}, times


def autotune_workers(instances, table, log_rate=1, hash_kind=0, num_queries=100, candidates=(1, 2, 4, 8, 12, 16),
                     stride=4, repeats=2):
    """Worker count that proves every `stride`-th instance fastest, after one untimed warm-up
    at the largest candidate (its threads and buffers). Returns (best, {workers: seconds})."""
    sub = [instances[i] for i in range(0, len(instances), stride)] or list(instances)
    prove_batch(sub, table, log_rate, hash_kind, num_queries, workers=max(candidates))
    times = {}
    for w in candidates:
        best = None
        for _ in range(repeats):
            t0 = time.perf_counter()
            prove_batch(sub, table, log_rate, hash_kind, num_queries, workers=w)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        times[w] = best
    return min(times, key=times.get), times
