"""Benchmark of the prover hot path on one MI355X.

Default workload (the headline metric's configuration): Basefold commit-and-open of one
multilinear polynomial with 2^22 uniform Goldilocks evaluations, rate 1/2 (2^23 codeword),
Monolith-w12 Merkle trees, 22 fold rounds with Ext2 challenges from the Poseidon2-w8
transcript, 100 query openings. A step is one full commit + open with the evaluations
already resident in HBM; the folding table (setup) is built once before timing.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload all|basefold|sumcheck|ajtai|batch|serve|...]

Prints one JSON line. `value` is proved elements per second (2^22 per step). The default
workload `all` adds, in the same line, a record per other benchmark configuration (the 2^20
product sumcheck, the int8 Ajtai commit, the 64-instance batch, the matmul proving service with
its proofs/s and p50/p99 latency), each with its roofline, CPU baseline and verification.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "proved Goldilocks elems/s (sumcheck+Basefold, 2^22); proofs/s + p99 latency"
HBM_PEAK_GBS = 8000.0
# VALU peak: 32-bit lane-operations per second, 256 CUs x 64 lanes x 2.4 GHz (AMD's MI355X FP32
# vector figure, 157.3 TFLOP/s, counts an FMA as 2 flops on 2 packed lanes: 157.3 / 4). Hash
# rooflines divide the algorithmic operation count of the work (counted from the definition,
# ALG_OPS_PER_PERM) by the measured time and by this peak. The VALU probe (valu_peak()) is
# reported beside it: chains of one instruction form issue at ~4.1 cycles per wave instruction.
VALU_NOMINAL_TOPS = 39.3


VALU_PROBE = {}


def valu_peak(M):
    """Measured VALU issue rate (wave instructions x 64 lanes, Tops/s): the fastest of the
    single-instruction chain probes (v_xad_u32, v_mad_u64_u32, v_lshl_add_u64, v_add_u32),
    8192 workgroups x 256 threads x 2048 x 8 steps each."""
    blocks, iters = 8192, 2048
    names = ("v_xad_u32", "v_mad_u64_u32", "v_lshl_add_u64", "v_add_u32")
    for kind, name in enumerate(names):
        ms = ctypes.c_double()
        M.check(M.lib.mam_valu_probe(kind, blocks, iters, ctypes.byref(ms)), "mam_valu_probe")
        waves = blocks * 256 // 64
        VALU_PROBE[name] = waves * iters * 8 * 64 / (ms.value / 1e3) / 1e12
    return max(VALU_PROBE.values())
# Algorithmic 32-bit VALU operations of one Monolith-w12 permutation, counted from its
# definition (goldilocks-cuda-rs/src/cpu_monolith.rs:65-171) by a fixed rule:
#   7 concrete layers x 144 multiply-adds of a 64-bit word by a small constant, 2 ops each
#     (one 32 x 32 + 64 multiply-add per 32-bit half), plus 12 row reductions of 3 ops;
#   5 x 12 round-constant additions, 2 ops each;
#   6 x 11 bricks x_j += x_(j-1)^2 mod p, 12 ops each (three 32 x 32 products, carries and a
#     128 -> 64 bit reduction);
#   6 x 4 bars, 24 ops each (per 32-bit half: three byte rotations of 3 ops, the chi step
#     and the final rotation).
MONO_ALG_OPS = {"concrete": 7 * (144 * 2 + 12 * 3), "round_constants": 5 * 12 * 2, "bricks": 6 * 11 * 12,
                "bars": 6 * 4 * 24}
ALG_OPS_PER_PERM = float(sum(MONO_ALG_OPS.values()))
# VALU lane-operations the leaf kernel executes per Monolith permutation: SQ_INSTS_VALU x 64
# over the permutations of one k_tree_levels<MonoHash, SRC_GL> dispatch (rocprofv3 --pmc,
# profiles/merkle_leaf_kernel_2p22.md). Reported beside the algorithmic count.
VALU_OPS_PER_PERM = 4979.0
# VALU instructions one lane executes for one full Monolith permutation with every lane
# active: SQ_INSTS_VALU / SQ_WAVES of k_mono_permute_n (rocprofv3 --pmc,
# profiles/monolith_permutation_valu.md). Lane-operations/s = permutations/s x this.
MONO_PERM_VALU = 5291.0
# HBM bytes of that dispatch in the 2^22 proof (2^21 leaves: one copy of the replicated
# codeword, digests stored at both copies): FETCH_SIZE x 2 (a 256 MiB copy reads back as
# 128 MiB on this counter) + WRITE_SIZE (exact for the copy), same profile file.
LEAF_LAUNCH_TRAFFIC_BYTES = (32.6 + 128.0) * 2 ** 20
# HBM bytes of one int8 Ajtai GEMM launch (witness 2^22, 256 rows, batch 1): FETCH_SIZE
# 540758 KiB x 2.0 (copy calibration) + WRITE_SIZE 8192 KiB, from the rocprofv3 --pmc runs.
AJTAI_GEMM_TRAFFIC_BYTES = (540758.0 * 2.0 + 8192.0) * 1024
# Wave-level VALU instructions of k_ring_dense (from its gfx950 ISA): one column fill
# (ChaCha8 block, rejection check, two LDS rows) and one set bit of one 7-row group.
RING_PRG_VALU = 440.0
RING_BIT_VALU = 18.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=["all", "basefold", "sumcheck", "ajtai", "ring_ajtai", "batch", "hash",
                                           "serve", "batch_open"],
                    default="all")
    ap.add_argument("--instances", type=int, default=64, help="instances per batch (batch workload)")
    ap.add_argument("--workers", type=int, default=0, help="batch workers (0: autotune)")
    ap.add_argument("--batch", type=int, default=1, help="witnesses per Ajtai commit")
    ap.add_argument("--polys", type=int, default=8, help="batch_open: committed polynomials")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="serve: mean arrival rate, proofs/s (0: 1.25 x the one-at-a-time capacity)")
    ap.add_argument("--hbm-budget-gib", type=int, default=96, help="serve: admission budget")
    ap.add_argument("--nv", type=int, default=None)
    ap.add_argument("--queries", type=int, default=100)
    ap.add_argument("--hash", choices=["monolith", "poseidon2"], default="monolith",
                    help="Merkle hash of the basefold workload (the reference default is Monolith-w12)")
    ap.add_argument("--cpu-sample-nv", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-sweep", action="store_true",
                    help="only time a plain device copy at 1 MiB .. 512 MiB (the achievable HBM rate by size)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy (PMC byte-counter calibration)")
    return ap.parse_args()


def gen_inputs(nv, seed=1):
    """Uniform Goldilocks words from a seeded generator (rejection of values >= p)."""
    P = 0xFFFFFFFF00000001
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 1 << 64, 1 << nv, dtype=np.uint64)
    bad = x >= np.uint64(P)
    while bad.any():
        x[bad] = rng.integers(0, 1 << 64, int(bad.sum()), dtype=np.uint64)
        bad = x >= np.uint64(P)
    return x


def profile_collect(lib):
    ms = (ctypes.c_double * 7)()
    cnt = (ctypes.c_long * 7)()
    work = (ctypes.c_double * 7)()
    assert lib.mam_profile_collect(ms, cnt, work) == 0
    return [(ms[i], cnt[i], work[i]) for i in range(7)]


def cpu_facts():
    """CPU model, visible cores and the thread count of the all-core runs (the job's CPU share:
    OMP_NUM_THREADS when set, else every visible core)."""
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    try:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or nproc
    except ValueError:
        threads = nproc
    return {"cpu_model": model, "nproc": nproc, "threads_all": min(threads, nproc)}


def oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    return O


def timed_threads(O, threads, fn):
    O.set_threads(threads)
    t0 = time.perf_counter()
    r = fn()
    return time.perf_counter() - t0, r


def cpu_basefold_all(nv_all, nv_single, lr, nq):
    """The oracle's commit + open_ext2 (the opening takes the commitment's tree, as the
    reference's CPU opening takes it from the commitment cache) on every core at nv_all and on
    one core at nv_single; value is the all-core rate."""
    O = oracle()
    f = cpu_facts()

    def run(nv):
        ev = gen_inputs(nv, seed=3)
        pt = np.resize(gen_inputs(6, seed=5), 2 * nv)
        table = O.basefold_table(nv, lr, nv, 42)

        def go():
            cw, bh, tree = O.basefold_commit(ev, nv, lr, O.MONOLITH)
            O.basefold_open_ext2(cw, bh, tree[-4:], nv, lr, O.MONOLITH, pt, table, nv, O.Transcript(b"bench"), nq,
                                 tree=tree)
        return go
    da, _ = timed_threads(O, f["threads_all"], run(nv_all))
    ds, _ = timed_threads(O, 1, run(nv_single))
    return {"value": (1 << nv_all) / da, "unit": "elems/s", "cores": f["threads_all"], "kind": "port",
            "sample": f"oracle commit + open_ext2 of 2^{nv_all} evaluations (rate 1/2, Monolith, {nq} queries) on "
                      f"{f['threads_all']} threads, {da:.2f} s",
            "single_thread": {"value": (1 << nv_single) / ds, "unit": "elems/s", "cores": 1,
                              "sample": f"the same at 2^{nv_single} on one thread, {ds:.2f} s"},
            "cpu_model": f["cpu_model"], "nproc": f["nproc"]}


def cpu_sumcheck_all(nv):
    O = oracle()
    f = cpu_facts()
    fg = gen_inputs(nv + 1, seed=6)
    ext = np.zeros((2, 1 << nv, 2), dtype=np.uint64)
    ext[:, :, 0] = fg.reshape(2, -1)
    out = {}
    for name, thr in (("all", f["threads_all"]), ("single", 1)):
        dt, _ = timed_threads(O, thr, lambda: O.sumcheck_prove_ext2(O.Transcript(b"bench"), ext.copy(), 2, nv))
        out[name] = dt
    return {"value": (2 << nv) / out["all"], "unit": "elems/s", "cores": f["threads_all"], "kind": "port",
            "sample": f"oracle Ext2 sumcheck of f*g, 2 x 2^{nv} evaluations, {f['threads_all']} threads, "
                      f"{out['all']:.2f} s",
            "single_thread": {"value": (2 << nv) / out["single"], "unit": "elems/s", "cores": 1,
                              "sample": f"the same on one thread, {out['single']:.2f} s"},
            "cpu_model": f["cpu_model"], "nproc": f["nproc"]}


def cpu_ajtai_all(nv):
    O = oracle()
    f = cpu_facts()
    from morpheus_amd.ajtai import key_from_seed
    z = gen_inputs(nv, seed=11)
    out = {}
    for name, thr in (("all", f["threads_all"]), ("single", 1)):
        out[name], _ = timed_threads(O, thr, lambda: O.ajtai_i8_commit(key_from_seed(2024), 256, 1 << nv, z, 1))
    return {"value": (1 << nv) / out["all"], "unit": "elems/s", "cores": f["threads_all"], "kind": "port",
            "sample": f"oracle int8 Ajtai (ChaCha8 matrix rows + exact matvec), 256 x 2^{nv}, {f['threads_all']} "
                      f"threads, {out['all']:.2f} s",
            "single_thread": {"value": (1 << nv) / out["single"], "unit": "elems/s", "cores": 1,
                              "sample": f"the same on one thread, {out['single']:.2f} s"},
            "cpu_model": f["cpu_model"], "nproc": f["nproc"]}


def cpu_batch_sample(data, stride, lr, nq):
    """The oracle proving every stride-th batched instance (product sumcheck, commit, open on
    one transcript) on every core; value is proved elements per second over the sample."""
    O = oracle()
    f = cpu_facts()
    sample = data[::stride]
    tnv = max(int(fv.size).bit_length() - 1 for fv, _, _ in sample)
    table = O.basefold_table(tnv, lr, tnv, 42)

    def go():
        for fv, gv, label in sample:
            nv = int(fv.size).bit_length() - 1
            ext = np.zeros((2, 1 << nv, 2), dtype=np.uint64)
            ext[0, :, 0], ext[1, :, 0] = fv, gv
            t = O.Transcript(label)
            sc = O.sumcheck_prove_ext2(t, ext, 2, nv)
            cw, bh, tree = O.basefold_commit(fv, nv, lr, O.MONOLITH)
            O.basefold_open_ext2(cw, bh, tree[-4:], nv, lr, O.MONOLITH, sc["challenges"].ravel(), table, tnv, t, nq,
                                 tree=tree)
    dt, _ = timed_threads(O, f["threads_all"], go)
    elems = sum(int(fv.size) for fv, _, _ in sample)
    return {"value": elems / dt, "unit": "elems/s", "cores": f["threads_all"], "kind": "port",
            "sample": f"oracle sumcheck + commit + open of every {stride}th instance ({len(sample)} instances, "
                      f"{elems} padded elements) on {f['threads_all']} threads, {dt:.2f} s",
            "cpu_model": f["cpu_model"], "nproc": f["nproc"]}


def cpu_baseline_basefold(nv, lr, nq):
    """Times the oracle (single-threaded C restatement) on a bounded sample."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    ev = gen_inputs(nv, seed=3)
    pt = np.resize(gen_inputs(6, seed=5), 2 * nv)
    table = O.basefold_table(nv, lr, nv, 42)
    t0 = time.perf_counter()
    cw, bh, tree = O.basefold_commit(ev, nv, lr, O.MONOLITH)
    O.basefold_open_ext2(cw, bh, tree[-4:], nv, lr, O.MONOLITH, pt, table, nv, O.Transcript(b"bench"), nq)
    dt = time.perf_counter() - t0
    return {"value": (1 << nv) / dt, "unit": "elems/s", "cores": 1, "kind": "port",
            "sample": f"oracle commit+open_ext2 of 2^{nv} evaluations (rate 1/2, Monolith, {nq} queries), "
                      f"single thread, {dt:.2f} s"}


def cpu_baseline_sumcheck(nv):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    f = gen_inputs(nv + 1, seed=6)
    ext = np.zeros((2, 1 << nv, 2), dtype=np.uint64)
    ext[:, :, 0] = f.reshape(2, -1)
    t0 = time.perf_counter()
    O.sumcheck_prove_ext2(O.Transcript(b"bench"), ext, 2, nv)
    dt = time.perf_counter() - t0
    return {"value": (2 << nv) / dt, "unit": "elems/s", "cores": 1, "kind": "port",
            "sample": f"oracle Ext2 sumcheck of f*g, 2 x 2^{nv} evaluations, single thread, {dt:.2f} s"}


def cpu_baseline_hash(nv):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    n = 1 << nv
    st = np.random.default_rng(7).integers(0, 1 << 63, 12 * n, dtype=np.uint64)
    t0 = time.perf_counter()
    for i in range(n):
        O.L.or_monolith_permute(O._p(st[12 * i:12 * i + 12]))
    dt = time.perf_counter() - t0
    return {"value": n / dt, "unit": "perms/s", "cores": 1, "kind": "port",
            "sample": f"oracle Monolith-w12 permutation, 2^{nv} states (ctypes call per state), single thread, "
                      f"{dt:.2f} s"}


def cpu_baseline_ring_ajtai(log_cols):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    n = 1 << log_cols
    z = np.random.default_rng(12).integers(0, 1 << 64, n, dtype=np.uint64)
    t0 = time.perf_counter()
    O.ring_commit_binary(RING_KEY, z, n, 1)
    dt = time.perf_counter() - t0
    return {"value": 64 * n / dt, "unit": "elems/s", "cores": 1, "kind": "port",
            "sample": f"oracle ring Ajtai (ChaCha8 matrix + negacyclic products), 42 rows x 2^{log_cols} ring "
                      f"columns of 64 binary coefficients, single thread, {dt:.2f} s"}


def cpu_baseline_ajtai(nv):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    from morpheus_amd.ajtai import key_from_seed
    z = gen_inputs(nv, seed=11)
    t0 = time.perf_counter()
    O.ajtai_i8_commit(key_from_seed(2024), 256, 1 << nv, z, 1)
    dt = time.perf_counter() - t0
    return {"value": (1 << nv) / dt, "unit": "elems/s", "cores": 1, "kind": "port",
            "sample": f"oracle int8 Ajtai (ChaCha8 matrix generation + exact matvec), 256 x 2^{nv}, "
                      f"single thread, {dt:.2f} s"}


def bench_basefold(args, M):
    nv = args.nv or 22
    p2 = getattr(args, "hash", "monolith") == "poseidon2"
    lr, nq, hash_kind = 1, args.queries, (M.HASH_POSEIDON2 if p2 else M.HASH_MONOLITH)
    lib = M.lib
    ev = gen_inputs(nv)
    pt = np.resize(gen_inputs(8, seed=2), 2 * nv)
    table = M.BasefoldTable.generate(nv, lr, nv, 42)
    d_table = table.upload()
    d_ev = M.DeviceBuffer.from_numpy(ev)
    cw_len = 1 << (nv + lr)
    d_cw, d_bh, d_tree = M.DeviceBuffer(cw_len), M.DeviceBuffer(1 << nv), M.DeviceBuffer((cw_len - 1) * 4)
    root = np.zeros(4, dtype=np.uint64)
    U = ctypes.POINTER(ctypes.c_uint64)
    # the proof lands in a page-locked buffer allocated once (mam_host_alloc)
    proof_buf = M.PinnedArray(M.proof_len(nv, lr, nq))
    proof = proof_buf.a

    # argument objects made once (the calls themselves are the same as a caller's)
    root_p, pt_p, proof_p = root.ctypes.data_as(U), pt.ctypes.data_as(U), proof.ctypes.data_as(U)
    h = ctypes.c_void_p()
    h_ref = ctypes.byref(h)

    def step():
        M.check(lib.mam_basefold_commit_into(d_ev.ptr, nv, lr, hash_kind, d_cw.ptr, d_bh.ptr, d_tree.ptr, root_p),
                "commit")
        M.check(lib.mam_basefold_commitment_wrap(d_cw.ptr, d_bh.ptr, d_tree.ptr, root_p, nv, lr, hash_kind, h_ref),
                "wrap")
        t = M.Transcript(b"bench")
        M.check(lib.mam_basefold_open_ext2(h.value, pt_p, d_table.ptr, nv, ctypes.byref(t.t), nq, proof_p), "open")
        lib.mam_basefold_commitment_free(h.value)

    for _ in range(args.warmup):
        step()
    M.synchronize()
    # timed region: HIP events around the dominant kernel only (two events per step)
    lib.mam_profile_enable_mask(1 << 4)
    lat = []
    M.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        lat.append(time.perf_counter() - s0)
    M.synchronize()
    elapsed = time.perf_counter() - t0
    prof = profile_collect(lib)
    # every kernel group, on three extra untimed steps
    lib.mam_profile_enable(1)
    for _ in range(3):
        step()
    M.synchronize()
    prof_all = profile_collect(lib)
    lib.mam_profile_enable(0)
    verified = M.verify_ext2(root, pt, nv, lr, hash_kind, proof, nq, table, M.Transcript(b"bench"))
    if p2:
        # Poseidon2-w8 trees: no algorithmic op count is fixed for this permutation here, so the
        # line reports the hashing rate of every Merkle build and the proof rate only
        mk_ms, mk_n, mk_perm = prof_all[0]
        lat_ms = np.array(lat) * 1e3
        return {"metric": METRIC, "value": args.steps * (1 << nv) / elapsed, "unit": "elems/s", "n_gpus": 1,
                "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
                "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
                "data": "synthetic",
                "config": {"workload": "basefold_commit_open", "num_vars": nv, "log_rate": lr, "codeword_len": cw_len,
                           "merkle_hash": "poseidon2-w8", "transcript": "poseidon2-w8", "fold_rounds": nv,
                           "num_queries": nq, "table_seed": 42, "parallelism": "single-gpu"},
                "proofs_per_s": args.steps / elapsed, "p50_ms": float(np.percentile(lat_ms, 50)),
                "p99_ms": float(np.percentile(lat_ms, 99)), "verified": verified,
                "roofline": None,
                "merkle": {"perms_per_s": mk_perm / (mk_ms / 1e3) if mk_ms > 0 else 0.0,
                           "avg_build_ms": mk_ms / max(mk_n, 1), "builds_per_step": mk_n / 3}}
    # Dominant kernel: the commit tree's leaf launch, k_tree_levels<MonoHash, SRC_GL> hashing the
    # 2^21 base-field leaves of one copy of the replicated codeword (2^21 Monolith permutations).
    lk_ms, lk_n, lk_perm = prof[4]
    lk_avg_ms = lk_ms / max(lk_n, 1)
    perms_per_launch = lk_perm / max(lk_n, 1)
    perm_rate = perms_per_launch / (lk_avg_ms / 1e3) if lk_ms > 0 else 0.0
    ach = perm_rate * ALG_OPS_PER_PERM / 1e12
    roof = {"kernel": f"k_tree_levels<MonoHash, SRC_GL> (the commitment's leaf level: 2^{nv - 1} leaves of one "
                      f"codeword copy)",
            "bound": "valu", "unit": "Tops/s", "achieved": ach, "peak": VALU_NOMINAL_TOPS,
            "frac": ach / VALU_NOMINAL_TOPS,
            "peak_source": "MI355X FP32 vector 157.3 TFLOP/s / 4 (no packing, no FMA): one 32-bit lane-op per "
                           "lane per clock, 256 CUs x 64 lanes x 2.4 GHz",
            "achieved_from": "permutations per second of the launch x algorithmic 32-bit ops per permutation "
                             "(counted from the definition)",
            "alg_ops_per_perm": ALG_OPS_PER_PERM, "alg_ops_breakdown": dict(MONO_ALG_OPS),
            "traffic": LEAF_LAUNCH_TRAFFIC_BYTES if nv == 22 else None,
            "avg_launch_ms": lk_avg_ms, "launches": lk_n, "perms_per_launch": perms_per_launch,
            "perms_per_s": perm_rate,
            "executed": {"valu_ops_per_perm": VALU_OPS_PER_PERM, "tops": perm_rate * VALU_OPS_PER_PERM / 1e12,
                         "note": "SQ_INSTS_VALU x 64 per permutation (rocprofv3 --pmc); the kernel's instruction "
                                 "mix issues faster than one wave instruction per 4 cycles per SIMD"},
            "valu_probe_tops": dict(VALU_PROBE)}
    mk_ms, mk_n, mk_perm = prof_all[0]
    roof["all_merkle_builds"] = {"perms_per_s": mk_perm / (mk_ms / 1e3) if mk_ms > 0 else 0.0,
                                 "frac": (mk_perm / (mk_ms / 1e3) * ALG_OPS_PER_PERM / 1e12 / VALU_NOMINAL_TOPS)
                                 if mk_ms > 0 else 0.0,
                                 "avg_build_ms": mk_ms / max(mk_n, 1), "builds_per_step": mk_n / 3,
                                 "from": "3 profiled steps after the timed region",
                                 "note": "events around each tree's launches: an opening round's time also holds "
                                         "the product-form job riding on its first launch and, when the rounds are "
                                         "pre-queued, the wait for the host's next challenge at the end of its "
                                         "last launch"}
    hbm = {}
    for cat, name in ((1, "open_fold_oracle"), (2, "commit_transform")):
        ms, n, by = prof_all[cat]
        if ms > 0:
            gbs = by / (ms / 1e3) / 1e9
            hbm[name] = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "avg_ms": ms / n, "launch_groups": n}
    lat_ms = np.array(lat) * 1e3
    out = {"metric": METRIC, "value": args.steps * (1 << nv) / elapsed, "unit": "elems/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
           "config": {"workload": "basefold_commit_open", "num_vars": nv, "log_rate": lr, "codeword_len": cw_len,
                      "merkle_hash": "monolith-w12", "transcript": "poseidon2-w8", "fold_rounds": nv,
                      "num_queries": nq, "table_seed": 42, "parallelism": "single-gpu"},
           "proofs_per_s": args.steps / elapsed, "p50_ms": float(np.percentile(lat_ms, 50)),
           "p99_ms": float(np.percentile(lat_ms, 99)), "verified": verified, "roofline": roof,
           "hbm_kernels": hbm}
    return out


def verify_sumcheck_proof(M, got, fv, gv, nv, label=b"bench"):
    """The product's host checks of one f*g sumcheck proof: the verifier accepts it for the
    claim sum f g (numpy, mod p) and f(r), g(r) (host MLE evaluation at the challenges) are the
    final evaluations whose product ends the proof."""
    from morpheus_amd.hostfield import e2mul, gl_dot_np, mle_eval
    claim = (gl_dot_np(fv, gv), 0)
    ok, chs = M.sumcheck_verify(M.Transcript(label), claim, 2, nv, got["round_msgs"], got["final_eval"])
    r = np.array(chs, dtype=np.uint64).ravel()
    fr, gr = mle_eval(fv, r), mle_eval(gv, r)
    fe = [tuple(int(x) for x in got["final_evals"][i]) for i in range(2)]
    return bool(ok and fe == [fr, gr] and e2mul(fr, gr) == tuple(int(x) for x in got["final_eval"]))


def bench_sumcheck(args, M):
    nv = args.nv or 20
    f = gen_inputs(nv + 1)
    d = M.DeviceBuffer.from_numpy(f)
    lib = M.lib

    def step():
        return M.sumcheck_prove(M.Transcript(b"bench"), d, 2, nv, base=True)

    got = None
    for _ in range(args.warmup):
        got = step()
    M.synchronize()
    lib.mam_profile_enable_mask(1 << 3)  # the pass kernels only
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        step()
        lat.append(time.perf_counter() - s0)
    M.synchronize()
    elapsed = time.perf_counter() - t0
    prof = profile_collect(lib)
    lib.mam_profile_enable(0)
    ms, n, by = prof[3]
    gbs = by / (ms / 1e3) / 1e9 if ms > 0 else 0.0
    lat_ms = np.array(lat) * 1e3
    verified = verify_sumcheck_proof(M, step(), f[:1 << nv], f[1 << nv:], nv)
    return {"metric": METRIC, "value": args.steps * (2 << nv) / elapsed, "unit": "elems/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64",
            "data": "synthetic",
            "config": {"workload": "sumcheck_product_f_g", "num_vars": nv, "num_polys": 2,
                       "transcript": "poseidon2-w8", "parallelism": "single-gpu"},
            "proofs_per_s": args.steps / elapsed, "p50_ms": float(np.percentile(lat_ms, 50)),
            "p99_ms": float(np.percentile(lat_ms, 99)), "verified": verified,
            "roofline": {"kernel": "k_sc_pass (two rounds per pass: folds on read + message + next message "
                                   "as a polynomial in the challenge); 2^20 is latency-bound, the host transcript "
                                   "(20 rounds x ~20 us) dominates", "bound": "hbm", "achieved": gbs,
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                         "avg_ms": ms / max(n, 1), "launches": n}}


def bench_ajtai(args, M):
    """int8-matrix Ajtai commit: 2^22-element witness(es), 256-row ChaCha8 matrix, INT8 MFMA."""
    from morpheus_amd.ajtai import AjtaiI8, key_from_seed
    nv = args.nv or 22
    n, rows, batch = 1 << nv, 256, args.batch
    lib = M.lib
    t0 = time.perf_counter()
    A = AjtaiI8(key_from_seed(2024), rows, n)
    setup_s = time.perf_counter() - t0
    z = np.concatenate([gen_inputs(nv, seed=10 + w) for w in range(batch)])
    dz, dout = M.DeviceBuffer.from_numpy(z), M.DeviceBuffer(batch * rows)
    for _ in range(args.warmup):
        A.commit_device(dz, batch, dout)
    M.synchronize()
    lib.mam_profile_enable(1)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        A.commit_device(dz, batch, dout)
        lat.append(time.perf_counter() - s0)
    M.synchronize()
    elapsed = time.perf_counter() - t0
    prof = profile_collect(lib)
    lib.mam_profile_enable(0)
    ms, cnt, by = prof[5]
    avg = ms / max(cnt, 1)
    gbs = by / max(cnt, 1) / (avg / 1e3) / 1e9 if ms > 0 else 0.0
    mfma_tops = 2.0 * rows * n * 8 * batch / (avg / 1e3) / 1e12 if ms > 0 else 0.0
    lat_ms = np.array(lat) * 1e3
    # sampled rows of the commitment against exact integer products with the set-up's matrix
    out_words = dout.to_numpy().reshape(batch, rows)
    verified = True
    m32 = np.uint64(0xFFFFFFFF)
    for r in (0, 1, rows // 2, rows - 1):
        mrow = A.matrix_row(r, 0, n).astype(np.int64)
        for w in range(batch):
            zw = z[w * n:(w + 1) * n] % np.uint64(M.P)
            lo = int(np.dot(mrow, (zw & m32).astype(np.int64)))   # |term| < 2^39, sum < 2^61
            hi = int(np.dot(mrow, (zw >> np.uint64(32)).astype(np.int64)))
            verified &= (lo + (hi << 32)) % M.P == int(out_words[w, r])
    return {"metric": METRIC, "value": args.steps * batch * n / elapsed, "unit": "elems/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int8->int32 (u64 out)",
            "data": "synthetic",
            "config": {"workload": "ajtai_int8_commit", "witness_len": n, "rows": rows, "batch": batch,
                       "limbs": 8, "matrix": "ChaCha8 int8, expanded once", "parallelism": "single-gpu"},
            "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
            "setup_s": setup_s, "verified": bool(verified),
            "roofline": {"kernel": "k_aj_gemm (v_mfma_i32_16x16x64_i8, LDS-staged, fused limb split)",
                         "bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS,
                         "traffic": AJTAI_GEMM_TRAFFIC_BYTES if (nv == 22 and batch == 1) else None,
                         "avg_launch_ms": avg, "mfma_int8_tops": mfma_tops, "mfma_int8_frac_of_5pops": mfma_tops / 5000.0}}


def bench_hash(args, M):
    """Raw permutation throughput, one permutation per lane with every lane active:
    Monolith-w12 (mam_monolith_permute_batch) and Poseidon2-w8 (poseidon2_hash_batch_ffi)
    over 2^nv states resident in HBM."""
    nv = args.nv or 22
    n = 1 << nv
    lib = M.lib
    rng = np.random.default_rng(6)
    d12 = M.DeviceBuffer.from_numpy(rng.integers(0, M.P, 12 * n, dtype=np.uint64))
    d8 = M.DeviceBuffer.from_numpy(rng.integers(0, M.P, 8 * n, dtype=np.uint64))
    o8 = M.DeviceBuffer(8 * n)
    res = {}
    for name, call in (("monolith_w12", lambda: lib.mam_monolith_permute_batch(d12.ptr, n)),
                       ("poseidon2_w8", lambda: lib.poseidon2_hash_batch_ffi(d8.ptr, o8.ptr, n))):
        for _ in range(args.warmup):
            M.check(call(), name)
        M.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            M.check(call(), name)
        M.synchronize()
        res[name] = args.steps * n / (time.perf_counter() - t0)
    ach = res["monolith_w12"] * ALG_OPS_PER_PERM / 1e12
    return {"metric": METRIC, "value": res["monolith_w12"], "unit": "perms/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": n / res["monolith_w12"] * 1e3, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "u64 (Goldilocks)", "data": "synthetic",
            "config": {"workload": "hash_permutations", "states": n, "hashes": ["monolith_w12", "poseidon2_w8"],
                       "parallelism": "single-gpu"},
            "perms_per_s": res,
            "roofline": {"kernel": "k_mono_permute_n (one Monolith permutation per lane)", "bound": "valu",
                         "achieved": ach, "peak": VALU_NOMINAL_TOPS, "unit": "Tops/s", "frac": ach / VALU_NOMINAL_TOPS,
                         "traffic": None, "alg_ops_per_perm": ALG_OPS_PER_PERM,
                         "executed_valu_per_perm": MONO_PERM_VALU, "valu_probe_tops": dict(VALU_PROBE)}}


def merkle_perms(nv, lr):
    """Monolith permutations executed by one commit plus one opening. The commitment's
    codeword is 2^lr copies of one block, so its tree hashes one copy's subtree
    (2^nv - 1 permutations) and lr levels above it; the opening's folded trees hash in full
    (sum of flen - 1)."""
    total = (1 << nv) - 1 + lr
    for r in range(nv - 1):
        total += (1 << (nv + lr - 1 - r)) - 1
    return total


def verify_batch_instance(M, o, fv, gv, label, table, lr, nq):
    """Host checks of one batched instance: the sumcheck verifier accepts the proof for the
    claim sum f g, f(r) and g(r) are its final evaluations, and the Basefold verifier accepts the
    opening of f at r (evaluation f(r)) on the transcript the sumcheck leaves."""
    from morpheus_amd.hostfield import e2mul, gl_dot_np, mle_eval
    nv = int(fv.size).bit_length() - 1
    t = M.Transcript(label)
    fe = [tuple(int(x) for x in o["final_evals"][i]) for i in range(2)]
    ok, chs = M.sumcheck_verify(t, (gl_dot_np(fv, gv), 0), 2, nv, o["round_msgs"], e2mul(fe[0], fe[1]))
    r = np.array(chs, dtype=np.uint64).ravel()
    ok = ok and fe == [mle_eval(fv, r), mle_eval(gv, r)] and tuple(int(x) for x in o["proof"][:2]) == fe[0]
    return bool(ok and M.verify_ext2(o["root"], r, nv, lr, 0, o["proof"], nq, table, t))


def bench_batch(args, M):
    """64 sumcheck + Basefold instances of log-uniform size 2^12..2^22, proved concurrently by
    persistent worker streams (count autotuned) against one-at-a-time proving."""
    from morpheus_amd import workloads as W
    from morpheus_amd.batch import Instance, autotune_graph, autotune_workers, graph_phase_ms, prove_batch, \
        prove_batch_graph
    lr, nq = 1, args.queries
    data = W.batch_instance_data(args.instances)
    sizes, nvs = W.batch_sizes(args.instances)
    insts = [Instance(f, g, label) for f, g, label in data]
    tnv = int(nvs.max())
    table = M.BasefoldTable.generate(tnv, lr, tnv, 42)
    table.upload()
    total = int(sizes.sum())
    perms = sum(merkle_perms(int(v), lr) for v in nvs)

    # proofs land in page-locked buffers allocated once (one set per form, compared below)
    def pinned_set():
        return [M.PinnedArray(M.proof_len(int(v), lr, nq)) for v in nvs]
    pb_seq, pb_bat, pb_g = pinned_set(), pinned_set(), pinned_set()

    def run(w, pb):
        t0 = time.perf_counter()
        out = prove_batch(insts, table, lr, 0, nq, workers=w, proof_bufs=[p.a for p in pb])
        return time.perf_counter() - t0, out

    # worker count autotuned on every fourth instance (untimed)
    cands = (args.workers,) if args.workers else (1, 2, 4, 8, 12, 16)
    best, tune = autotune_workers(insts, table, lr, 0, nq, candidates=cands)
    for _ in range(max(args.warmup, 1)):
        run(best, pb_bat)
        run(1, pb_seq)
    seq_t, seq_lat, bat_t, bat_lat = [], [], [], []
    seq_out = bat_out = None
    for _ in range(args.steps):
        dt, seq_out = run(1, pb_seq)
        seq_t.append(dt)
        seq_lat += [o["end_ms"] - o["start_ms"] for o in seq_out]
        dt, bat_out = run(best, pb_bat)
        bat_t.append(dt)
        bat_lat += [o["end_ms"] - o["start_ms"] for o in bat_out]
    seq_s, bat_s = float(np.mean(seq_t)), float(np.mean(bat_t))
    # graph form: settings autotuned on the whole batch (each setting captures its graph once)
    gset, gtune = autotune_graph(insts, table, lr, 0, nq, branches=(0, 2, 4), max_wgs=(64, 256),
                                 workers=(8, 12, 16), repeats=2)
    g_t, g_lat, g_phase = [], [], []
    g_out = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        g_out = prove_batch_graph(insts, table, lr, 0, nq, gset["workers"], gset["branches"], gset["max_wg"],
                                  proof_bufs=[p.a for p in pb_g])
        g_t.append(time.perf_counter() - t0)
        g_lat += [o["end_ms"] - o["start_ms"] for o in g_out]
        g_phase.append(graph_phase_ms())
    g_s = float(np.mean(g_t))
    # verification: batched proofs (both forms) equal one-at-a-time proofs word for word, and
    # the host verifiers accept the largest, the smallest and two middle instances
    keys = ("round_msgs", "challenges", "final_evals", "root", "proof")
    same = all(np.array_equal(a[k], b[k]) for a, b in zip(seq_out, bat_out) for k in keys)
    g_same = all(np.array_equal(a[k], b[k]) for a, b in zip(seq_out, g_out) for k in keys)
    graph_rec = {"value": total / g_s, "ms_per_batch": g_s * 1e3, "proofs_per_s": len(insts) / g_s,
                 "p50_ms": float(np.percentile(g_lat, 50)), "p99_ms": float(np.percentile(g_lat, 99)),
                 "all_ready_ms": float(np.mean([p[0] for p in g_phase])),
                 "note": "openings start per instance as its commitment and sumcheck finish inside the graph; "
                         "all_ready_ms is when the last instance was ready",
                 "settings": gset, "autotune_s": {f"b{k[0]}_wg{k[1]}_w{k[2]}": v for k, v in gtune.items()},
                 "equals_one_at_a_time": bool(g_same)}
    # the graph form is the headline only when it is clearly faster (its p99 is the batch's end)
    form = "graph" if g_s < 0.97 * bat_s else "threads"
    best_s = min(g_s, bat_s)
    best_lat = g_lat if form == "graph" else bat_lat
    order = np.argsort(sizes)
    sample = sorted({int(order[-1]), int(order[0]), int(order[len(order) // 2]), int(order[len(order) // 4])})
    host_ok = all(verify_batch_instance(M, bat_out[i], data[i][0], data[i][1], data[i][2], table, lr, nq)
                  for i in sample)
    ach = perms * ALG_OPS_PER_PERM / best_s / 1e12
    return {"metric": METRIC, "value": total / best_s, "unit": "elems/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": best_s * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64 (Goldilocks, Ext2 challenges)", "data": "synthetic",
            "config": {"workload": "batched_sumcheck_basefold", "instances": len(insts), "total_elems": total,
                       "log2_size": "U[12,22] seeded, zero-padded to 2^ceil", "num_vars": [int(v) for v in nvs],
                       "log_rate": lr, "num_queries": nq, "merkle_hash": "monolith-w12", "workers": best,
                       "autotune_s": {str(k): v for k, v in tune.items()}, "parallelism": "single-gpu"},
            "form": form,
            "proofs_per_s": len(insts) / best_s,
            "verified": bool(same and g_same and host_ok),
            "verification": {"batched_equals_one_at_a_time": bool(same), "graph_equals_one_at_a_time": bool(g_same),
                             "host_verifiers_accept": bool(host_ok), "host_checked_instances": sample},
            "threads": {"value": total / bat_s, "ms_per_batch": bat_s * 1e3, "proofs_per_s": len(insts) / bat_s,
                        "p50_ms": float(np.percentile(bat_lat, 50)), "p99_ms": float(np.percentile(bat_lat, 99)),
                        "workers": best},
            "graph": graph_rec,
            "sequential": {"value": total / seq_s, "ms_per_batch": seq_s * 1e3, "proofs_per_s": len(insts) / seq_s,
                           "p50_ms": float(np.percentile(seq_lat, 50)), "p99_ms": float(np.percentile(seq_lat, 99))},
            "speedup_vs_sequential": seq_s / best_s,
            "p50_ms": float(np.percentile(best_lat, 50)), "p99_ms": float(np.percentile(best_lat, 99)),
            "roofline": {"kernel": "all Monolith Merkle work of the batch over its wall time", "bound": "valu",
                         "achieved": ach, "peak": VALU_NOMINAL_TOPS, "unit": "Tops/s", "frac": ach / VALU_NOMINAL_TOPS,
                         "traffic": None, "perms": perms, "alg_ops_per_perm": ALG_OPS_PER_PERM}}


def batch_open_claims(polys):
    """Claims of the batch-opening workload: poly i at point i % 2, and poly 0 at both points."""
    return [(i, i % 2) for i in range(polys)] + [(0, 1)]


def bench_batch_open(args, M):
    """Batch opening (batch_open_ext2): `polys` committed polynomials of 2^nv evaluations opened
    at 2 Ext2 points for polys + 1 claims in one proof, against one open_ext2 per claim."""
    from morpheus_amd.prover import EvaluationExt2, batch_open_ext2, batch_verify_ext2
    nv, lr, nq, k = args.nv or 20, 1, args.queries, args.polys
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    rng = np.random.default_rng(17)
    polys = [gen_inputs(nv, seed=100 + i) for i in range(k)]
    pts = rng.integers(0, M.P, (2, nv, 2), dtype=np.uint64)
    comms = [M.BasefoldCommitment.commit(f, nv, lr, M.HASH_MONOLITH) for f in polys]
    table = M.BasefoldTable.generate(nv, lr, nv, 42)
    table.upload()
    claims = batch_open_claims(k)
    # claimed values: the batch verifier's sum check needs the true evaluations
    vals = {}
    for pi, qi in claims:
        d = M.DeviceBuffer.from_numpy(polys[pi])
        dr = M.DeviceBuffer.from_numpy(pts[qi].ravel())
        out = M.DeviceBuffer(1 << nv)  # the first layer writes 2^(nv-1) Ext2 entries
        M.check(M.lib.partial_eval_ext2_from_gl_ffi(d.ptr, out.ptr, dr.ptr, nv, nv), "partial_eval")
        vals[(pi, qi)] = out.to_numpy(2)
    evals = [EvaluationExt2(pi, qi, vals[(pi, qi)]) for pi, qi in claims]

    def batch():
        return batch_open_ext2(comms, pts, evals, table, M.Transcript(b"bench"), nq)

    def singles():
        for pi, qi in claims:
            comms[pi].open_ext2(pts[qi], table, M.Transcript(b"bench"), nq)

    for _ in range(max(args.warmup, 1)):
        proof = batch()
        singles()
    bt, st = [], []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        proof = batch()
        bt.append(time.perf_counter() - t0)
        t0 = time.perf_counter()
        singles()
        st.append(time.perf_counter() - t0)
    roots = np.array([c.root for c in comms])
    ok = batch_verify_ext2(roots, pts, evals, proof, table, M.Transcript(b"bench"), nv, lr, nq)
    b_s, s_s = float(np.mean(bt)), float(np.mean(st))
    opened = k * (1 << nv)
    perms = merkle_perms(nv, lr)  # the combined codeword's tree and the folded trees
    ach = perms * ALG_OPS_PER_PERM / b_s / 1e12
    return {"metric": METRIC, "value": opened / b_s, "unit": "elems/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": b_s * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64 (Goldilocks, Ext2 challenges)", "data": "synthetic",
            "config": {"workload": "basefold_batch_open", "num_vars": nv, "log_rate": lr, "polys": k, "points": 2,
                       "claims": len(claims), "num_queries": nq, "merkle_hash": "monolith-w12",
                       "parallelism": "single-gpu"},
            "verified": ok, "proof_words": int(proof.size),
            "one_open_per_claim": {"ms": s_s * 1e3, "value": opened / s_s},
            "speedup_vs_one_open_per_claim": s_s / b_s,
            "roofline": {"kernel": "Monolith Merkle work of the batch opening over its wall time", "bound": "valu",
                         "achieved": ach, "peak": VALU_NOMINAL_TOPS, "unit": "Tops/s",
                         "frac": ach / VALU_NOMINAL_TOPS, "traffic": None, "perms": perms,
                         "alg_ops_per_perm": ALG_OPS_PER_PERM}}


def cpu_baseline_batch_open(nv, polys, nq):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    rng = np.random.default_rng(17)
    fs = [gen_inputs(nv, seed=100 + i) for i in range(polys)]
    comms = []
    for f in fs:
        cw, bh, tree = O.basefold_commit(f, nv, 1, O.MONOLITH)
        comms.append((cw, bh, tree[-4:]))
    pts = rng.integers(0, O.P, (2, nv, 2), dtype=np.uint64)
    evals = []
    for pi, qi in batch_open_claims(polys):
        e = np.zeros((1 << nv, 2), dtype=np.uint64)
        e[:, 0] = fs[pi]
        evals.append((pi, qi, O.partial_eval_ext2(e, nv, pts[qi].ravel(), nv).ravel()[:2]))
    table = O.basefold_table(nv, 1, nv, 42)
    t0 = time.perf_counter()
    O.basefold_batch_open_ext2(comms, nv, 1, O.MONOLITH, pts, evals, table, nv, O.Transcript(b"bench"), nq)
    dt = time.perf_counter() - t0
    return {"value": polys * (1 << nv) / dt, "unit": "elems/s", "cores": 1, "kind": "port",
            "sample": f"oracle batch_open_ext2 of {polys} polys x 2^{nv} evaluations, {polys + 1} claims at 2 points, "
                      f"rate 1/2, Monolith, {nq} queries, single thread, {dt:.2f} s"}


def bench_serve(args, M):
    """Matmul proving service: requests C = A B with m, k, n uniform over {256, ..., 4096},
    int8 A, B lifted into Goldilocks, C resident with them in HBM; each proof is the Basefold
    commits of A, B, C, an int8 Ajtai commit of C (256 rows), the degree-3 sumcheck over k and
    three openings. Poisson arrivals at a fixed mean rate; continuous batching on W worker
    streams with HBM-footprint admission, against the same trace proved one at a time."""
    from morpheus_amd import workloads as W
    from morpheus_amd.ajtai import AjtaiI8, key_from_seed
    from morpheus_amd.matmul import MatmulJob, serve, verify_matmul
    lr, nq = 1, args.queries
    dims = W.serve_trace(args.instances)
    jobs, elems, perms = [], 0, 0
    for i, ((m, k, n), (a, b)) in enumerate(zip(dims, W.matmul_operands(dims, seed=9))):
        jobs.append(MatmulJob.from_int8(a, b, b"c5-%d" % i))
        lm, lk, ln = (int(x).bit_length() - 1 for x in (m, k, n))
        elems += int(m * k + k * n + m * n)
        perms += sum(merkle_perms(nv, lr) for nv in (lm + lk, lk + ln, lm + ln))
    max_nv = int(max(int(m * n).bit_length() - 1 for m, _, n in dims))
    tnv = int(max(max(int(m * k).bit_length(), int(k * n).bit_length(), int(m * n).bit_length()) - 1
                  for m, k, n in dims))
    table = M.BasefoldTable.generate(tnv, lr, tnv, 42)
    table.upload()
    aj = AjtaiI8(key_from_seed(2024), 256, 1 << max_nv)
    budget = args.hbm_budget_gib << 30

    def run(w, arrivals, sub=None):
        js = jobs if sub is None else [jobs[i] for i in sub]
        arr = arrivals if sub is None else arrivals[sub]
        t0 = time.perf_counter()
        outs, admit, done = serve(js, arr, table, lr, 0, nq, aj, w, budget)
        return time.perf_counter() - t0, admit, done, outs

    zero = np.zeros(len(jobs))
    # worker count autotuned on every fourth request, all arriving at once (untimed)
    cands = [args.workers] if args.workers else [2, 4, 8]
    sub = np.arange(0, len(jobs), 4)
    run(cands[-1], zero, sub)
    tune = {}
    for w in cands:
        tune[w] = min(run(w, zero, sub)[0] for _ in range(2))
    best = min(tune, key=tune.get)
    run(1, zero)
    # saturation: the whole trace at t = 0
    sat_seq = min(run(1, zero)[0] for _ in range(max(args.warmup, 1)))
    sat_srv = min(run(best, zero)[0] for _ in range(max(args.warmup, 1)))
    rate = args.rate or 1.25 * len(jobs) / sat_seq
    arrivals = np.cumsum(np.random.default_rng(10).exponential(1e3 / rate, len(jobs)))
    arrivals -= arrivals[0]
    res, last = {}, {}
    for name, w in (("sequential", 1), ("served", best)):
        spans, lat = [], []
        for _ in range(args.steps):
            dt, admit, done, outs = run(w, arrivals)
            spans.append(float(done.max()))
            lat += list(done - arrivals)
        last[name] = outs
        span = float(np.mean(spans)) / 1e3
        res[name] = {"value": elems / span, "proofs_per_s": len(jobs) / span, "ms_per_trace": span * 1e3,
                     "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99))}
    # verification: each served result equals the one-at-a-time result, and the host verifier
    # accepts the largest request's proof
    same = all(np.array_equal(a[k], b[k]) for a, b in zip(last["sequential"], last["served"]) for k in a)
    big = int(np.argmax([int(m) * int(k) * int(n) for m, k, n in dims]))
    jb = jobs[big]
    host_ok = verify_matmul(last["served"][big], jb.lm, jb.lk, jb.ln, jb.label, table, lr, 0, nq)
    srv = res["served"]
    ach = perms * ALG_OPS_PER_PERM / (srv["ms_per_trace"] / 1e3) / 1e12
    return {"metric": METRIC, "value": srv["value"], "unit": "elems/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": srv["ms_per_trace"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64 (Goldilocks, Ext2 challenges)", "data": "synthetic",
            "config": {"workload": "matmul_proving_service", "requests": len(jobs),
                       "dims": "m, k, n uniform over {256, 512, 1024, 2048, 4096} (seeded)",
                       "entries": "int8 uniform lifted to p", "proved_elems_per_trace": elems,
                       "arrivals": f"Poisson, mean rate {rate:.1f} proofs/s (1.25 x one-at-a-time capacity)"
                       if not args.rate else f"Poisson, mean rate {rate:.1f} proofs/s",
                       "ajtai": "int8, 256 rows, prefix of a 2^%d-column set-up" % max_nv, "log_rate": lr,
                       "num_queries": nq, "merkle_hash": "monolith-w12", "workers": best,
                       "hbm_budget_gib": args.hbm_budget_gib, "autotune_s": {str(k): v for k, v in tune.items()},
                       "parallelism": "single-gpu"},
            "proofs_per_s": srv["proofs_per_s"], "p50_ms": srv["p50_ms"], "p99_ms": srv["p99_ms"],
            "verified": bool(same and host_ok),
            "verification": {"served_equals_one_at_a_time": bool(same), "host_verifier_accepts_largest": bool(host_ok),
                             "largest_request": [int(x) for x in dims[big]]},
            "sequential": res["sequential"],
            "saturated": {"sequential_proofs_per_s": len(jobs) / sat_seq, "served_proofs_per_s": len(jobs) / sat_srv,
                          "speedup": sat_seq / sat_srv},
            "roofline": {"kernel": "all Monolith Merkle work of the trace over its span", "bound": "valu",
                         "achieved": ach, "peak": VALU_NOMINAL_TOPS, "unit": "Tops/s", "frac": ach / VALU_NOMINAL_TOPS,
                         "traffic": None, "perms": perms, "alg_ops_per_perm": ALG_OPS_PER_PERM}}


def cpu_baseline_serve(lg):
    """The oracle proving one matmul request (m = k = n = 2^lg) end to end, on every core and on
    one."""
    O = oracle()
    f = cpu_facts()
    from morpheus_amd.ajtai import key_from_seed
    rng = np.random.default_rng(9)
    n = 1 << lg
    a = rng.integers(-128, 128, (n, n))
    b = rng.integers(-128, 128, (n, n))
    table = O.basefold_table(2 * lg, 1, 2 * lg, 42)
    out = {}
    for name, thr in (("all", f["threads_all"]), ("single", 1)):
        out[name], _ = timed_threads(O, thr, lambda: O.matmul_prove(a, b, b"c5-cpu", table, 2 * lg, 1, 100,
                                                                    key_from_seed(2024), 256))
    return {"value": 3 * n * n / out["all"], "unit": "elems/s", "cores": f["threads_all"], "kind": "port",
            "sample": f"oracle matmul proof, m = k = n = {n} (A, B, C Basefold commit + open, int8 Ajtai of C, "
                      f"sumcheck) on {f['threads_all']} threads, {out['all']:.2f} s ({1 / out['all']:.3f} proofs/s "
                      f"at this size)",
            "single_thread": {"value": 3 * n * n / out["single"], "unit": "elems/s", "cores": 1,
                              "sample": f"the same on one thread, {out['single']:.2f} s"},
            "cpu_model": f["cpu_model"], "nproc": f["nproc"]}


RING_KEY = bytes((7 * i + 3) % 256 for i in range(32))


def bench_ring_ajtai(args, M):
    """Ring Ajtai commit over F_q[X]/(X^64+1), 42 rows: 2^nv binary coefficients per witness
    (2^(nv-6) ring columns, density 1/2), B witnesses per launch, matrix regenerated by ChaCha8."""
    from morpheus_amd import ring_ajtai as R
    nv = args.nv or 22
    n, b = 1 << (nv - 6), args.batch
    lib = M.lib
    z = np.random.default_rng(13).integers(0, 1 << 64, b * n, dtype=np.uint64)
    seed = R.Seed.from_bytes(RING_KEY)
    dk, dz, dout = R._key_buffer(seed), M.DeviceBuffer.from_numpy(z), M.DeviceBuffer(b * R.KAPPA * R.RING_DIM)
    chunk = R.pick_default_chunk(n, b)
    for _ in range(args.warmup):
        M.check(lib.ajtai_commit_dense_batched_ffi(dk.ptr, dz.ptr, n, b, chunk, dout.ptr), "ring commit")
    lib.mam_profile_enable(1)
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s0 = time.perf_counter()
        M.check(lib.ajtai_commit_dense_batched_ffi(dk.ptr, dz.ptr, n, b, chunk, dout.ptr), "ring commit")
        lat.append(time.perf_counter() - s0)
    elapsed = time.perf_counter() - t0
    prof = profile_collect(lib)
    lib.mam_profile_enable(0)
    ms, cnt, _ = prof[6]
    avg = ms / max(cnt, 1)
    blocks = 42 * n * 8 * ((b + 3) // 4 if b >= 4 else 1)  # ChaCha8 blocks per launch
    set_bits = int(sum(bin(int(w)).count("1") for w in z))
    lat_ms = np.array(lat) * 1e3
    # VALU issue slots per launch: each 64-lane wave issues RING_PRG_VALU instructions per
    # ChaCha8 column fill (56 blocks) and RING_BIT_VALU per set bit per 7-row group.
    fills = 6 * n * ((b + 3) // 4 if b >= 4 else 1)
    slots = (fills * RING_PRG_VALU + set_bits * 6 * RING_BIT_VALU) * 64
    ach = slots / (avg / 1e3) / 1e12 if avg > 0 else 0.0
    return {"metric": METRIC, "value": args.steps * b * (1 << nv) / elapsed, "unit": "elems/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": elapsed / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64 (mod q)",
            "data": "synthetic",
            "config": {"workload": "ring_ajtai_commit", "binary_coeffs": 1 << nv, "ring_columns": n, "kappa": 42,
                       "ring_dim": 64, "batch": b, "chunk": chunk, "matrix": "ChaCha8, regenerated per launch",
                       "parallelism": "single-gpu"},
            "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
            "roofline": {"kernel": "k_ring_dense (ChaCha8 + LDS negacyclic rows)", "bound": "valu",
                         "achieved": ach, "peak": VALU_NOMINAL_TOPS, "unit": "Tops/s",
                         "frac": ach / VALU_NOMINAL_TOPS, "achieved_from": "wave instructions of its gfx950 ISA x 64",
                         "traffic": None, "avg_launch_ms": avg, "chacha_blocks_per_s": blocks / (avg / 1e3),
                         "set_bits": set_bits}}


SHARED_KEYS = ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data")


def bench_all(args, M):
    """The headline (Basefold commit + open at 2^22) and, in the same line, one record per other
    benchmark configuration: the 2^20 product sumcheck, the int8 Ajtai commit at 2^22 x 256, the
    64-instance batch and the matmul proving service."""
    from morpheus_amd import workloads as W

    def sub(**kw):
        return argparse.Namespace(**{**vars(args), **kw})
    out = bench_basefold(sub(nv=22), M)
    c1 = bench_sumcheck(sub(nv=20, steps=max(args.steps, 20), warmup=max(args.warmup, 2)), M)
    c3 = bench_ajtai(sub(nv=22, batch=1, steps=max(args.steps, 20), warmup=max(args.warmup, 2)), M)
    c4 = bench_batch(sub(instances=64, steps=min(max(args.steps, 3), 5), warmup=1, workers=0), M)
    c5 = bench_serve(sub(instances=64, steps=min(max(args.steps, 2), 3), warmup=1, workers=0, rate=0.0), M)
    c2p = bench_basefold(sub(nv=22, hash="poseidon2", steps=min(max(args.steps, 3), 10), warmup=1), M)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_basefold_all(22, 22, 1, args.queries)
        c1["cpu_baseline"] = cpu_sumcheck_all(20)
        c3["cpu_baseline"] = cpu_ajtai_all(22)
        c4["cpu_baseline"] = cpu_batch_sample(W.batch_instance_data(64), 8, 1, args.queries)
        c5["cpu_baseline"] = cpu_baseline_serve(8)
    configs = {}
    for name, rec in (("sumcheck_2p20", c1), ("ajtai_int8_2p22", c3), ("batched_64", c4), ("serve_matmul", c5),
                      ("basefold_2p22_poseidon2_trees", c2p)):
        configs[name] = {k: v for k, v in rec.items() if k not in SHARED_KEYS}
    out["configs"] = configs
    out["served"] = {"proofs_per_s": c5["proofs_per_s"], "p50_ms": c5["p50_ms"], "p99_ms": c5["p99_ms"],
                     "one_at_a_time": {k: c5["sequential"][k] for k in ("proofs_per_s", "p50_ms", "p99_ms")},
                     "verified": c5["verified"]}
    out["verified_all"] = bool(out["verified"] and all(r.get("verified") for r in (c1, c3, c4, c5, c2p)))
    return out


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}, WORLD_SIZE {world}); "
              "this build has no multi-GPU path", flush=True)
        sys.exit(2)
    import morpheus_amd as M
    if M.lib.goldilocks_cuda_init() != 0:
        print("no usable GPU", file=sys.stderr)
        sys.exit(1)
    valu_peak(M)
    if args.copy_sweep:
        for mib in (1, 2, 4, 8, 16, 32, 64, 96, 128, 256, 512):
            ms = ctypes.c_double()
            M.check(M.lib.mam_copy_probe(mib << 20, 20, ctypes.byref(ms)), "mam_copy_probe")
            moved = 2 * (mib << 20)
            print(json.dumps({"copy_mib_each_way": mib, "bytes_moved": moved, "ms": ms.value,
                              "gb_s": moved / ms.value / 1e6, "frac_of_8tb": moved / ms.value / 1e6 / 8000.0}),
                  flush=True)
        return
    if args.calibrate_copy:
        n = 1 << 25
        a, b = M.DeviceBuffer(n), M.DeviceBuffer(n)
        M.check(M.lib.gl_memset_zero(a.ptr, n), "zero")
        for _ in range(2):
            t0 = time.perf_counter()
            M.check(M.lib.cuda_memcpy_dtod(b.ptr, a.ptr, n * 8), "copy")
            dt = time.perf_counter() - t0
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "wall_ms": dt * 1e3,
                          "valu_probe_tops": VALU_PROBE}), flush=True)
        return
    if args.workload == "all":
        out = bench_all(args, M)
    elif args.workload == "basefold":
        out = bench_basefold(args, M)
    elif args.workload == "sumcheck":
        out = bench_sumcheck(args, M)
    elif args.workload == "ajtai":
        out = bench_ajtai(args, M)
    elif args.workload == "batch":
        out = bench_batch(args, M)
    elif args.workload == "hash":
        out = bench_hash(args, M)
    elif args.workload == "serve":
        out = bench_serve(args, M)
    elif args.workload == "batch_open":
        out = bench_batch_open(args, M)
    else:
        out = bench_ring_ajtai(args, M)
    if not args.no_cpu and args.workload != "all":
        if args.workload == "basefold":
            out["cpu_baseline"] = cpu_basefold_all(args.cpu_sample_nv or 22, 20, 1, args.queries)
        elif args.workload == "sumcheck":
            out["cpu_baseline"] = cpu_baseline_sumcheck(args.cpu_sample_nv or 20)
        elif args.workload == "ajtai":
            out["cpu_baseline"] = cpu_baseline_ajtai(args.cpu_sample_nv or 16)
        elif args.workload == "batch":
            out["cpu_baseline"] = cpu_baseline_basefold(args.cpu_sample_nv or 18, 1, args.queries)
        elif args.workload == "hash":
            out["cpu_baseline"] = cpu_baseline_hash(args.cpu_sample_nv or 18)
        elif args.workload == "serve":
            out["cpu_baseline"] = cpu_baseline_serve(args.cpu_sample_nv or 8)
        elif args.workload == "batch_open":
            out["cpu_baseline"] = cpu_baseline_batch_open(args.cpu_sample_nv or 16, args.polys, args.queries)
        else:
            out["cpu_baseline"] = cpu_baseline_ring_ajtai(args.cpu_sample_nv or 13)
        out["cpu_baseline"].setdefault("cpu_model", cpu_facts()["cpu_model"])
        out["cpu_baseline"].setdefault("nproc", os.cpu_count())
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
