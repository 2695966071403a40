// Batched proving of independent instances of uneven size: each instance proves
// sum_x f(x) g(x) with the Ext2 product sumcheck and then opens a Basefold commitment of f
// at the sumcheck's challenge point, on one Fiat-Shamir transcript:
//
//   T = Transcript::new(label)
//   (round messages, r, final evals) = sumcheck_prove(T, [f, g])        mam_sumcheck_prove
//   root = BasefoldCommitment::commit(f)                                commit_into
//   proof = open_ext2(root, point = r, T, num_queries)                  mam_basefold_open_ext2
//
// (zk-torch-3/src/sumcheck/gpu_prover.rs:21-89 followed by goldilocks-cuda-rs/src/basefold.rs
// :382-414 and :578-863; the composition is this library's, the pieces are the reference's.)
//
// MI355X scheduling: W persistent host worker threads take instances from a shared
// counter, largest first. Every thread drives its own per-thread stream (hipStreamPerThread) with its own
// scratch arenas, pinned mailbox and commitment buffers, so the latency-bound rounds of small
// instances (one host round trip per sumcheck round and per Basefold fold) overlap with the
// bandwidth- and hash-bound kernels of large ones, and the CPU transcripts of different
// instances run on different cores. W = 1 is one-at-a-time proving on the same code path.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <string.h>
#include <thread>
#include <vector>

#include "basefold.h"
#include "common.h"
#include "merkle.h"
#include "profile.h"
#include "sumcheck.h"
#include "transcript.h"
#include "workers.h"

extern "C" {
int mam_sumcheck_prove(mam_transcript *t, const u64 *d_polys, int num_polys, int num_vars, int input_is_base,
                       u64 *round_msgs, u64 *challenges, u64 *final_evals, u64 *final_eval);
int mam_basefold_commit_into(const u64 *d_evals, int num_vars, int log_rate, int hash, u64 *d_codeword,
                             u64 *d_bh_evals, u64 *d_tree, u64 *root_out);
int mam_basefold_commitment_wrap(u64 *d_codeword, u64 *d_bh_evals, u64 *d_tree, const u64 *root4, int num_vars,
                                 int log_rate, int hash, void **handle);
int mam_basefold_commitment_free(void *handle);
long mam_basefold_proof_len(int num_vars, int log_rate, int num_queries);
int mam_basefold_open_ext2(void *handle, const u64 *point2, const u64 *d_table, int table_nv, mam_transcript *t,
                           int num_queries, u64 *proof);
}

// One instance: inputs on the device, outputs on the host (sizes from num_vars).
struct mam_batch_instance {
    const u64 *d_fg;       // f then g, 2^num_vars base words each
    int num_vars;
    const uint8_t *label;  // transcript label
    long label_len;
    u64 *round_msgs;       // [num_vars][3][2]
    u64 *challenges;       // [num_vars][2]
    u64 *final_evals;      // [2][2]
    u64 *root;             // [4]
    u64 *proof;            // mam_basefold_proof_len(num_vars, log_rate, num_queries) words
    double start_ms, end_ms;  // relative to the batch start
    int status;
};

namespace {

WorkerPool &pool() {
    static WorkerPool *p = new WorkerPool();
    return *p;
}

// Per-thread commitment buffers for instances of up to 2^nv evaluations at rate 2^-lr.
struct CommitBufs {
    u64 *cw = nullptr, *bh = nullptr, *tree = nullptr;
    int nv = -1, lr = -1;
    ~CommitBufs() { release(); }
    void release() {
        if (cw) (void)hipFree(cw);
        if (bh) (void)hipFree(bh);
        if (tree) (void)hipFree(tree);
        cw = bh = tree = nullptr;
        nv = lr = -1;
    }
    bool reserve(int want_nv, int want_lr) {
        if (want_nv <= nv && want_lr == lr) return true;
        (void)hipStreamSynchronize(mam_stream());
        release();
        const size_t cwl = (size_t)1 << (want_nv + want_lr);
        if (hipMalloc((void **)&cw, cwl * sizeof(u64)) != hipSuccess ||
            hipMalloc((void **)&bh, ((size_t)1 << want_nv) * sizeof(u64)) != hipSuccess ||
            hipMalloc((void **)&tree, (cwl - 1) * 4 * sizeof(u64)) != hipSuccess) {
            release();
            return false;
        }
        nv = want_nv;
        lr = want_lr;
        return true;
    }
};
thread_local CommitBufs t_bufs;

int prove_one(mam_batch_instance &I, const u64 *d_table, int table_nv, int lr, int hash, int nq, u64 *cw, u64 *bh,
              u64 *tree) {
    mam_transcript t;
    mam::tr_init(&t, I.label, I.label_len);
    u64 fe[2];
    int rc = mam_sumcheck_prove(&t, I.d_fg, 2, I.num_vars, 1, I.round_msgs, I.challenges, I.final_evals, fe);
    if (rc) return rc;
    rc = mam_basefold_commit_into(I.d_fg, I.num_vars, lr, hash, cw, bh, tree, I.root);
    if (rc) return rc;
    void *h = nullptr;
    rc = mam_basefold_commitment_wrap(cw, bh, tree, I.root, I.num_vars, lr, hash, &h);
    if (rc) return rc;
    rc = mam_basefold_open_ext2(h, I.challenges, d_table, table_nv, &t, nq, I.proof);
    mam_basefold_commitment_free(h);
    return rc;
}

}  // namespace

// Proves n instances with `workers` threads (1 = sequential). d_table: Basefold folding
// table generated for (table_nv, log_rate, table_nv), table_nv >= every num_vars.
// Returns 0 if every instance succeeded, else the first failing status (per-instance
// statuses are in inst[i].status).
MAM_API int mam_prove_batch(mam_batch_instance *inst, int n, const u64 *d_table, int table_nv, int log_rate,
                            int hash, int num_queries, int workers) {
    if (!inst || n < 0 || workers < 1 || workers > 64 || !d_table) return -2;
    int max_nv = 0;
    for (int i = 0; i < n; i++) {
        if (inst[i].num_vars < 2 || inst[i].num_vars > table_nv || !inst[i].d_fg) return -2;
        max_nv = std::max(max_nv, inst[i].num_vars);
    }
    // largest first, so the long instances start early and the small ones fill the gaps
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return inst[a].num_vars > inst[b].num_vars; });
    std::atomic<int> next(0);
    std::atomic<int> first_err(0);
    const auto t0 = std::chrono::steady_clock::now();
    auto ms_since = [&]() {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    auto work = [&]() {
        const bool ok = t_bufs.reserve(max_nv, log_rate);
        for (int k; (k = next.fetch_add(1)) < n;) {
            mam_batch_instance &I = inst[order[k]];
            I.start_ms = ms_since();
            I.status = ok ? prove_one(I, d_table, table_nv, log_rate, hash, num_queries, t_bufs.cw, t_bufs.bh,
                                      t_bufs.tree)
                          : -1;
            I.end_ms = ms_since();
            if (I.status) {
                int z = 0;
                first_err.compare_exchange_strong(z, I.status);
            }
        }
        (void)hipStreamSynchronize(mam_stream());
    };
    pool().run(workers, work);
    return first_err.load();
}

// ---------------------------------------------------------------------------
// Graph form of the batch. Phase 1 is one hipGraph launch that holds every instance's
// sumcheck and Basefold commitment: the sumchecks advance together with the transcript on
// the device (mam_bsc_launch_pass / _round: one launch pair per pass for the whole batch, each
// instance's workgroups sized by its own array), and the commitments (codeword transform,
// Merkle tree, root copy) run on up to `branches` side streams of the capture, which the
// graph runs concurrently with the sumchecks. Phase 2 opens every commitment at its
// sumcheck point on the worker pool, each transcript continuing from the device state. The
// graph and its buffers are kept for the next call with the same instances (sizes, input
// buffers, rate, hash, branches and workgroup cap), so a repeated batch costs one graph
// launch, two copies and the openings. The proofs are the per-instance ones word for word.
// ---------------------------------------------------------------------------
namespace {

// Readiness flags of the graph form: the value at *seq (the call's sequence number, uploaded
// with the transcripts) released to a pinned host word once the work before it on the stream
// has finished, so an opening can start while the graph still runs.
__global__ void k_ready_flag(const u64 *seq, u64 *host_flag) {
    if (threadIdx.x == 0) {
        const u64 v = *seq;
        __threadfence_system();
        __hip_atomic_store(host_flag, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

struct GraphBatch {
    std::vector<int> nv;
    std::vector<const u64 *> fg;
    int lr = -1, hash = -1, branches = -1, max_wg = -1;
    hipStream_t s = nullptr;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> ev;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void *mem = nullptr;  // every device buffer below, one allocation
    u64 *host = nullptr;  // pinned staging of the output blocks, then the call's sequence number
    u64 *flags = nullptr; // pinned: commitment ready [n], then sumcheck pass done [passes]
    u64 seq = 0;
    std::vector<int> last_pass;
    size_t out_words = 0;
    u64 *out = nullptr;   // per instance: tr 18 | ch 4 | msgs 6 nv | chs 2 nv | finals 4 | root 4
    std::vector<size_t> off;
    std::vector<u64 *> cw, bh, tree;

    static size_t block_words(int nv) { return 30 + 8 * (size_t)nv; }
    void release() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        if (mem) (void)hipFree(mem);
        if (host) (void)hipHostFree(host);
        if (flags) (void)hipHostFree(flags);
        mem = nullptr;
        host = nullptr;
        flags = nullptr;
        for (hipStream_t t : side) (void)hipStreamDestroy(t);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        if (s) (void)hipStreamDestroy(s);
        side.clear();
        ev.clear();
        s = nullptr;
        nv.clear();
        fg.clear();
    }
};

GraphBatch &graph_batch() {
    static GraphBatch *g = new GraphBatch();
    return *g;
}
std::mutex g_graph_mu;
double g_graph_phase_ms[2] = {0, 0};

size_t align64(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

// Builds buffers, job tables and the graph for this batch.
int graph_build(GraphBatch &G, const mam_batch_instance *inst, int n, int lr, int hash, int branches, int max_wg) {
    G.release();
    for (int i = 0; i < n; i++) {
        G.nv.push_back(inst[i].num_vars);
        G.fg.push_back(inst[i].d_fg);
    }
    G.lr = lr;
    G.hash = hash;
    G.branches = branches;
    G.max_wg = max_wg;
    // pass schedule per instance (as the whole-proof host prover's: two rounds per pass, the
    // last pass of an odd count alone)
    struct PassPlan {
        int round, base_in, folds, look, use_buf;  // use_buf: 0/1 output buffer, -1 none
        size_t items, in_len;                       // in_len: elements per polynomial of the input
        int in_buf;                                 // -1: the instance's input
        size_t in_stride_words, out_stride_words;
    };
    std::vector<std::vector<PassPlan>> plans(n);
    int max_passes = 0;
    for (int i = 0; i < n; i++) {
        const int nv = inst[i].num_vars;
        const size_t N = (size_t)1 << nv;
        const size_t stride[2] = {N, N / 2 > 1 ? N / 2 : 2};  // words per polynomial in buf0 / buf1
        int folds = 0, nxt = 0, src_buf = -1, src_base = 1;
        size_t size = N, src_stride = N;
        for (int r = 0; r < nv;) {
            PassPlan p;
            p.round = r;
            p.look = nv - r >= 2;
            p.folds = folds;
            p.base_in = src_base;
            const size_t m = size >> folds;
            p.items = m / (p.look ? 4 : 2);
            p.in_buf = src_buf;
            p.in_stride_words = src_stride;
            p.use_buf = folds ? nxt : -1;
            p.out_stride_words = folds ? stride[nxt] : 0;
            plans[i].push_back(p);
            if (folds) {
                src_buf = nxt;
                src_stride = stride[nxt];
                src_base = 0;
                size = m;
                nxt ^= 1;
            }
            folds = p.look ? 2 : 1;
            r += p.look ? 2 : 1;
        }
        max_passes = std::max(max_passes, (int)plans[i].size());
    }
    // device layout
    std::vector<size_t> o_buf0(n), o_buf1(n), o_cw(n), o_bh(n), o_tree(n);
    size_t bytes = 0;
    auto take = [&](size_t b) {
        const size_t o = bytes;
        bytes += align64(b);
        return o;
    };
    G.off.resize(n);
    G.out_words = 0;
    for (int i = 0; i < n; i++) {
        G.off[i] = G.out_words;
        G.out_words += GraphBatch::block_words(inst[i].num_vars);
    }
    const size_t o_out = take((G.out_words + 1) * 8);  // + the sequence word
    for (int i = 0; i < n; i++) {
        const size_t N = (size_t)1 << inst[i].num_vars, L = N << lr;
        o_buf0[i] = take(2 * N * 8);
        o_buf1[i] = take(2 * (N / 2 > 1 ? N / 2 : 2) * 8);
        o_cw[i] = take(L * 8);
        o_bh[i] = take(N * 8);
        o_tree[i] = take((L - 1) * 4 * 8);
    }
    // job tables: pass k holds the jobs of the instances with more than k passes
    std::vector<std::vector<BscJob>> jobs(max_passes);
    std::vector<unsigned> nblk_total(max_passes, 0);
    size_t max_blocks = 1;
    const size_t o_insts = take((size_t)n * sizeof(BscInst));
    size_t njobs_all = 0;
    for (int k = 0; k < max_passes; k++)
        for (int i = 0; i < n; i++)
            if (k < (int)plans[i].size()) njobs_all++;
    const size_t o_jobs = take(njobs_all * sizeof(BscJob));
    // partials: sized after the tables are known (taken last)
    std::vector<BscInst> insts(n);
    for (int k = 0; k < max_passes; k++) {
        unsigned blk = 0;
        for (int i = 0; i < n; i++) {
            if (k >= (int)plans[i].size()) continue;
            const PassPlan &p = plans[i][k];
            BscJob J = {};
            J.items = p.items;
            J.base_in = p.base_in;
            J.folds = p.folds;
            J.look = p.look;
            J.inst = i;
            J.round = p.round;
            J.last = k + 1 == (int)plans[i].size();
            J.in_stride = p.in_stride_words;
            J.out_stride = p.out_stride_words;
            size_t want = (p.items + MAM_BLOCK - 1) / MAM_BLOCK;
            if (want < 1) want = 1;
            if (want > (size_t)max_wg) want = (size_t)max_wg;
            J.blk0 = blk;
            J.nblk = (unsigned)want;
            blk += J.nblk;
            // pointers are filled in once the allocation exists (offsets for now)
            J.in = (const u64 *)(uintptr_t)(p.in_buf < 0 ? 0 : 1 + p.in_buf);
            J.out = (u64 *)(uintptr_t)(p.use_buf < 0 ? 0 : 1 + p.use_buf);
            jobs[k].push_back(J);
        }
        nblk_total[k] = blk;
        max_blocks = std::max(max_blocks, (size_t)blk);
    }
    const size_t o_partials = take(max_blocks * MAM_BSC_WORDS * 8);
    if (hipMalloc(&G.mem, bytes) != hipSuccess) {
        G.mem = nullptr;
        return -1;
    }
    if (hipHostMalloc((void **)&G.host, (G.out_words + 1) * 8, hipHostMallocDefault) != hipSuccess) {
        G.host = nullptr;
        return -1;
    }
    if (hipHostMalloc((void **)&G.flags, (n + max_passes) * 8, hipHostMallocDefault) != hipSuccess) {
        G.flags = nullptr;
        return -1;
    }
    memset(G.flags, 0, (n + max_passes) * 8);
    G.seq = 0;
    G.last_pass.resize(n);
    for (int i = 0; i < n; i++) G.last_pass[i] = (int)plans[i].size() - 1;
    char *base = (char *)G.mem;
    G.out = (u64 *)(base + o_out);
    G.cw.resize(n);
    G.bh.resize(n);
    G.tree.resize(n);
    for (int i = 0; i < n; i++) {
        G.cw[i] = (u64 *)(base + o_cw[i]);
        G.bh[i] = (u64 *)(base + o_bh[i]);
        G.tree[i] = (u64 *)(base + o_tree[i]);
        u64 *blk = G.out + G.off[i];
        const int nv = inst[i].num_vars;
        insts[i].tr = blk;
        insts[i].ch = blk + 18;
        insts[i].msgs = blk + 22;
        insts[i].chs = blk + 22 + 6 * (size_t)nv;
        insts[i].finals = blk + 22 + 8 * (size_t)nv;
    }
    // resolve the job pointers
    std::vector<BscJob> flat;
    std::vector<size_t> job_off(max_passes);
    for (int k = 0; k < max_passes; k++) {
        job_off[k] = flat.size();
        for (BscJob J : jobs[k]) {
            const int i = J.inst;
            u64 *bufs[2] = {(u64 *)(base + o_buf0[i]), (u64 *)(base + o_buf1[i])};
            const uintptr_t in_tag = (uintptr_t)J.in, out_tag = (uintptr_t)J.out;
            J.in = in_tag ? bufs[in_tag - 1] : inst[i].d_fg;
            J.out = out_tag ? bufs[out_tag - 1] : nullptr;
            if (J.last) {
                J.fin_src = J.folds ? J.out : J.in;
                J.fin_stride = J.folds ? J.out_stride : J.in_stride;
                J.fin_base = J.folds ? 0 : J.base_in;
                J.fin_folds = J.look ? 2 : 1;
            }
            flat.push_back(J);
        }
    }
    BscInst *d_insts = (BscInst *)(base + o_insts);
    BscJob *d_jobs = (BscJob *)(base + o_jobs);
    u64 *partials = (u64 *)(base + o_partials);
    if (hipStreamCreateWithFlags(&G.s, hipStreamNonBlocking) != hipSuccess) return -1;
    if (hipMemcpyAsync(d_insts, insts.data(), n * sizeof(BscInst), hipMemcpyHostToDevice, G.s) != hipSuccess ||
        hipMemcpyAsync(d_jobs, flat.data(), flat.size() * sizeof(BscJob), hipMemcpyHostToDevice, G.s) != hipSuccess ||
        hipStreamSynchronize(G.s) != hipSuccess)
        return -1;
    G.side.resize(branches);
    for (int b = 0; b < branches; b++)
        if (hipStreamCreateWithFlags(&G.side[b], hipStreamNonBlocking) != hipSuccess) return -1;
    G.ev.resize(1 + branches);
    for (auto &e : G.ev)
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return -1;
    // capture
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return inst[a].num_vars > inst[b].num_vars; });
    if (hipStreamBeginCapture(G.s, hipStreamCaptureModeThreadLocal) != hipSuccess) return -1;
    mam_prof_suspend(true);
    int rc = hipEventRecord(G.ev[0], G.s) == hipSuccess ? 0 : -1;
    for (int b = 0; b < branches && !rc; b++)
        if (hipStreamWaitEvent(G.side[b], G.ev[0], 0) != hipSuccess) rc = -1;
    // commitments, largest first, round-robin over the side streams (or after the sumchecks)
    auto commit = [&](int i, hipStream_t st) {
        const int nv = inst[i].num_vars;
        const size_t leaves = (size_t)1 << (nv + lr - 1);
        if (mam_commit_codeword(inst[i].d_fg, G.bh[i], G.cw[i], nv, lr, st)) return -1;
        if (mam_merkle_build_copies(G.cw[i], G.tree[i], leaves, hash, 0, 1 << lr, st)) return -1;
        u64 *root = G.out + G.off[i] + 26 + 8 * (size_t)nv;
        if (hipMemcpyAsync(root, G.tree[i] + mam_merkle_root_offset(leaves), 4 * 8, hipMemcpyDeviceToDevice, st) !=
            hipSuccess)
            return -1;
        k_ready_flag<<<1, 64, 0, st>>>(G.out + G.out_words, G.flags + i);
        return hipGetLastError() == hipSuccess ? 0 : -1;
    };
    if (branches > 0)
        for (int k = 0; k < n && !rc; k++) rc = commit(order[k], G.side[k % branches]);
    for (int k = 0; k < max_passes && !rc; k++) {
        rc = mam_bsc_launch_pass(d_jobs + job_off[k], (int)jobs[k].size(), nblk_total[k], d_insts, partials, G.s);
        if (!rc) rc = mam_bsc_launch_round(d_jobs + job_off[k], (int)jobs[k].size(), d_insts, partials, G.s);
        if (!rc) {
            k_ready_flag<<<1, 64, 0, G.s>>>(G.out + G.out_words, G.flags + n + k);
            if (hipGetLastError() != hipSuccess) rc = -1;
        }
    }
    if (branches == 0)
        for (int k = 0; k < n && !rc; k++) rc = commit(order[k], G.s);
    for (int b = 0; b < branches && !rc; b++)
        if (hipEventRecord(G.ev[1 + b], G.side[b]) != hipSuccess || hipStreamWaitEvent(G.s, G.ev[1 + b], 0) != hipSuccess)
            rc = -1;
    mam_prof_suspend(false);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(G.s, &g);
    if (ec != hipSuccess) rc = -1;
    G.graph = g;
    if (!rc && hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0) != hipSuccess) rc = -1;
    return rc;
}

bool graph_matches(const GraphBatch &G, const mam_batch_instance *inst, int n, int lr, int hash, int branches,
                   int max_wg) {
    if (!G.exec || (int)G.nv.size() != n || G.lr != lr || G.hash != hash || G.branches != branches ||
        G.max_wg != max_wg)
        return false;
    for (int i = 0; i < n; i++)
        if (G.nv[i] != inst[i].num_vars || G.fg[i] != inst[i].d_fg) return false;
    return true;
}

const uint8_t L_NUM_VAR[] = "num_var";
const uint8_t L_NUM_POLY[] = "num_poly";

}  // namespace

// Same instances, outputs and statuses as mam_prove_batch. branches: side streams for the
// commitments inside the graph (0..4; 0 runs them after the sumchecks); max_wg: workgroups
// per instance and pass (1..4096); workers: threads of the opening phase.
MAM_API int mam_prove_batch_graph(mam_batch_instance *inst, int n, const u64 *d_table, int table_nv, int log_rate,
                                  int hash, int num_queries, int workers, int branches, int max_wg) {
    if (!inst || n < 1 || workers < 1 || workers > 64 || !d_table || branches < 0 || branches > 4 || max_wg < 1 ||
        max_wg > 4096 || log_rate < 0 || (hash != 0 && hash != 1))
        return -2;
    for (int i = 0; i < n; i++)
        if (inst[i].num_vars < 2 || inst[i].num_vars > table_nv || inst[i].num_vars + log_rate > 30 || !inst[i].d_fg)
            return -2;
    std::lock_guard<std::mutex> lk(g_graph_mu);
    GraphBatch &G = graph_batch();
    const auto t0 = std::chrono::steady_clock::now();
    auto ms_since = [&]() {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    };
    if (!graph_matches(G, inst, n, log_rate, hash, branches, max_wg)) {
        (void)hipStreamSynchronize(mam_stream());
        if (graph_build(G, inst, n, log_rate, hash, branches, max_wg)) {
            G.release();
            return -1;
        }
    }
    // transcripts: new(label), then the prover's num_var / num_poly appends
    memset(G.host, 0, G.out_words * 8);
    for (int i = 0; i < n; i++) {
        mam_transcript t;
        mam::tr_init(&t, inst[i].label, inst[i].label_len);
        mam::tr_append_u64(&t, L_NUM_VAR, 7, (u64)inst[i].num_vars);
        mam::tr_append_u64(&t, L_NUM_POLY, 8, 2);
        memcpy(G.host + G.off[i], &t, sizeof(t));
    }
    const u64 seq = ++G.seq;
    G.host[G.out_words] = seq;
    // the caller's inputs may have been written on its own stream: order the graph after it
    if (hipStreamSynchronize(mam_stream()) != hipSuccess) return -1;
    if (hipMemcpyAsync(G.out, G.host, (G.out_words + 1) * 8, hipMemcpyHostToDevice, G.s) != hipSuccess ||
        hipGraphLaunch(G.exec, G.s) != hipSuccess)
        return -1;
    // openings start as their instance's commitment and sumcheck are ready (flags), while the
    // graph still runs the rest
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return inst[a].num_vars > inst[b].num_vars; });
    std::atomic<int> next(0), first_err(0);
    std::atomic<long long> ready_us(0);
    auto wait_flag = [&](const u64 *f) {
        for (unsigned it = 1;; it++) {
            if (__atomic_load_n(f, __ATOMIC_ACQUIRE) == seq) return 0;
            if ((it & 4095) == 0) {
                const hipError_t e = hipStreamQuery(G.s);
                if (e == hipSuccess) return __atomic_load_n(f, __ATOMIC_ACQUIRE) == seq ? 0 : -1;
                if (e != hipErrorNotReady) return -1;
            }
            __builtin_ia32_pause();
        }
    };
    auto work = [&]() {
        std::vector<u64> blk;
        for (int k; (k = next.fetch_add(1)) < n;) {
            const int i = order[k];
            mam_batch_instance &I = inst[i];
            I.start_ms = 0;
            const int nv = I.num_vars;
            int rc = wait_flag(G.flags + i);
            if (!rc) rc = wait_flag(G.flags + n + G.last_pass[i]);
            if (!rc) {
                long long t = (long long)(ms_since() * 1000.0), cur = ready_us.load();
                while (t > cur && !ready_us.compare_exchange_weak(cur, t)) {
                }
                blk.resize(GraphBatch::block_words(nv));
                if (hipMemcpyAsync(blk.data(), G.out + G.off[i], blk.size() * 8, hipMemcpyDeviceToHost,
                                   mam_stream()) != hipSuccess ||
                    hipStreamSynchronize(mam_stream()) != hipSuccess)
                    rc = -1;
            }
            if (!rc) {
                mam_transcript t;
                memcpy(&t, blk.data(), sizeof(t));
                memcpy(I.round_msgs, blk.data() + 22, 6 * (size_t)nv * 8);
                memcpy(I.challenges, blk.data() + 22 + 6 * (size_t)nv, 2 * (size_t)nv * 8);
                memcpy(I.final_evals, blk.data() + 22 + 8 * (size_t)nv, 4 * 8);
                memcpy(I.root, blk.data() + 26 + 8 * (size_t)nv, 4 * 8);
                void *h = nullptr;
                rc = mam_basefold_commitment_wrap(G.cw[i], G.bh[i], G.tree[i], I.root, nv, log_rate, hash, &h);
                if (!rc) rc = mam_basefold_open_ext2(h, I.challenges, d_table, table_nv, &t, num_queries, I.proof);
                if (h) mam_basefold_commitment_free(h);
            }
            I.status = rc;
            I.end_ms = ms_since();
            if (rc) {
                int z = 0;
                first_err.compare_exchange_strong(z, rc);
            }
        }
        (void)hipStreamSynchronize(mam_stream());
    };
    pool().run(workers, work);
    if (hipStreamSynchronize(G.s) != hipSuccess && !first_err.load()) first_err.store(-1);
    g_graph_phase_ms[0] = ready_us.load() / 1000.0;
    g_graph_phase_ms[1] = ms_since();
    return first_err.load();
}

// Times of the last graph-form call in ms from its start: [0] when the last instance's
// commitment and sumcheck were ready, [1] the end of the call (the openings overlap the graph).
MAM_API int mam_batch_graph_phase_ms(double *out2) {
    if (!out2) return -2;
    out2[0] = g_graph_phase_ms[0];
    out2[1] = g_graph_phase_ms[1];
    return 0;
}

// Frees the cached graph and its buffers (the next graph-form call builds them again).
MAM_API int mam_batch_graph_release() {
    std::lock_guard<std::mutex> lk(g_graph_mu);
    graph_batch().release();
    return 0;
}

