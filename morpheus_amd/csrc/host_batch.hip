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

#include "common.h"
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
