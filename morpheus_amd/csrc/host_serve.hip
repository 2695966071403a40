// Proving service for matmul proofs (the benchmark's fifth configuration): requests arrive
// on a clock, are admitted while their HBM footprint fits a budget, and are proved by W
// persistent worker threads, each on its own stream with its own buffers, so requests that
// arrive while others are on the GPU start as soon as a worker and memory are free and their
// kernels overlap on the device (continuous batching, not a queue drained one at a time).
//
// Admission: a request's footprint is what its proof allocates on the device: for each of A,
// B, C with 2^nv entries, the evaluations' codeword at rate 2^-lr, the bit-reversed copy and
// the Merkle tree, plus the opening's folded codewords and trees (about 2x the codeword
// layers) -- see matmul_footprint(). Requests are admitted in arrival order (FIFO); a request
// larger than the whole budget is admitted alone.
//
// Timing: arrival times are offsets (ms) from the service start; the dispatcher releases each
// request at its arrival time; latency = completion - arrival. Inputs (A, B, C) are resident
// on the device before the service starts, as the outputs of the layer being proved would be.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"
#include "matmul.h"
#include "workers.h"


struct mam_serve_request {
    mam_matmul_instance *inst;  // statement and output buffers
    double arrival_ms;          // offset from the service start
    double admit_ms, done_ms;   // filled: admission and completion times
    int status;
};

namespace {

size_t poly_footprint(int nv, int lr) {
    const size_t cw = (size_t)1 << (nv + lr);
    // codeword + bh + tree, then the opening's folded codewords (Ext2) and their trees
    return cw * 8 + ((size_t)1 << nv) * 8 + (cw - 1) * 32 + cw * 16 + cw * 32;
}

size_t matmul_footprint(const mam_matmul_instance &h, int lr) {
    return poly_footprint(h.lm + h.lk, lr) + poly_footprint(h.lk + h.ln, lr) + poly_footprint(h.lm + h.ln, lr);
}

}  // namespace

// Runs the service over n requests (sorted or not; released by arrival time) with `workers`
// prover threads and an HBM admission budget in bytes. Returns 0 when every request
// succeeded, else the first failing status.
MAM_API int mam_serve_matmul(mam_serve_request *req, int n, const u64 *d_table, int table_nv, int log_rate, int hash,
                             int num_queries, void *ajtai_handle, int ajtai_rows, int workers, uint64_t hbm_budget) {
    if (!req || n < 0 || workers < 1 || workers > 64 || !d_table) return -2;
    std::vector<int> order(n);
    for (int i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival_ms < req[b].arrival_ms; });
    std::vector<size_t> foot(n);
    for (int i = 0; i < n; i++) foot[i] = matmul_footprint(*req[i].inst, log_rate);

    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> ready;     // admitted, waiting for a worker
    size_t in_flight = 0;      // bytes of admitted, unfinished requests
    int next_arrival = 0;      // index into order of the next request to arrive
    int finished = 0;
    bool stop = false;
    std::atomic<int> first_err(0);
    const auto t0 = std::chrono::steady_clock::now();
    auto now_ms = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
    std::deque<int> arrived;  // arrived, waiting for admission (FIFO)

    // Admit arrived requests in order while the budget allows (caller holds mu).
    auto admit = [&]() {
        while (!arrived.empty()) {
            const int i = arrived.front();
            if (in_flight != 0 && in_flight + foot[i] > hbm_budget) break;
            arrived.pop_front();
            in_flight += foot[i];
            req[i].admit_ms = now_ms();
            ready.push_back(i);
            cv.notify_one();
        }
    };

    auto worker = [&]() {
        for (;;) {
            int i;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return stop || !ready.empty(); });
                if (ready.empty()) return;  // stop requested and nothing left
                i = ready.front();
                ready.pop_front();
            }
            const int rc = mam_matmul_prove(req[i].inst, d_table, table_nv, log_rate, hash, num_queries, ajtai_handle,
                                            ajtai_rows);
            (void)hipStreamSynchronize(mam_stream());
            std::lock_guard<std::mutex> lk(mu);
            req[i].status = rc;
            req[i].done_ms = now_ms();
            if (rc) {
                int z = 0;
                first_err.compare_exchange_strong(z, rc);
            }
            in_flight -= foot[i];
            finished++;
            admit();
            cv.notify_all();
        }
    };

    static WorkerPool *serve_pool = new WorkerPool();
    std::thread runner([&] { serve_pool->run(workers, worker); });
    // dispatcher: release requests at their arrival times
    while (next_arrival < n) {
        const int i = order[next_arrival];
        const double wait = req[i].arrival_ms - now_ms();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(wait));
        std::lock_guard<std::mutex> lk(mu);
        arrived.push_back(i);
        next_arrival++;
        admit();
    }
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return finished == n; });
        stop = true;
        cv.notify_all();
    }
    runner.join();
    return first_err.load();
}
