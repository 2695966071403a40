// Persistent host worker threads for the multi-instance entry points (batched proving, the
// proving service). Each worker keeps its per-thread stream, arenas, mailbox and buffers
// between calls; a pool is never torn down and its idle threads wait on a condition
// variable. Concurrent run() calls on one pool are serialized.
#pragma once

#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>

#include "common.h"

class WorkerPool {
   public:
    // Run f on `n` pool threads (created on first need) and wait until every one returned.
    void run(int n, const std::function<void()> &f) {
        std::lock_guard<std::mutex> one_run(run_mu_);
        std::unique_lock<std::mutex> lk(m_);
        while ((int)count_ < n) {
            std::thread(&WorkerPool::loop, this, (int)count_).detach();
            count_++;
        }
        job_ = f;
        active_ = n;
        remaining_ = n;
        gen_++;
        cv_.notify_all();
        done_.wait(lk, [&] { return remaining_ == 0; });
    }

   private:
    void loop(int id) {
        // workers share the hardware queues with each other: no pre-queued rounds here
        mam_prequeue_set(false);
        unsigned seen = 0;
        for (;;) {
            std::function<void()> j;
            {
                std::unique_lock<std::mutex> lk(m_);
                cv_.wait(lk, [&] { return gen_ != seen && id < active_; });
                seen = gen_;
                j = job_;
            }
            j();
            std::lock_guard<std::mutex> lk(m_);
            if (--remaining_ == 0) done_.notify_all();
        }
    }
    std::mutex m_, run_mu_;
    std::condition_variable cv_, done_;
    std::function<void()> job_;
    unsigned gen_ = 0, count_ = 0;
    int active_ = 0, remaining_ = 0;
};
