// stage_pool.h -- host threads that split the copies of a large batched step
// into the slot's pinned staging: the variable updates (with their range
// check), the clause literals (with their maximum) and the per-clause
// records.  A step of ~140k updates and ~45k clauses stages ~5 MB; one
// thread copying and checking it is most of the host's time per step.
//
// One engine owns one pool and uses it from one thread at a time (the
// engine is not thread-safe).  start() hands a job to every worker; wait()
// returns once each worker has run it.  Workers spin for a short while after
// a job (the next phase of the same step follows within microseconds) and
// then sleep on a condition variable.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace gss {

class StagePool {
public:
    explicit StagePool(int workers) {
        for (int i = 0; i < workers; i++) th_.emplace_back([this, i] { loop(i); });
    }
    ~StagePool() {
        {
            std::lock_guard<std::mutex> l(m_);
            stop_.store(true, std::memory_order_relaxed);
        }
        cv_.notify_all();
        for (std::thread &t : th_) t.join();
    }
    StagePool(const StagePool &) = delete;
    StagePool &operator=(const StagePool &) = delete;

    int size() const { return (int)th_.size(); }

    // Runs fn(0) .. fn(size() - 1), one part per worker, without waiting.
    // The previous job must have been waited for.
    void start(std::function<void(int)> fn) {
        job_ = std::move(fn);
        left_.store(size(), std::memory_order_relaxed);
        {
            std::lock_guard<std::mutex> l(m_);
            gen_.fetch_add(1, std::memory_order_release);
        }
        cv_.notify_all();
        running_ = true;
    }
    // Returns once every part of the started job has ended (no-op if none).
    void wait() {
        if (!running_) return;
        while (left_.load(std::memory_order_acquire) != 0) std::this_thread::yield();
        running_ = false;
    }
    void run(std::function<void(int)> fn) {
        start(std::move(fn));
        wait();
    }

private:
    void loop(int part) {
        uint64_t seen = 0;
        for (;;) {
            const auto t0 = std::chrono::steady_clock::now();
            while (gen_.load(std::memory_order_acquire) == seen &&
                   !stop_.load(std::memory_order_relaxed)) {
                if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200)) {
                    std::unique_lock<std::mutex> l(m_);
                    cv_.wait(l, [&] {
                        return stop_.load(std::memory_order_relaxed) ||
                               gen_.load(std::memory_order_relaxed) != seen;
                    });
                    break;
                }
                std::this_thread::yield();
            }
            if (stop_.load(std::memory_order_relaxed)) return;
            seen = gen_.load(std::memory_order_acquire);
            job_(part);
            left_.fetch_sub(1, std::memory_order_release);
        }
    }

    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_;
    std::atomic<uint64_t> gen_{0};
    std::atomic<int> left_{0};
    std::atomic<bool> stop_{false};
    std::function<void(int)> job_;
    bool running_ = false;  // touched by the owning thread only
};

// Waits for the pool's job when it leaves scope, so that no early return
// leaves workers reading the caller's arrays.
struct StageJoin {
    StagePool *pool = nullptr;
    ~StageJoin() {
        if (pool) pool->wait();
    }
};

}  // namespace gss
