// ssw_pool.h -- the host worker pool of a solver context (no HIP: the pool is
// also built alone on the CPU, under the sanitizers, by its stress test).
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace ssw {

// Host worker threads kept for a context's lifetime (started on first use):
// the host entry points pack, check and copy on them instead of starting
// threads on every call, which cost 0.1-0.25 ms per call (15 threads for the
// packing, up to 16 more for each copy to the caller).  run() hands fn(0) ..
// fn(k - 1) to the workers and returns at once; wait() returns when they are
// done; stop() lets the job handed out last finish.  One job at a time, from
// the context's calling thread.
class HostPool {
public:
    // threads = 0: one fewer than min(hardware threads, 16), at least one
    explicit HostPool(int threads = 0) : want_(threads) {}
    ~HostPool() { stop(); }
    int size() {
        start();
        return (int)threads_.size();
    }
    void run(std::function<void(int)> fn, int k) {
        start();
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return pending_ == 0; });
        job_ = std::move(fn);
        job_k_ = std::min<int>(k, (int)threads_.size());
        pending_ = (int)threads_.size();
        ++gen_;
        cv_.notify_all();
    }
    void wait() {
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return pending_ == 0; });
    }
    bool idle() {
        std::lock_guard<std::mutex> lk(mu_);
        return pending_ == 0;
    }
    // the job handed out last runs to its end first
    void stop() {
        {
            std::unique_lock<std::mutex> lk(mu_);
            done_cv_.wait(lk, [this] { return pending_ == 0; });
            quit_ = true;
        }
        cv_.notify_all();
        for (auto &t : threads_) t.join();
        threads_.clear();
    }

private:
    void start() {
        if (!threads_.empty()) return;
        const int hw = std::max(1, (int)std::thread::hardware_concurrency());
        const int n = want_ > 0 ? want_ : std::max(1, std::min(hw, 16) - 1);
        uint64_t g0;
        {
            std::lock_guard<std::mutex> lk(mu_);
            quit_ = false;
            g0 = gen_;  // a restarted pool's workers wait for the next job, not the last one
        }
        for (int i = 0; i < n; ++i)
            threads_.emplace_back([this, i, g0] {
                uint64_t seen = g0;
                for (;;) {
                    std::function<void(int)> fn;
                    int k = 0;
                    {
                        std::unique_lock<std::mutex> lk(mu_);
                        cv_.wait(lk, [&] { return quit_ || gen_ != seen; });
                        if (quit_) return;
                        seen = gen_;
                        fn = job_;
                        k = job_k_;
                    }
                    if (i < k) fn(i);
                    std::lock_guard<std::mutex> lk(mu_);
                    if (--pending_ == 0) done_cv_.notify_all();
                }
            });
    }
    std::vector<std::thread> threads_;
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    std::function<void(int)> job_;
    int want_ = 0, job_k_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
    bool quit_ = false;
};

}  // namespace ssw
