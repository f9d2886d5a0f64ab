// Stress test of the host worker pool (csrc/ssw_pool.h), built for the CPU
// alone under ThreadSanitizer and AddressSanitizer by tests/test_pool_sanitized.py:
// jobs of every width handed out back to back (with and without waiting in
// between), each worker index run exactly once per job, idle() as the host
// entry points use it, and stop / restart / destruction with a job running.
#include "ssw_pool.h"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <vector>

static int fail(const char *what, long a, long b) {
    std::printf("pool stress FAILED: %s (%ld, %ld)\n", what, a, b);
    return 1;
}

int main() {
    for (int threads : {1, 3, 7}) {
        ssw::HostPool pool(threads);
        if (pool.size() != threads) return fail("size", pool.size(), threads);
        if (!pool.idle()) return fail("idle at start", 0, 0);
        std::vector<std::atomic<int>> hits(16);
        for (int round = 0; round < 2000; ++round) {
            const int k = 1 + round % 10;
            for (auto &h : hits) h.store(0);
            std::atomic<long> sum{0};
            pool.run([&](int i) {
                hits[(size_t)i].fetch_add(1);
                sum.fetch_add(i + 1);
            }, k);
            pool.wait();
            if (!pool.idle()) return fail("idle after wait", round, 0);
            const int ran = k < threads ? k : threads;
            for (int i = 0; i < 16; ++i)
                if (hits[(size_t)i].load() != (i < ran ? 1 : 0)) return fail("hits", i, hits[(size_t)i].load());
            if (sum.load() != (long)ran * (ran + 1) / 2) return fail("sum", sum.load(), ran);
        }
        // jobs handed out without waiting: run() waits for the previous one
        std::atomic<long> total{0};
        for (int round = 0; round < 500; ++round)
            pool.run([&](int i) { total.fetch_add(i + 1); }, threads);
        pool.wait();
        if (total.load() != 500L * threads * (threads + 1) / 2) return fail("back to back", total.load(), threads);
        // a slow job seen as busy, then stop while another runs, then reuse
        std::atomic<bool> go{false};
        pool.run([&](int) {
            while (!go.load()) std::this_thread::sleep_for(std::chrono::microseconds(50));
        }, threads);
        if (pool.idle()) return fail("busy job reported idle", 0, 0);
        go.store(true);
        pool.wait();
        pool.run([&](int) { std::this_thread::sleep_for(std::chrono::milliseconds(2)); }, threads);
        pool.stop();
        std::atomic<int> after{0};
        pool.run([&](int) { after.fetch_add(1); }, threads);
        pool.wait();
        if (after.load() != threads) return fail("restart", after.load(), threads);
        pool.run([&](int) { std::this_thread::sleep_for(std::chrono::milliseconds(1)); }, threads);
        // destroyed with that job running
    }
    std::printf("pool stress passed\n");
    return 0;
}
