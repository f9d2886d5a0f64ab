// ssw_server_core.h -- the request loop behind the serving entry points
// (ssw_server_* in ssw.h), apart from the device: tickets, the two request
// queues, admission in arrival order under a memory budget, results handed out
// once, latency statistics and shutdown.
//
// The engines that solve requests are backends.  ssw_server.cpp binds them to
// the HIP solvers (batched rounds for small n, the device-wide engine for
// large n); the loop itself includes no HIP header, so it also builds on the
// CPU alone against stub backends (tests/cpp/server_tsan.cpp, under
// ThreadSanitizer).
#pragma once

#include "ssw.h"

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace ssw {

using ServerClock = std::chrono::steady_clock;

struct ServerRequest {
    uint64_t ticket = 0;
    int n = 0;
    uint8_t delta[2 * SSW_MAX_N];
    ServerClock::time_point t_submit, t_done;
    bool done = false;
    bool collected = false;  // the result was handed out; the ticket is gone
    bool cancelled = false;  // taken out of its queue by cancel; the ticket is gone
    int waiters = 0;         // callers blocked in wait on this request
    ssw_result res;
};

// Solves one round of small requests together.
struct SmallBackend {
    virtual ~SmallBackend() = default;
    // m requests in the packed batch format of ssw_batch (n_per, offsets,
    // transitions); writes length[m], expanded[m] and, when witness is not
    // null, m witness rows of `stride` bytes.  0 or an ssw_status (*err set).
    virtual int solve(size_t m, const uint8_t *n_per, const uint64_t *offs, const uint8_t *delta, int32_t *length,
                      int64_t *expanded, uint8_t *witness, int stride, std::string *err) = 0;
};

// Instances that run side by side, admitted and retired between ticks.
struct LargeBackend {
    struct Finished {
        int slot;
        int32_t length;
        int64_t expanded;
        bool witness_valid;
        uint8_t witness[sizeof(((ssw_result *)nullptr)->witness)];
    };
    virtual ~LargeBackend() = default;
    // called on the engine thread before anything else / after everything
    virtual int setup(std::string *err) = 0;
    virtual void teardown() = 0;
    virtual int capacity() const = 0;                          // concurrent instances
    virtual uint64_t default_budget() const = 0;               // bytes, when the config gives none
    virtual uint64_t admit_bytes(int n, bool witness) const = 0;  // what one instance holds when admitted
    virtual uint64_t bytes_in_use() const = 0;
    virtual int running() const = 0;                           // admitted and not finished
    // *slot = -1 when its memory is not free yet (nothing admitted)
    virtual int admit(int n, const uint8_t *delta, bool witness, uint64_t budget, int *slot, std::string *err) = 0;
    // advances every running instance by some levels
    virtual int tick(std::string *err) = 0;
    // instances finished by the last tick, retired (their slots free again)
    virtual int harvest(bool witness, std::vector<Finished> *out, std::string *err) = 0;
};

class ServerCore {
public:
    // Latencies kept for the percentiles: the most recent kLatencyWindow
    // completed requests (a long-running server holds no more than that).
    static constexpr size_t kLatencyWindow = 1u << 20;
    // Requests one small round takes at most, so one round never holds back
    // the requests behind it for long.
    static constexpr size_t kMaxRound = 16384;

    ServerCore(const ssw_server_config &cfg, std::vector<SmallBackend *> small, LargeBackend *large);
    ~ServerCore();
    // starts one thread per small backend and one for the large backend
    void start();
    int submit(int32_t n, const int32_t *a, const int32_t *b, uint64_t *ticket);
    int poll(uint64_t ticket, ssw_result *out);
    int wait(uint64_t ticket, int32_t timeout_ms, ssw_result *out);
    int cancel(uint64_t ticket);
    int stats(ssw_server_stats *out);
    // the message of the error that stopped the loop ("" if none); stable
    // once set (it is set once)
    const char *last_error();
    // Stops taking submissions, lets the engines finish every request already
    // submitted, joins them, wakes the callers still blocked in wait (they
    // return their results or SSW_ERR_STATE) and returns once none is left
    // inside the loop's calls.  Requests not collected are freed.
    void shutdown();
    const ssw_server_config &config() const { return cfg_; }
    uint64_t large_budget() const;

private:
    void small_loop(SmallBackend *be);
    void large_loop();
    void fail_all(int code, const std::string &msg);
    void finish(ServerRequest *r, int32_t length, int64_t expanded, const uint8_t *wit, bool wit_valid);
    void take(ServerRequest *r, ssw_result *out);

    ssw_server_config cfg_;
    std::vector<SmallBackend *> small_;
    LargeBackend *large_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_, cv_idle_;
    std::deque<ServerRequest *> small_q_, large_q_;
    std::unordered_map<uint64_t, ServerRequest *> reqs_;
    uint64_t next_ticket_ = 1;
    bool stop_ = false;    // no more submissions: the engines drain and leave
    bool closed_ = false;  // the engines are gone: blocked waiters return
    int waiting_ = 0;      // callers blocked in wait
    int error_ = SSW_OK;
    std::string error_msg_;
    std::vector<std::thread> threads_;
    bool started_ = false;
    bool shut_ = false;
    // statistics
    std::vector<double> lat_us_;  // ring of the last kLatencyWindow latencies
    size_t lat_next_ = 0;
    uint64_t submitted_ = 0, completed_ = 0, rejected_ = 0, subsets_ = 0;
    ServerClock::time_point first_submit_, last_done_;
    bool any_submit_ = false;
};

}  // namespace ssw
