// The serving request loop (ssw_server_core.cpp) on the CPU alone, against
// stub engines, under a sanitizer (tests/test_server_sanitized.py builds this
// file with -fsanitize=thread and with -fsanitize=address).
//
// The stub engines solve requests with the CPU oracle, so every result that
// comes back through a ticket is checked against the oracle's answer for that
// ticket's automaton.  The large stub keeps instances running for a number of
// ticks that depends on the automaton, so requests overlap, finish out of
// order and compete for slots and for the memory budget, as on the device.
//
// Scenarios: many submitting/waiting/polling threads at once; two callers
// waiting on one ticket (one result, one SSW_ERR_NOTFOUND); a timed-out wait
// keeping the ticket; the budget check at submit agreeing with admission;
// destroy while callers are blocked in wait and requests are still queued;
// cancellation racing the engines, with callers blocked on the cancelled
// tickets.
#include "ssw_server_core.h"

#include <atomic>
#include <condition_variable>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

extern "C" int64_t oracle_wide_chunk(int n, int64_t count, const int32_t *A, const int32_t *B, int32_t *length,
                                     uint8_t *sync, int64_t *expanded, int32_t bound);

using namespace ssw;

static std::atomic<int> failures{0};
#define CHECK(cond, ...)                                              \
    do {                                                              \
        if (!(cond)) {                                                \
            failures.fetch_add(1);                                    \
            fprintf(stderr, "FAIL %s:%d: ", __FILE__, __LINE__);      \
            fprintf(stderr, __VA_ARGS__);                             \
            fprintf(stderr, "\n");                                    \
        }                                                             \
    } while (0)

static void oracle_one(int n, const uint8_t *delta, int32_t *len, int64_t *exp) {
    std::vector<int32_t> a(n), b(n);
    for (int q = 0; q < n; ++q) {
        a[q] = delta[q];
        b[q] = delta[n + q];
    }
    oracle_wide_chunk(n, 1, a.data(), b.data(), len, nullptr, exp, -1);
}

struct StubSmall final : SmallBackend {
    std::atomic<int> rounds{0};
    int solve(size_t m, const uint8_t *n_per, const uint64_t *offs, const uint8_t *delta, int32_t *length,
              int64_t *expanded, uint8_t *witness, int stride, std::string *) override {
        rounds.fetch_add(1);
        for (size_t i = 0; i < m; ++i) oracle_one(n_per[i], delta + offs[i], &length[i], &expanded[i]);
        if (witness)
            for (size_t i = 0; i < m; ++i) witness[i * (size_t)stride] = (uint8_t)length[i];
        std::this_thread::sleep_for(std::chrono::microseconds(200 + 50 * (int)(m % 7)));
        return SSW_OK;
    }
};

// Instances live for ticks_for(n, delta) ticks; only the large-engine thread
// calls the mutating methods, the budget queries come from any thread.
struct StubLarge final : LargeBackend {
    struct Inst {
        bool used = false;
        int n = 0;
        uint8_t delta[2 * SSW_MAX_N];
        int ticks_left = 0;
        uint64_t bytes = 0;
    };
    static constexpr int kSlots = 6;
    Inst slots[kSlots];
    uint64_t in_use = 0;
    uint64_t budget_default = 1ull << 30;
    bool set_up = false;
    int setup(std::string *) override {
        set_up = true;
        return SSW_OK;
    }
    void teardown() override { set_up = false; }
    int capacity() const override { return kSlots; }
    uint64_t default_budget() const override { return budget_default; }
    uint64_t admit_bytes(int n, bool witness) const override {
        return (1ull << n) / 8 + 4 * (1ull << std::min(n, 16)) + (witness ? 4096 : 0);
    }
    uint64_t bytes_in_use() const override { return in_use; }
    int running() const override {
        int r = 0;
        for (const Inst &s : slots) r += s.used ? 1 : 0;
        return r;
    }
    int admit(int n, const uint8_t *delta, bool witness, uint64_t budget, int *slot, std::string *err) override {
        CHECK(set_up, "admit before setup");
        *slot = -1;
        const uint64_t need = admit_bytes(n, witness);
        if (in_use + need > budget && in_use > 0) return SSW_OK;
        if (need > budget) {
            *err = "request larger than the budget reached admission";
            return SSW_ERR_BUDGET;
        }
        for (int k = 0; k < kSlots; ++k)
            if (!slots[k].used) {
                Inst &s = slots[k];
                s.used = true;
                s.n = n;
                std::memcpy(s.delta, delta, 2 * (size_t)n);
                s.ticks_left = 1 + (delta[0] + delta[n - 1] + n) % 5;
                s.bytes = need;
                in_use += need;
                *slot = k;
                return SSW_OK;
            }
        return SSW_OK;
    }
    int tick(std::string *) override {
        std::this_thread::sleep_for(std::chrono::microseconds(300));
        for (Inst &s : slots)
            if (s.used && s.ticks_left > 0) --s.ticks_left;
        return SSW_OK;
    }
    int harvest(bool witness, std::vector<Finished> *out, std::string *) override {
        for (int k = 0; k < kSlots; ++k) {
            Inst &s = slots[k];
            if (!s.used || s.ticks_left > 0) continue;
            Finished f;
            std::memset(&f, 0, sizeof f);
            f.slot = k;
            oracle_one(s.n, s.delta, &f.length, &f.expanded);
            f.witness_valid = witness && f.length >= 0;
            f.witness[0] = (uint8_t)f.length;
            s.used = false;
            in_use -= s.bytes;
            out->push_back(f);
        }
        return SSW_OK;
    }
};

struct Req {
    int n;
    std::vector<int32_t> a, b;
    int32_t want_len;
    int64_t want_exp;
};

static Req make_req(std::mt19937 &rng, int nlo, int nhi) {
    Req r;
    r.n = std::uniform_int_distribution<int>(nlo, nhi)(rng);
    r.a.resize(r.n);
    r.b.resize(r.n);
    for (int q = 0; q < r.n; ++q) {
        r.a[q] = std::uniform_int_distribution<int>(0, r.n - 1)(rng);
        r.b[q] = std::uniform_int_distribution<int>(0, r.n - 1)(rng);
    }
    oracle_wide_chunk(r.n, 1, r.a.data(), r.b.data(), &r.want_len, nullptr, &r.want_exp, -1);
    return r;
}

static ssw_server_config config(int small_n_max, int workers, bool witness, uint64_t budget = 0) {
    ssw_server_config c;
    std::memset(&c, 0, sizeof c);
    c.mem_budget = budget;
    c.small_n_max = small_n_max;
    c.levels_per_tick = 4;
    c.max_large = 64;
    c.want_witness = witness ? 1 : 0;
    c.small_workers = workers;
    return c;
}

// Many threads submit, then wait on or poll their own tickets.
static void scenario_many_callers() {
    StubSmall s0, s1;
    StubLarge large;
    ServerCore core(config(12, 2, true), {&s0, &s1}, &large);
    core.start();
    std::vector<std::thread> ts;
    for (int t = 0; t < 6; ++t)
        ts.emplace_back([&core, t] {
            std::mt19937 rng(100 + t);
            std::vector<Req> reqs;
            std::vector<uint64_t> tickets;
            for (int i = 0; i < 120; ++i) {
                reqs.push_back(make_req(rng, 2, 18));
                uint64_t tk = 0;
                const int rc = core.submit(reqs.back().n, reqs.back().a.data(), reqs.back().b.data(), &tk);
                CHECK(rc == SSW_OK, "submit rc %d", rc);
                tickets.push_back(tk);
                if (i % 17 == 0) std::this_thread::sleep_for(std::chrono::microseconds(300));
            }
            for (size_t i = 0; i < reqs.size(); ++i) {
                ssw_result res;
                int rc;
                if ((i + (size_t)t) % 3 == 0) {
                    while ((rc = core.poll(tickets[i], &res)) == SSW_PENDING)
                        std::this_thread::sleep_for(std::chrono::microseconds(100));
                } else {
                    rc = core.wait(tickets[i], 60000, &res);
                }
                CHECK(rc == SSW_OK, "thread %d request %zu rc %d", t, i, rc);
                CHECK(res.length == reqs[i].want_len && res.expanded == reqs[i].want_exp && res.n == reqs[i].n,
                      "thread %d request %zu: got (%d, %lld) want (%d, %lld)", t, i, res.length,
                      (long long)res.expanded, reqs[i].want_len, (long long)reqs[i].want_exp);
                CHECK(!res.witness_valid || res.witness[0] == (uint8_t)res.length, "witness row of another request");
                CHECK(core.poll(tickets[i], &res) == SSW_ERR_NOTFOUND, "a result handed out twice");
            }
        });
    // a stats reader alongside
    std::atomic<bool> done{false};
    std::thread st([&] {
        while (!done.load()) {
            ssw_server_stats s;
            core.stats(&s);
            CHECK(s.completed <= s.submitted, "completed > submitted");
            std::this_thread::sleep_for(std::chrono::microseconds(500));
        }
    });
    for (auto &t : ts) t.join();
    done.store(true);
    st.join();
    ssw_server_stats s;
    core.stats(&s);
    CHECK(s.completed == 6 * 120 && s.submitted == 6 * 120, "completed %llu", (unsigned long long)s.completed);
    CHECK(s.p99_us >= s.p50_us && s.p50_us > 0, "percentiles");
    CHECK(s0.rounds.load() > 0 && s1.rounds.load() > 0, "both small workers took rounds (%d, %d)", s0.rounds.load(),
          s1.rounds.load());
    core.shutdown();
}

// Two callers block on one ticket: one result, one SSW_ERR_NOTFOUND.
static void scenario_two_waiters_one_ticket() {
    StubSmall s0;
    StubLarge large;
    ServerCore core(config(4, 1, false), {&s0}, &large);
    core.start();
    std::mt19937 rng(7);
    for (int round = 0; round < 20; ++round) {
        Req r = make_req(rng, 14, 18);
        uint64_t tk = 0;
        CHECK(core.submit(r.n, r.a.data(), r.b.data(), &tk) == SSW_OK, "submit");
        std::atomic<int> ok{0}, notfound{0};
        std::vector<std::thread> ws;
        for (int k = 0; k < 3; ++k)
            ws.emplace_back([&] {
                ssw_result res;
                const int rc = core.wait(tk, 60000, &res);
                if (rc == SSW_OK) {
                    ok.fetch_add(1);
                    CHECK(res.length == r.want_len, "wrong result");
                } else if (rc == SSW_ERR_NOTFOUND) {
                    notfound.fetch_add(1);
                } else {
                    CHECK(false, "wait rc %d", rc);
                }
            });
        for (auto &w : ws) w.join();
        CHECK(ok.load() == 1 && notfound.load() == 2, "ok %d notfound %d", ok.load(), notfound.load());
    }
    core.shutdown();
}

// A wait that times out keeps the ticket; the budget check at submit is the
// figure admission uses.
static void scenario_timeout_and_budget() {
    StubSmall s0;
    StubLarge large;
    const uint64_t budget = large.admit_bytes(20, true);
    ServerCore core(config(4, 1, true, budget), {&s0}, &large);
    core.start();
    std::mt19937 rng(9);
    Req r = make_req(rng, 20, 20);
    uint64_t tk = 0;
    CHECK(core.submit(r.n, r.a.data(), r.b.data(), &tk) == SSW_OK, "a request of exactly the budget is taken");
    Req big = make_req(rng, 21, 21);
    uint64_t tk2 = 0;
    CHECK(core.submit(big.n, big.a.data(), big.b.data(), &tk2) == SSW_ERR_BUDGET, "over the budget at submit");
    ssw_result res;
    const int rc = core.wait(tk, 0, &res);
    CHECK(rc == SSW_ERR_TIMEOUT || rc == SSW_OK, "wait(0) rc %d", rc);
    if (rc == SSW_ERR_TIMEOUT) CHECK(core.wait(tk, 60000, &res) == SSW_OK, "the ticket survives a time-out");
    CHECK(res.length == r.want_len, "result");
    ssw_server_stats s;
    core.stats(&s);
    CHECK(s.rejected == 1, "rejected %llu", (unsigned long long)s.rejected);
    core.shutdown();
}

// Destroy with requests queued and callers blocked in wait: the requests are
// finished, the callers return, and shutdown returns after them.
static void scenario_destroy_with_waiters() {
    for (int rep = 0; rep < 5; ++rep) {
        StubSmall s0, s1;
        StubLarge large;
        ServerCore *core = new ServerCore(config(10, 2, false), {&s0, &s1}, &large);
        core->start();
        std::mt19937 rng(50 + rep);
        std::vector<Req> reqs;
        std::vector<uint64_t> tickets;
        for (int i = 0; i < 80; ++i) {
            reqs.push_back(make_req(rng, 3, 19));
            uint64_t tk = 0;
            core->submit(reqs.back().n, reqs.back().a.data(), reqs.back().b.data(), &tk);
            tickets.push_back(tk);
        }
        std::atomic<int> returned{0};
        std::vector<std::thread> ws;
        for (int k = 0; k < 8; ++k)
            ws.emplace_back([&, k] {
                for (size_t i = (size_t)k; i < tickets.size(); i += 8) {
                    ssw_result res;
                    const int rc = core->wait(tickets[i], -1, &res);
                    CHECK(rc == SSW_OK || rc == SSW_ERR_STATE || rc == SSW_ERR_NOTFOUND, "wait rc %d", rc);
                    if (rc == SSW_OK) CHECK(res.length == reqs[i].want_len, "result after destroy began");
                    if (rc != SSW_OK) break;  // the server is closing: later tickets may be gone
                }
                returned.fetch_add(1);
            });
        std::this_thread::sleep_for(std::chrono::milliseconds(2 + rep));
        core->shutdown();
        for (auto &w : ws) w.join();
        CHECK(returned.load() == 8, "waiters returned %d", returned.load());
        uint64_t tk;
        const int32_t a5[5] = {1, 2, 3, 4, 0}, b5[5] = {0, 0, 2, 3, 4};
        CHECK(core->submit(5, a5, b5, &tk) == SSW_ERR_STATE, "submit after shutdown");
        delete core;
    }
}

// Cancel while the engines take requests: a cancel either wins (the ticket is
// gone, its blocked waiters return SSW_ERR_CANCELLED) or finds the request
// taken (SSW_ERR_STATE, and its result arrives); no other outcome, and every
// request not cancelled gets the oracle's result.
static void scenario_cancel() {
    for (int rep = 0; rep < 4; ++rep) {
        StubSmall s0, s1;
        StubLarge large;
        ServerCore core(config(10, 2, false), {&s0, &s1}, &large);
        core.start();
        std::mt19937 rng(90 + rep);
        const int m = 200;
        std::vector<Req> reqs;
        std::vector<uint64_t> tickets;
        for (int i = 0; i < m; ++i) {
            reqs.push_back(make_req(rng, 3, 19));
            uint64_t tk = 0;
            CHECK(core.submit(reqs.back().n, reqs.back().a.data(), reqs.back().b.data(), &tk) == SSW_OK, "submit");
            tickets.push_back(tk);
        }
        // a blocked waiter on every third ticket, then cancels of every
        // other ticket from two threads
        std::vector<int> cancel_rc(m, 1), wait_rc(m, 1);
        std::vector<std::thread> ws;
        for (int k = 0; k < 4; ++k)
            ws.emplace_back([&, k] {
                for (int i = k * 3; i < m; i += 12) {
                    ssw_result res;
                    wait_rc[i] = core.wait(tickets[i], 60000, &res);
                    if (wait_rc[i] == SSW_OK) CHECK(res.length == reqs[i].want_len, "result of %d", i);
                }
            });
        std::vector<std::thread> cs;
        for (int k = 0; k < 2; ++k)
            cs.emplace_back([&, k] {
                for (int i = k; i < m; i += 4) cancel_rc[i] = core.cancel(tickets[i]);
            });
        for (auto &c : cs) c.join();
        for (auto &w : ws) w.join();
        int cancelled = 0;
        for (int i = 0; i < m; ++i) {
            ssw_result res;
            if (i % 4 < 2) {
                // a waiter may collect the result before the cancel (then
                // the ticket is gone: NOTFOUND)
                const bool collected = i % 3 == 0 && wait_rc[i] == SSW_OK;
                CHECK(cancel_rc[i] == SSW_OK || cancel_rc[i] == SSW_ERR_STATE ||
                          (collected && cancel_rc[i] == SSW_ERR_NOTFOUND),
                      "cancel rc %d (waiter %d)", cancel_rc[i], wait_rc[i]);
                if (cancel_rc[i] == SSW_OK) {
                    ++cancelled;
                    // blocked before the cancel: CANCELLED; came after it: NOTFOUND
                    if (i % 3 == 0)
                        CHECK(wait_rc[i] == SSW_ERR_CANCELLED || wait_rc[i] == SSW_ERR_NOTFOUND,
                              "waiter on a cancelled ticket: %d", wait_rc[i]);
                    CHECK(core.poll(tickets[i], &res) == SSW_ERR_NOTFOUND, "a cancelled ticket is gone");
                    CHECK(core.cancel(tickets[i]) == SSW_ERR_NOTFOUND, "cancel twice");
                    continue;
                }
                if (collected) continue;
            }
            if (i % 3 == 0) {
                CHECK(wait_rc[i] == SSW_OK, "waiter %d rc %d", i, wait_rc[i]);
            } else {
                CHECK(core.wait(tickets[i], 60000, &res) == SSW_OK, "result %d", i);
                CHECK(res.length == reqs[i].want_len, "result of %d", i);
            }
        }
        CHECK(core.cancel(tickets[0] + 100000) == SSW_ERR_NOTFOUND, "unknown ticket");
        // what was done stays done
        CHECK(cancelled > 0 || rep > 0, "no cancel won the race (%d)", cancelled);
        core.shutdown();
    }
}

// A small engine that holds its round until released: a request it has
// taken cannot be cancelled, one still queued can, and a caller blocked on
// the queued one returns SSW_ERR_CANCELLED.
struct GatedSmall final : SmallBackend {
    std::mutex mu;
    std::condition_variable cv;
    bool open = false, entered = false;
    int solve(size_t m, const uint8_t *n_per, const uint64_t *offs, const uint8_t *delta, int32_t *length,
              int64_t *expanded, uint8_t *, int, std::string *) override {
        std::unique_lock<std::mutex> lk(mu);
        entered = true;
        cv.notify_all();
        cv.wait(lk, [&] { return open; });
        for (size_t i = 0; i < m; ++i) oracle_one(n_per[i], delta + offs[i], &length[i], &expanded[i]);
        return SSW_OK;
    }
};

static void scenario_cancel_blocked_waiter() {
    GatedSmall gate;
    StubLarge large;
    ServerCore core(config(10, 1, false), {&gate}, &large);
    core.start();
    std::mt19937 rng(123);
    Req ra = make_req(rng, 6, 9), rb = make_req(rng, 6, 9);
    uint64_t ta = 0, tb = 0;
    CHECK(core.submit(ra.n, ra.a.data(), ra.b.data(), &ta) == SSW_OK, "submit a");
    {
        std::unique_lock<std::mutex> lk(gate.mu);
        gate.cv.wait(lk, [&] { return gate.entered; });  // a's round is in the engine
    }
    CHECK(core.submit(rb.n, rb.a.data(), rb.b.data(), &tb) == SSW_OK, "submit b");
    std::atomic<int> wrc{1};
    std::thread w([&] {
        ssw_result res;
        wrc = core.wait(tb, 60000, &res);
    });
    std::this_thread::sleep_for(std::chrono::milliseconds(50));  // the waiter blocks on b
    CHECK(core.cancel(ta) == SSW_ERR_STATE, "a is being solved");
    CHECK(core.cancel(tb) == SSW_OK, "b is still queued");
    w.join();
    CHECK(wrc.load() == SSW_ERR_CANCELLED, "the blocked waiter: %d", wrc.load());
    {
        std::lock_guard<std::mutex> lk(gate.mu);
        gate.open = true;
    }
    gate.cv.notify_all();
    ssw_result res;
    CHECK(core.wait(ta, 60000, &res) == SSW_OK && res.length == ra.want_len, "a's result");
    CHECK(core.poll(tb, &res) == SSW_ERR_NOTFOUND, "b is gone");
    core.shutdown();
}

int main() {
    scenario_cancel_blocked_waiter();
    scenario_cancel();
    scenario_many_callers();
    scenario_two_waiters_one_ticket();
    scenario_timeout_and_budget();
    scenario_destroy_with_waiters();
    if (failures.load()) {
        fprintf(stderr, "%d failures\n", failures.load());
        return 1;
    }
    printf("server stress passed\n");
    return 0;
}
