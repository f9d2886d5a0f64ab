// ssw_server.cpp -- continuous-batching request loop (ssw_server_* in ssw.h).
//
// Two engines run side by side on the same device, each on its own stream:
//   * small requests (n <= small_n_max) are gathered while the previous round
//     is on the GPU and solved together by one batched persistent launch;
//   * large requests join the device-wide engine between graph replays of
//     `levels_per_tick` levels: a request that arrives while others are in
//     flight is admitted as soon as a slot and its memory (2^n/8 bitmap +
//     frontier ring) fit the budget, and finished instances leave at the same
//     points, so their work overlaps on the device.
// Every request is solved by the same kernels as a lone solve, so its result
// does not depend on what else is in flight.
#include "ssw_host.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <unordered_map>

using namespace ssw;
using Clock = std::chrono::steady_clock;

namespace {

struct Request {
    uint64_t ticket = 0;
    int n = 0;
    uint8_t delta[2 * SSW_MAX_N];
    Clock::time_point t_submit, t_done;
    bool done = false;
    int slot = -1;
    ssw_result res;
};

}  // namespace

struct ssw_server {
    ssw_context *owner = nullptr;
    ssw_context *small_ctx = nullptr;
    ssw_context *large_ctx = nullptr;
    ssw_server_config cfg;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request *> small_q, large_q;
    std::unordered_map<uint64_t, Request *> reqs;
    uint64_t next_ticket = 1;
    bool stop = false;
    int error = SSW_OK;
    std::string error_msg;
    std::thread small_thread, large_thread;
    // stats
    std::vector<double> lat_us;
    uint64_t submitted = 0, completed = 0, rejected = 0, subsets = 0;
    Clock::time_point first_submit, last_done;
    bool any_submit = false;

    void finish(Request *r, int32_t length, int64_t expanded, const uint8_t *wit, bool wit_valid) {
        r->t_done = Clock::now();
        r->res.length = length;
        r->res.n = r->n;
        r->res.expanded = expanded;
        r->res.sync = length >= 0 ? 1 : 0;
        r->res.witness_valid = wit_valid ? 1 : 0;
        if (wit_valid) std::memcpy(r->res.witness, wit, sizeof r->res.witness);
        r->res.latency_us = std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count();
        r->done = true;
        lat_us.push_back(r->res.latency_us);
        ++completed;
        subsets += (uint64_t)std::max<int64_t>(expanded, 0);
        last_done = r->t_done;
    }

    // A finished request is handed out once and then forgotten, so a
    // long-running server holds only the requests in flight or not yet
    // collected (the engines no longer touch a request once it is done).
    void release(std::unordered_map<uint64_t, Request *>::iterator it) {
        delete it->second;
        reqs.erase(it);
    }

    void fail_all(int code, const std::string &msg) {
        std::lock_guard<std::mutex> lk(mu);
        error = code;
        error_msg = msg;
        stop = true;
        cv_work.notify_all();
        cv_done.notify_all();
    }

    // ---- small engine: one batched launch per round ------------------------
    void small_loop() {
        std::vector<Request *> round;
        std::vector<uint8_t> n_per, delta, wit;
        std::vector<uint64_t> offs;
        std::vector<int32_t> length;
        std::vector<int64_t> expanded;
        const int stride = (int)sizeof(((ssw_result *)0)->witness);
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stop || !small_q.empty(); });
                if (stop && small_q.empty()) return;
                round.assign(small_q.begin(), small_q.end());
                small_q.clear();
            }
            const size_t m = round.size();
            n_per.resize(m);
            offs.resize(m);
            delta.clear();
            for (size_t i = 0; i < m; ++i) {
                n_per[i] = (uint8_t)round[i]->n;
                offs[i] = delta.size();
                delta.insert(delta.end(), round[i]->delta, round[i]->delta + 2 * round[i]->n);
            }
            length.assign(m, 0);
            expanded.assign(m, 0);
            if (cfg.want_witness) wit.assign(m * (size_t)stride, 0);
            ssw_batch b;
            std::memset(&b, 0, sizeof b);
            b.count = (int64_t)m;
            b.n_per = n_per.data();
            b.offset = offs.data();
            b.delta = delta.data();
            b.length = length.data();
            b.expanded = expanded.data();
            b.witness = cfg.want_witness ? wit.data() : nullptr;
            b.witness_stride = stride;
            b.bound = -1;
            int rc;
            {
                std::lock_guard<std::mutex> lk(small_ctx->mu);
                rc = small_ctx->solve_entry(&b, SSW_HOST, kModeBatch, nullptr);
            }
            if (rc) {
                fail_all(rc, small_ctx->err);
                return;
            }
            std::lock_guard<std::mutex> lk(mu);
            for (size_t i = 0; i < m; ++i)
                finish(round[i], length[i], expanded[i], cfg.want_witness ? wit.data() + i * stride : nullptr,
                       cfg.want_witness && length[i] >= 0);
            cv_done.notify_all();
        }
    }

    // ---- large engine: admit / tick / harvest --------------------------------
    void large_loop() {
        ssw_context &ctx = *large_ctx;
        LargeEngine &E = ctx.large;
        std::vector<Request *> in_slot(kLargeMaxInstances, nullptr);
        uint8_t *d_delta = nullptr, *d_wit = nullptr;
        const int stride = (int)sizeof(((ssw_result *)0)->witness);
        std::vector<uint8_t> hwit(stride);
        std::lock_guard<std::mutex> clk(ctx.mu);
        if (hipSetDevice(ctx.device) != hipSuccess || E.setup(ctx) ||
            E.set_outputs(ctx, nullptr, nullptr, nullptr, nullptr, nullptr) ||
            hipMalloc((void **)&d_delta, 64ull * kLargeMaxInstances) != hipSuccess ||
            hipMalloc((void **)&d_wit, (size_t)stride * kLargeMaxInstances) != hipSuccess) {
            fail_all(SSW_ERR_DEVICE, "large engine set-up failed: " + ctx.err);
            return;
        }
        E.tick_levels = std::max(1, cfg.levels_per_tick);
        const uint64_t budget = cfg.mem_budget ? cfg.mem_budget : ctx.budget / 2;
        for (;;) {
            std::vector<Request *> admit_now;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (E.running() == 0)
                    cv_work.wait(lk, [&] { return stop || !large_q.empty(); });
                if (stop && large_q.empty() && E.running() == 0) break;
                // admit in arrival order while slots and memory allow
                uint64_t planned = E.bytes_in_use;
                while (!large_q.empty()) {
                    Request *r = large_q.front();
                    const int in_flight = E.running() + (int)admit_now.size();
                    if (in_flight >= std::min(kLargeMaxInstances, std::max(1, cfg.max_large))) break;
                    const uint64_t need = LargeEngine::admit_bytes(r->n, cfg.want_witness != 0, 1ull << 16);
                    if (planned + need > budget && in_flight > 0) break;
                    planned += need;
                    admit_now.push_back(r);
                    large_q.pop_front();
                }
            }
            for (size_t k = 0; k < admit_now.size(); ++k) {
                Request *r = admit_now[k];
                int slot = -1;
                // stage the transitions where large_init reads them
                int free_idx = -1;
                for (int s = 0; s < kLargeMaxInstances; ++s)
                    if (!E.slots[s].used) {
                        free_idx = s;
                        break;
                    }
                if (free_idx < 0 ||
                    hipMemcpy(d_delta + 64ull * free_idx, r->delta, 2 * r->n, hipMemcpyHostToDevice) != hipSuccess) {
                    fail_all(SSW_ERR_DEVICE, "staging failed");
                    break;
                }
                int rc = E.admit(ctx, 0, r->n, d_delta + 64ull * free_idx, nullptr, -1, false,
                                 cfg.want_witness != 0, std::max<uint64_t>(budget, 1), &slot);
                if (rc) {
                    fail_all(rc, "admission failed: " + ctx.err);
                    break;
                }
                if (slot < 0) {
                    // memory is not free yet: the rest waits for the next harvest, in order
                    std::lock_guard<std::mutex> lk(mu);
                    for (size_t j = admit_now.size(); j > k; --j) large_q.push_front(admit_now[j - 1]);
                    break;
                }
                r->slot = slot;
                in_slot[slot] = r;
            }
            if (E.running() == 0) continue;
            if (E.tick(ctx, E.tick_levels)) {
                fail_all(SSW_ERR_DEVICE, "tick failed: " + ctx.err);
                break;
            }
            // harvest
            for (int s = 0; s < E.num; ++s) {
                Request *r = in_slot[s];
                if (!r) continue;
                const LargeInstance &I = E.hinst[s];
                if (I.status != kLFound && I.status != kLNoSync && I.status != kLError) continue;
                const int32_t length = I.status == kLError ? I.length : (I.status == kLFound ? I.length : -1);
                const int64_t expanded = (int64_t)I.expanded;
                bool wv = false;
                if (cfg.want_witness && I.status == kLFound) {
                    if (E.witness(ctx, s, d_wit + (size_t)stride * s, stride) == SSW_OK &&
                        hipMemcpy(hwit.data(), d_wit + (size_t)stride * s, stride, hipMemcpyDeviceToHost) ==
                            hipSuccess)
                        wv = true;
                }
                E.retire(s);
                E.clear_slot(ctx, s);
                in_slot[s] = nullptr;
                std::lock_guard<std::mutex> lk(mu);
                finish(r, length, expanded, wv ? hwit.data() : nullptr, wv);
                cv_done.notify_all();
            }
        }
        hipFree(d_delta);
        hipFree(d_wit);
    }
};

extern "C" {

int ssw_server_create(ssw_context *ctx, const ssw_server_config *cfg, ssw_server **out) {
    if (!ctx || !cfg || !out) return SSW_ERR_ARG;
    *out = nullptr;
    if (cfg->small_n_max < 0 || cfg->small_n_max > 20) return ctx->fail(SSW_ERR_ARG, "small_n_max must be <= 20");
    ssw_server *s = new (std::nothrow) ssw_server();
    if (!s) return SSW_ERR_NOMEM;
    s->owner = ctx;
    s->cfg = *cfg;
    if (s->cfg.levels_per_tick <= 0) s->cfg.levels_per_tick = 16;
    if (s->cfg.max_large <= 0 || s->cfg.max_large > kLargeMaxInstances) s->cfg.max_large = kLargeMaxInstances;
    const uint64_t half = ctx->budget / 2;
    int rc = ssw_create(ctx->device, half, 0, &s->small_ctx);
    if (rc == SSW_OK) rc = ssw_create(ctx->device, half, 0, &s->large_ctx);
    if (rc) {
        if (s->small_ctx) ssw_destroy(s->small_ctx);
        delete s;
        return ctx->fail(rc, "server contexts: %s", ssw_last_error(nullptr));
    }
    // the small engine uses the tuning of the owning context
    std::memcpy(s->small_ctx->tune, ctx->tune, sizeof ctx->tune);
    s->small_thread = std::thread([s] { s->small_loop(); });
    s->large_thread = std::thread([s] { s->large_loop(); });
    *out = s;
    return SSW_OK;
}

int ssw_server_submit(ssw_server *s, int32_t n, const int32_t *a, const int32_t *b, uint64_t *ticket) {
    if (!s || !a || !b || !ticket) return SSW_ERR_ARG;
    if (n < 1 || n > SSW_MAX_N) return SSW_ERR_ARG;
    for (int q = 0; q < n; ++q)
        if (a[q] < 0 || a[q] >= n || b[q] < 0 || b[q] >= n) return SSW_ERR_ARG;
    Request *r = new (std::nothrow) Request();
    if (!r) return SSW_ERR_NOMEM;
    r->n = n;
    for (int q = 0; q < n; ++q) {
        r->delta[q] = (uint8_t)a[q];
        r->delta[n + q] = (uint8_t)b[q];
    }
    std::memset(&r->res, 0, sizeof r->res);
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) {
        delete r;
        return SSW_ERR_STATE;
    }
    if (n > s->cfg.small_n_max) {
        const uint64_t budget = s->cfg.mem_budget ? s->cfg.mem_budget : s->large_ctx->budget / 2;
        if (LargeEngine::admit_bytes(n, s->cfg.want_witness != 0, 1ull << 16) > budget) {
            delete r;
            ++s->rejected;
            return SSW_ERR_BUDGET;
        }
    }
    r->ticket = s->next_ticket++;
    r->t_submit = Clock::now();
    if (!s->any_submit) {
        s->first_submit = r->t_submit;
        s->any_submit = true;
    }
    s->reqs[r->ticket] = r;
    ++s->submitted;
    (n <= s->cfg.small_n_max ? s->small_q : s->large_q).push_back(r);
    s->cv_work.notify_all();
    *ticket = r->ticket;
    return SSW_OK;
}

int ssw_server_poll(ssw_server *s, uint64_t ticket, ssw_result *out) {
    if (!s || !out) return SSW_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return SSW_ERR_NOTFOUND;
    if (!it->second->done) return s->error ? s->error : SSW_PENDING;
    *out = it->second->res;
    s->release(it);
    return SSW_OK;
}

int ssw_server_wait(ssw_server *s, uint64_t ticket, int32_t timeout_ms, ssw_result *out) {
    if (!s || !out) return SSW_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return SSW_ERR_NOTFOUND;
    Request *r = it->second;
    const auto pred = [&] { return r->done || s->error != SSW_OK; };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, pred);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred))
        return SSW_ERR_TIMEOUT;
    if (!r->done) return s->error;
    *out = r->res;
    s->release(s->reqs.find(ticket));
    return SSW_OK;
}

int ssw_server_stats_get(ssw_server *s, ssw_server_stats *out) {
    if (!s || !out) return SSW_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof *out);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->rejected = s->rejected;
    out->subsets_expanded = s->subsets;
    if (!s->lat_us.empty()) {
        std::vector<double> v = s->lat_us;
        std::sort(v.begin(), v.end());
        const auto pct = [&](double p) {
            const size_t k = (size_t)std::min<double>((double)v.size() - 1, std::ceil(p * (double)v.size()) - 1);
            return v[k];
        };
        out->p50_us = pct(0.50);
        out->p99_us = pct(0.99);
        out->max_us = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        out->mean_us = sum / (double)v.size();
        const double span = std::chrono::duration<double>(s->last_done - s->first_submit).count();
        out->automata_per_s = span > 0 ? (double)s->completed / span : 0;
    }
    return SSW_OK;
}

int ssw_server_destroy(ssw_server *s) {
    if (!s) return SSW_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->small_thread.joinable()) s->small_thread.join();
    if (s->large_thread.joinable()) s->large_thread.join();
    for (auto &kv : s->reqs) delete kv.second;
    ssw_destroy(s->small_ctx);
    ssw_destroy(s->large_ctx);
    delete s;
    return SSW_OK;
}

}  // extern "C"
