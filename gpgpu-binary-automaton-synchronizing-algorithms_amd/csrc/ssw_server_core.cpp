// ssw_server_core.cpp -- the continuous-batching request loop (see
// ssw_server_core.h), independent of the device.
//
// Small requests (n <= small_n_max) are taken by one of several small-engine
// threads, each solving whatever is queued when it becomes free as one round
// (at most kMaxRound requests): a request that arrives while one round is on
// the GPU is picked up by an idle thread at once instead of waiting for that
// round to end.  Large requests join the large engine between its ticks, in
// arrival order, as soon as a slot and their memory fit the budget, and
// finished instances leave at the same points, so their work overlaps on the
// device.  Every request is solved by the same kernels as a lone solve, so its
// result does not depend on what else is in flight.
#include "ssw_server_core.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace ssw {

ServerCore::ServerCore(const ssw_server_config &cfg, std::vector<SmallBackend *> small, LargeBackend *large)
    : cfg_(cfg), small_(std::move(small)), large_(large) {}

ServerCore::~ServerCore() { shutdown(); }

void ServerCore::start() {
    if (started_) return;
    started_ = true;
    for (SmallBackend *be : small_) threads_.emplace_back([this, be] { small_loop(be); });
    threads_.emplace_back([this] { large_loop(); });
}

uint64_t ServerCore::large_budget() const {
    return cfg_.mem_budget ? cfg_.mem_budget : large_->default_budget();
}

void ServerCore::finish(ServerRequest *r, int32_t length, int64_t expanded, const uint8_t *wit, bool wit_valid) {
    r->t_done = ServerClock::now();
    r->res.length = length;
    r->res.n = r->n;
    r->res.expanded = expanded;
    r->res.sync = length >= 0 ? 1 : 0;
    r->res.witness_valid = wit_valid ? 1 : 0;
    if (wit_valid) std::memcpy(r->res.witness, wit, sizeof r->res.witness);
    r->res.latency_us = std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count();
    r->done = true;
    if (lat_us_.size() < kLatencyWindow) {
        lat_us_.push_back(r->res.latency_us);
    } else {
        lat_us_[lat_next_] = r->res.latency_us;
        lat_next_ = (lat_next_ + 1) % kLatencyWindow;
    }
    ++completed_;
    subsets_ += (uint64_t)std::max<int64_t>(expanded, 0);
    last_done_ = r->t_done;
}

// Hands the result of a finished request to one caller; the ticket is then
// forgotten.  The request itself is freed by the last caller still blocked in
// wait on it (they wake, see it collected and return SSW_ERR_NOTFOUND), or
// here when there is none.  Called with mu_ held.
void ServerCore::take(ServerRequest *r, ssw_result *out) {
    *out = r->res;
    r->collected = true;
    reqs_.erase(r->ticket);
    if (r->waiters == 0)
        delete r;
    else
        cv_done_.notify_all();
}

void ServerCore::fail_all(int code, const std::string &msg) {
    std::lock_guard<std::mutex> lk(mu_);
    if (error_ == SSW_OK) {
        error_ = code;
        error_msg_ = msg;
    }
    stop_ = true;
    cv_work_.notify_all();
    cv_done_.notify_all();
}

// ---- small engine: one round at a time per thread ---------------------------
void ServerCore::small_loop(SmallBackend *be) {
    std::vector<ServerRequest *> round;
    std::vector<uint8_t> n_per, delta, wit;
    std::vector<uint64_t> offs;
    std::vector<int32_t> length;
    std::vector<int64_t> expanded;
    const int stride = (int)sizeof(((ssw_result *)nullptr)->witness);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_work_.wait(lk, [&] { return stop_ || !small_q_.empty(); });
            if (small_q_.empty() || error_ != SSW_OK) return;  // stopped and drained, or failed
            const size_t m = std::min(small_q_.size(), kMaxRound);
            round.assign(small_q_.begin(), small_q_.begin() + (std::ptrdiff_t)m);
            small_q_.erase(small_q_.begin(), small_q_.begin() + (std::ptrdiff_t)m);
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
        if (cfg_.want_witness) wit.assign(m * (size_t)stride, 0);
        std::string err;
        const int rc = be->solve(m, n_per.data(), offs.data(), delta.data(), length.data(), expanded.data(),
                                 cfg_.want_witness ? wit.data() : nullptr, stride, &err);
        if (rc) {
            fail_all(rc, err);
            return;
        }
        std::lock_guard<std::mutex> lk(mu_);
        for (size_t i = 0; i < m; ++i)
            finish(round[i], length[i], expanded[i], cfg_.want_witness ? wit.data() + i * stride : nullptr,
                   cfg_.want_witness && length[i] >= 0);
        cv_done_.notify_all();
    }
}

// ---- large engine: admit / tick / harvest ------------------------------------
void ServerCore::large_loop() {
    std::string err;
    if (large_->setup(&err)) {
        fail_all(SSW_ERR_DEVICE, "large engine set-up failed: " + err);
        large_->teardown();
        return;
    }
    const bool wit = cfg_.want_witness != 0;
    const uint64_t budget = std::max<uint64_t>(large_budget(), 1);
    const int max_in_flight = std::max(1, std::min(large_->capacity(), cfg_.max_large));
    std::vector<ServerRequest *> in_slot((size_t)large_->capacity(), nullptr);
    std::vector<LargeBackend::Finished> done;
    for (;;) {
        std::vector<ServerRequest *> admit_now;
        {
            std::unique_lock<std::mutex> lk(mu_);
            if (large_->running() == 0) cv_work_.wait(lk, [&] { return stop_ || !large_q_.empty(); });
            if (error_ != SSW_OK) break;
            if (stop_ && large_q_.empty() && large_->running() == 0) break;
            // admit in arrival order while slots and memory allow
            uint64_t planned = large_->bytes_in_use();
            while (!large_q_.empty()) {
                ServerRequest *r = large_q_.front();
                const int in_flight = large_->running() + (int)admit_now.size();
                if (in_flight >= max_in_flight) break;
                const uint64_t need = large_->admit_bytes(r->n, wit);
                if (planned + need > budget && in_flight > 0) break;
                planned += need;
                admit_now.push_back(r);
                large_q_.pop_front();
            }
        }
        bool failed = false;
        for (size_t k = 0; k < admit_now.size(); ++k) {
            ServerRequest *r = admit_now[k];
            int slot = -1;
            if (const int rc = large_->admit(r->n, r->delta, wit, budget, &slot, &err)) {
                fail_all(rc, "admission failed: " + err);
                failed = true;
                break;
            }
            if (slot < 0) {
                // memory is not free yet: the rest waits for the next harvest, in order
                std::lock_guard<std::mutex> lk(mu_);
                for (size_t j = admit_now.size(); j > k; --j) large_q_.push_front(admit_now[j - 1]);
                break;
            }
            in_slot[(size_t)slot] = r;
        }
        if (failed) break;
        if (large_->running() == 0) continue;
        if (large_->tick(&err)) {
            fail_all(SSW_ERR_DEVICE, "tick failed: " + err);
            break;
        }
        done.clear();
        if (large_->harvest(wit, &done, &err)) {
            fail_all(SSW_ERR_DEVICE, "harvest failed: " + err);
            break;
        }
        if (done.empty()) continue;
        std::lock_guard<std::mutex> lk(mu_);
        for (const LargeBackend::Finished &f : done) {
            ServerRequest *r = in_slot[(size_t)f.slot];
            in_slot[(size_t)f.slot] = nullptr;
            if (r) finish(r, f.length, f.expanded, f.witness_valid ? f.witness : nullptr, f.witness_valid);
        }
        cv_done_.notify_all();
    }
    large_->teardown();
}

// ---- API ----------------------------------------------------------------------
int ServerCore::submit(int32_t n, const int32_t *a, const int32_t *b, uint64_t *ticket) {
    if (!a || !b || !ticket) return SSW_ERR_ARG;
    if (n < 1 || n > SSW_MAX_N) return SSW_ERR_ARG;
    for (int q = 0; q < n; ++q)
        if (a[q] < 0 || a[q] >= n || b[q] < 0 || b[q] >= n) return SSW_ERR_ARG;
    const bool small = n <= cfg_.small_n_max;
    // a request that can never fit the budget is refused here, by the same
    // figure admission uses
    const bool too_big = !small && large_->admit_bytes(n, cfg_.want_witness != 0) > large_budget();
    ServerRequest *r = too_big ? nullptr : new (std::nothrow) ServerRequest();
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) {
        delete r;
        return error_ != SSW_OK ? error_ : SSW_ERR_STATE;
    }
    if (too_big) {
        ++rejected_;
        return SSW_ERR_BUDGET;
    }
    if (!r) return SSW_ERR_NOMEM;
    r->n = n;
    for (int q = 0; q < n; ++q) {
        r->delta[q] = (uint8_t)a[q];
        r->delta[n + q] = (uint8_t)b[q];
    }
    std::memset(&r->res, 0, sizeof r->res);
    r->ticket = next_ticket_++;
    r->t_submit = ServerClock::now();
    if (!any_submit_) {
        first_submit_ = r->t_submit;
        any_submit_ = true;
    }
    reqs_[r->ticket] = r;
    ++submitted_;
    (small ? small_q_ : large_q_).push_back(r);
    cv_work_.notify_all();
    *ticket = r->ticket;
    return SSW_OK;
}

int ServerCore::poll(uint64_t ticket, ssw_result *out) {
    if (!out) return SSW_ERR_ARG;
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return SSW_ERR_NOTFOUND;
    ServerRequest *r = it->second;
    if (!r->done) return error_ != SSW_OK ? error_ : SSW_PENDING;
    take(r, out);
    return SSW_OK;
}

int ServerCore::wait(uint64_t ticket, int32_t timeout_ms, ssw_result *out) {
    if (!out) return SSW_ERR_ARG;
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return SSW_ERR_NOTFOUND;
    ServerRequest *r = it->second;
    // the request stays allocated while this caller sleeps: take() and
    // shutdown() leave it to the last waiter
    ++r->waiters;
    ++waiting_;
    const auto ready = [&] { return r->done || r->collected || r->cancelled || error_ != SSW_OK || closed_; };
    bool in_time = true;
    if (timeout_ms < 0)
        cv_done_.wait(lk, ready);
    else
        in_time = cv_done_.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready);
    --r->waiters;
    --waiting_;
    int rc;
    if (r->cancelled) {
        rc = SSW_ERR_CANCELLED;
        if (r->waiters == 0) delete r;
    } else if (r->collected) {
        // another caller took the result while this one slept
        rc = SSW_ERR_NOTFOUND;
        if (r->waiters == 0) delete r;
    } else if (r->done) {
        take(r, out);
        rc = SSW_OK;
    } else if (!in_time) {
        rc = SSW_ERR_TIMEOUT;  // the ticket stays valid
    } else {
        rc = error_ != SSW_OK ? error_ : SSW_ERR_STATE;
    }
    if (closed_ && waiting_ == 0) cv_idle_.notify_all();
    return rc;
}

int ServerCore::cancel(uint64_t ticket) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) return SSW_ERR_NOTFOUND;
    ServerRequest *r = it->second;
    if (r->done) return SSW_ERR_STATE;
    // only from its queue: a request an engine has taken is being solved
    std::deque<ServerRequest *> &q = r->n <= cfg_.small_n_max ? small_q_ : large_q_;
    const auto qi = std::find(q.begin(), q.end(), r);
    if (qi == q.end()) return SSW_ERR_STATE;
    q.erase(qi);
    reqs_.erase(it);
    r->cancelled = true;
    if (r->waiters == 0)
        delete r;  // else the last waiter frees it
    else
        cv_done_.notify_all();
    return SSW_OK;
}

const char *ServerCore::last_error() {
    std::lock_guard<std::mutex> lk(mu_);
    return error_ != SSW_OK ? error_msg_.c_str() : "";
}

int ServerCore::stats(ssw_server_stats *out) {
    if (!out) return SSW_ERR_ARG;
    std::lock_guard<std::mutex> lk(mu_);
    std::memset(out, 0, sizeof *out);
    out->submitted = submitted_;
    out->completed = completed_;
    out->rejected = rejected_;
    out->subsets_expanded = subsets_;
    if (!lat_us_.empty()) {
        std::vector<double> v = lat_us_;
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
        const double span = std::chrono::duration<double>(last_done_ - first_submit_).count();
        out->automata_per_s = span > 0 ? (double)completed_ / span : 0;
    }
    return SSW_OK;
}

void ServerCore::shutdown() {
    if (shut_) return;
    shut_ = true;
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_work_.notify_all();
    for (std::thread &t : threads_)
        if (t.joinable()) t.join();
    threads_.clear();
    std::unique_lock<std::mutex> lk(mu_);
    closed_ = true;
    cv_done_.notify_all();
    cv_idle_.wait(lk, [&] { return waiting_ == 0; });
    for (auto &kv : reqs_) delete kv.second;
    reqs_.clear();
    small_q_.clear();
    large_q_.clear();
}

}  // namespace ssw
