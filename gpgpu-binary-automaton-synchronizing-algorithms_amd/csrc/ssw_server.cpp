// ssw_server.cpp -- the serving entry points (ssw_server_* in ssw.h): the
// request loop of ssw_server_core.cpp bound to the HIP solvers.
//
//   * small requests (n <= small_n_max): batched rounds on `small_workers`
//     private contexts (each its own streams), so one round can run while the
//     next is gathered and started;
//   * large requests: the device-wide engine of a private context, one
//     captured graph of `levels_per_tick` level steps per tick, instances
//     admitted and retired between ticks.
#include "ssw_host.h"
#include "ssw_server_core.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>

using namespace ssw;

namespace {

constexpr int kMaxSmallWorkers = 4;

// One batched launch per round, on a context of its own.
struct HipSmall final : SmallBackend {
    ssw_context *ctx = nullptr;
    int solve(size_t m, const uint8_t *n_per, const uint64_t *offs, const uint8_t *delta, int32_t *length,
              int64_t *expanded, uint8_t *witness, int stride, std::string *err) override {
        ssw_batch b;
        std::memset(&b, 0, sizeof b);
        b.count = (int64_t)m;
        b.n_per = n_per;
        b.offset = offs;
        b.delta = delta;
        b.length = length;
        b.expanded = expanded;
        b.witness = witness;
        b.witness_stride = stride;
        b.bound = -1;
        std::lock_guard<std::mutex> lk(ctx->mu);
        const int rc = ctx->solve_entry(&b, SSW_HOST, kModeBatch, nullptr);
        if (rc) *err = ctx->err;
        return rc;
    }
};

// The device-wide engine of a private context.  Every method but the
// budget queries runs on the server's large-engine thread, which holds the
// context from setup() to teardown().
struct HipLarge final : LargeBackend {
    ssw_context *ctx = nullptr;
    int levels = 16;  // most levels per graph replay
    // levels of the next replay: halved while a replay takes longer than
    // kTickTarget and doubled (up to `levels`) while it takes less than a
    // quarter of it, so a small-request round queued behind a replay on a
    // shared hardware queue waits about kTickTarget at most even when the
    // running instances' levels are wide
    int cur = 16;
    static constexpr double kTickTarget = 2e-3;
    std::unique_lock<std::mutex> hold;
    uint8_t *d_delta = nullptr, *d_wit = nullptr;
    std::vector<uint8_t> hwit;
    static constexpr int kStride = (int)sizeof(((ssw_result *)nullptr)->witness);

    int setup(std::string *err) override {
        hold = std::unique_lock<std::mutex>(ctx->mu);
        LargeEngine &E = ctx->large;
        hwit.assign(kStride, 0);
        if (hipSetDevice(ctx->device) != hipSuccess || E.setup(*ctx) ||
            E.set_outputs(*ctx, nullptr, nullptr, nullptr, nullptr, nullptr) ||
            hipMalloc((void **)&d_delta, 64ull * kLargeMaxInstances) != hipSuccess ||
            hipMalloc((void **)&d_wit, (size_t)kStride * kLargeMaxInstances) != hipSuccess) {
            *err = ctx->err;
            return SSW_ERR_DEVICE;
        }
        E.tick_levels = std::max(1, levels);
        cur = E.tick_levels;
        return SSW_OK;
    }
    void teardown() override {
        if (d_delta) hipFree(d_delta);
        if (d_wit) hipFree(d_wit);
        d_delta = d_wit = nullptr;
        if (hold.owns_lock()) hold.unlock();
    }
    int capacity() const override { return kLargeMaxInstances; }
    uint64_t default_budget() const override { return ctx->budget / 2; }
    uint64_t admit_bytes(int n, bool witness) const override { return LargeEngine::admit_bytes(n, witness); }
    uint64_t bytes_in_use() const override { return ctx->large.bytes_in_use; }
    int running() const override { return ctx->large.running(); }

    int admit(int n, const uint8_t *delta, bool witness, uint64_t budget, int *slot, std::string *err) override {
        LargeEngine &E = ctx->large;
        *slot = -1;
        // the transitions are staged where large_init reads them: at the slot
        // admit() will take, the first free one
        int free_idx = -1;
        for (int s = 0; s < kLargeMaxInstances; ++s)
            if (!E.slots[s].used) {
                free_idx = s;
                break;
            }
        if (free_idx < 0) return SSW_OK;
        // on the engine's own stream: a copy on the legacy stream would wait
        // on (and be refused beside) other contexts' streams in the process
        hipError_t e = hipMemcpyAsync(d_delta + 64ull * free_idx, delta, 2 * (size_t)n, hipMemcpyHostToDevice,
                                      ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) {
            *err = std::string("staging the transitions failed: ") + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
            return SSW_ERR_DEVICE;
        }
        const int rc = E.admit(*ctx, 0, n, d_delta + 64ull * free_idx, nullptr, -1, false, witness, budget, slot);
        if (rc) *err = ctx->err;
        return rc;
    }
    int tick(std::string *err) override {
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = ctx->large.tick(*ctx, cur);
        if (rc) {
            *err = ctx->err;
            return rc;
        }
        const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (dt > kTickTarget && cur > 1)
            cur = std::max(1, cur / 2);
        else if (dt < kTickTarget / 4 && cur < ctx->large.tick_levels)
            cur = std::min(ctx->large.tick_levels, 2 * cur);
        return SSW_OK;
    }
    int harvest(bool witness, std::vector<Finished> *out, std::string *err) override {
        LargeEngine &E = ctx->large;
        for (int s = 0; s < E.num; ++s) {
            if (!E.slots[s].used) continue;
            const LargeInstance &I = E.hinst[s];
            if (I.status != kLFound && I.status != kLNoSync && I.status != kLError) continue;
            Finished f;
            std::memset(&f, 0, sizeof f);
            f.slot = s;
            f.length = I.status == kLNoSync ? -1 : I.length;
            f.expanded = (int64_t)I.expanded;
            if (witness && I.status == kLFound) {
                if (E.witness(*ctx, s, d_wit + (size_t)kStride * s, kStride) == SSW_OK &&
                    hipMemcpyAsync(hwit.data(), d_wit + (size_t)kStride * s, kStride, hipMemcpyDeviceToHost,
                                   ctx->stream) == hipSuccess &&
                    hipStreamSynchronize(ctx->stream) == hipSuccess) {
                    f.witness_valid = true;
                    std::memcpy(f.witness, hwit.data(), kStride);
                }
            }
            E.retire(s);
            if (const int rc = E.clear_slot(*ctx, s)) {
                *err = ctx->err;
                return rc;
            }
            out->push_back(f);
        }
        return SSW_OK;
    }
};

}  // namespace

struct ssw_server {
    ssw_context *small_ctx[kMaxSmallWorkers] = {};
    ssw_context *large_ctx = nullptr;
    HipSmall small[kMaxSmallWorkers];
    HipLarge large;
    std::unique_ptr<ServerCore> core;
    int workers = 0;

    void release_contexts() {
        for (int k = 0; k < kMaxSmallWorkers; ++k)
            if (small_ctx[k]) ssw_destroy(small_ctx[k]);
        if (large_ctx) ssw_destroy(large_ctx);
    }
};

extern "C" {

int ssw_server_create(ssw_context *ctx, const ssw_server_config *cfg, ssw_server **out) {
    if (!ctx || !cfg || !out) return SSW_ERR_ARG;
    *out = nullptr;
    if (cfg->small_n_max < 0 || cfg->small_n_max > 20) return ctx->fail(SSW_ERR_ARG, "small_n_max must be <= 20");
    if (cfg->small_workers < 0 || cfg->small_workers > kMaxSmallWorkers)
        return ctx->fail(SSW_ERR_ARG, "small_workers must be 0..%d", kMaxSmallWorkers);
    ssw_server *s = new (std::nothrow) ssw_server();
    if (!s) return SSW_ERR_NOMEM;
    ssw_server_config c = *cfg;
    if (c.levels_per_tick <= 0) c.levels_per_tick = 16;
    if (c.max_large <= 0 || c.max_large > kLargeMaxInstances) c.max_large = kLargeMaxInstances;
    if (c.small_workers == 0) c.small_workers = 2;
    s->workers = c.small_workers;
    // the small engines share half of the owner's budget, the large one has
    // the other half
    const uint64_t half = ctx->budget / 2;
    int rc = SSW_OK;
    for (int k = 0; k < s->workers && rc == SSW_OK; ++k)
        rc = ssw_create(ctx->device, half / (uint64_t)s->workers, 0, &s->small_ctx[k]);
    if (rc == SSW_OK) rc = ssw_create(ctx->device, half, 0, &s->large_ctx);
    if (rc) {
        const std::string why = ssw_last_error(nullptr);
        s->release_contexts();
        delete s;
        return ctx->fail(rc, "server contexts: %s", why.c_str());
    }
    std::vector<SmallBackend *> smalls;
    for (int k = 0; k < s->workers; ++k) {
        // the small engines use the tuning of the owning context
        std::memcpy(s->small_ctx[k]->tune, ctx->tune, sizeof ctx->tune);
        s->small[k].ctx = s->small_ctx[k];
        smalls.push_back(&s->small[k]);
    }
    s->large.ctx = s->large_ctx;
    s->large.levels = c.levels_per_tick;
    s->core.reset(new ServerCore(c, smalls, &s->large));
    s->core->start();
    *out = s;
    return SSW_OK;
}

int ssw_server_submit(ssw_server *s, int32_t n, const int32_t *a, const int32_t *b, uint64_t *ticket) {
    if (!s) return SSW_ERR_ARG;
    return s->core->submit(n, a, b, ticket);
}

int ssw_server_poll(ssw_server *s, uint64_t ticket, ssw_result *out) {
    if (!s) return SSW_ERR_ARG;
    return s->core->poll(ticket, out);
}

int ssw_server_wait(ssw_server *s, uint64_t ticket, int32_t timeout_ms, ssw_result *out) {
    if (!s) return SSW_ERR_ARG;
    return s->core->wait(ticket, timeout_ms, out);
}

int ssw_server_stats_get(ssw_server *s, ssw_server_stats *out) {
    if (!s) return SSW_ERR_ARG;
    return s->core->stats(out);
}

int ssw_server_cancel(ssw_server *s, uint64_t ticket) {
    if (!s) return SSW_ERR_ARG;
    return s->core->cancel(ticket);
}

const char *ssw_server_last_error(ssw_server *s) {
    if (!s) return "";
    return s->core->last_error();
}

int ssw_server_destroy(ssw_server *s) {
    if (!s) return SSW_ERR_ARG;
    s->core->shutdown();
    s->core.reset();
    s->release_contexts();
    delete s;
    return SSW_OK;
}

}  // extern "C"
