// ssw_tune.cpp -- per-n autotuner of the batched kernel and its JSON table.
//
// For every n the tuner solves `samples` random automata (the library's
// seeded generator) under candidate launch configurations and keeps the one
// with the shortest kernel time (HIP events on the context's stream, best of
// two runs).  It searches one parameter at a time, starting from the current
// table: threads per automaton, workgroup size, work-queue chunk, frontier
// tile in LDS, frontier entries per lane, and (n <= 16) the ring of the
// second launch that takes the automata outgrowing the first one's ring.
#include "ssw_host.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using namespace ssw;

namespace {

struct Timed {
    ssw_tuning t;
    double ms;
};

int time_config(ssw_context &ctx, int n, const ssw_tuning &t, const DeviceBatch &db, double *ms) {
    ctx.tune[n] = t;
    double best = 1e30;
    for (int rep = 0; rep < 2; ++rep) {
        int rc = ctx.solve_device(db, nullptr, false, nullptr);
        if (rc) return rc;
        best = std::min(best, ctx.last_ms);
    }
    *ms = best;
    return SSW_OK;
}

bool valid_config(const ssw_tuning &t) {
    const int G = t.group_lanes;
    if (G != 16 && G != 32 && G != 64 && G != 256) return false;
    if (t.block_threads && (t.block_threads % 64 || t.block_threads > 256 || t.block_threads % G)) return false;
    if (G == 256 && t.block_threads && t.block_threads != 256) return false;
    const int SG = t.spill_group_lanes;
    if (SG != 0 && SG != 16 && SG != 32 && SG != 64 && SG != 128 && SG != 256) return false;
    if (t.narrow_bits != 0 && (t.narrow_bits < 8 || t.narrow_bits > 16)) return false;
    return t.chunk >= 1 && t.arena_lds >= 4 && t.spill_arena_lds >= 0 && t.spill_lane_entries >= 0;
}

// Tries the values of one parameter around `best`, keeps the fastest.
template <class Setter>
int sweep(ssw_context &ctx, int n, const DeviceBatch &db, Timed &best, const std::vector<int> &values, Setter set) {
    for (int v : values) {
        ssw_tuning t = best.t;
        set(t, v);
        if (!valid_config(t) || std::memcmp(&t, &best.t, sizeof t) == 0) continue;
        double ms = 0;
        const int rc = time_config(ctx, n, t, db, &ms);
        if (rc == SSW_ERR_ARG) continue;  // does not fit (LDS)
        if (rc) return rc;
        if (ms < best.ms) best = Timed{t, ms};
    }
    ctx.tune[n] = best.t;
    return SSW_OK;
}

// Coordinate search over the launch parameters for automata of size n on db.
int tune_one(ssw_context &ctx, int n, const DeviceBatch &db, Timed *out) {
    int rc;
    Timed best{ctx.tune[n], 0};
    if ((rc = time_config(ctx, n, best.t, db, &best.ms))) return rc;
    const bool small = n <= kMaxLdsN;
    const std::vector<int> lanes = small ? std::vector<int>{16, 32, 64} : std::vector<int>{64, 256};
    if ((rc = sweep(ctx, n, db, best, lanes, [](ssw_tuning &t, int v) {
             t.group_lanes = v;
             if (v == 256) t.block_threads = 256;
             else if (t.block_threads == 256 && v != 256) t.block_threads = 0;
         })))
        return rc;
    if (best.t.group_lanes != 256 &&
        (rc = sweep(ctx, n, db, best, {0, 64, 128, 192, 256}, [](ssw_tuning &t, int v) { t.block_threads = v; })))
        return rc;
    if ((rc = sweep(ctx, n, db, best, {1, 2, 4, 8, 16}, [](ssw_tuning &t, int v) { t.chunk = v; }))) return rc;
    const std::vector<int> tiles = small ? std::vector<int>{256, 512, 1024, 2048, 4096}
                                         : std::vector<int>{1024, 2048, 4096, 8192};
    if ((rc = sweep(ctx, n, db, best, tiles, [](ssw_tuning &t, int v) { t.arena_lds = v; }))) return rc;
    if (small && (rc = sweep(ctx, n, db, best, {1, 2, 4}, [](ssw_tuning &t, int v) { t.lane_entries = v; })))
        return rc;
    // n > 16: automata with |a(Q) | b(Q)| <= 16 first through the relabelled
    // n <= 16 solver
    if (!small && (rc = sweep(ctx, n, db, best, {0, 16}, [](ssw_tuning &t, int v) { t.narrow_bits = v; })))
        return rc;
    if (small) {
        // ring spill: automata that outgrow a small first-launch ring go to a
        // second launch with a large ring; then the first ring can shrink
        // joint moves (spill ring, first ring): spilling only pays with a
        // small first ring, which alone would move wide frontiers to HBM
        if ((rc = sweep(ctx, n, db, best, {0, 4096 * 8192 + 0, 4096 * 8192 + 256, 4096 * 8192 + 512,
                                           4096 * 8192 + 1024, 2048 * 8192 + 512, 8192 * 8192 + 512},
                        [](ssw_tuning &t, int v) {
                            t.spill_arena_lds = v / 8192;
                            if (v % 8192) t.arena_lds = v % 8192;
                            if (t.spill_arena_lds && !t.spill_lane_entries) t.spill_lane_entries = 4;
                        })))
            return rc;
        if (best.t.spill_arena_lds > 0) {
            if ((rc = sweep(ctx, n, db, best, {128, 256, 512, 1024},
                            [](ssw_tuning &t, int v) { t.arena_lds = v; })))
                return rc;
            if ((rc = sweep(ctx, n, db, best, {1, 2, 4}, [](ssw_tuning &t, int v) { t.lane_entries = v; })))
                return rc;
            if ((rc = sweep(ctx, n, db, best, {1, 2, 4, 8}, [](ssw_tuning &t, int v) { t.spill_lane_entries = v; })))
                return rc;
            if ((rc = sweep(ctx, n, db, best, {0, 128, 256}, [](ssw_tuning &t, int v) { t.spill_group_lanes = v; })))
                return rc;
            if (best.t.spill_group_lanes > 64 &&
                (rc = sweep(ctx, n, db, best, {1, 2, 4, 8}, [](ssw_tuning &t, int v) { t.spill_lane_entries = v; })))
                return rc;
            if (n > 8 && (rc = sweep(ctx, n, db, best, {0, n - 1, n - 2}, [](ssw_tuning &t, int v) {
                              t.narrow_bits = v >= 8 ? v : 0;
                          })))
                return rc;
            if (best.t.narrow_bits &&
                (rc = sweep(ctx, n, db, best, {0, 64, 128, 192, 256}, [](ssw_tuning &t, int v) { t.block_threads = v; })))
                return rc;
        }
    }
    *out = best;
    return SSW_OK;
}

void write_json(FILE *f, ssw_context &ctx, int n_lo, int n_hi, int64_t samples, const std::vector<double> &ms) {
    hipDeviceProp_t prop;
    hipGetDeviceProperties(&prop, ctx.device);
    fprintf(f, "{\n  \"device\": \"%s\",\n  \"cus\": %d,\n  \"samples\": %lld,\n  \"tuning\": {\n", prop.gcnArchName,
            ctx.cus, (long long)samples);
    for (int n = n_lo; n <= n_hi; ++n) {
        const ssw_tuning &t = ctx.tune[n];
        fprintf(f,
                "    \"%d\": {\"group_lanes\": %d, \"block_threads\": %d, \"chunk\": %d, \"arena_lds\": %d, "
                "\"lane_entries\": %d, \"spill_arena_lds\": %d, \"spill_lane_entries\": %d, "
                "\"spill_group_lanes\": %d, \"narrow_bits\": %d, \"ms\": %.4f}%s\n",
                n, t.group_lanes, t.block_threads, t.chunk, t.arena_lds, t.lane_entries, t.spill_arena_lds,
                t.spill_lane_entries, t.spill_group_lanes, t.narrow_bits, ms[n - n_lo], n < n_hi ? "," : "");
    }
    fprintf(f, "  }\n}\n");
}

// Reads "key": number inside the object text [p, end).
bool read_int(const std::string &obj, const char *key, int *out) {
    const std::string k = std::string("\"") + key + "\"";
    const size_t at = obj.find(k);
    if (at == std::string::npos) return false;
    size_t c = obj.find(':', at + k.size());
    if (c == std::string::npos) return false;
    *out = (int)strtol(obj.c_str() + c + 1, nullptr, 10);
    return true;
}

}  // namespace

extern "C" {

int ssw_autotune(ssw_context *ctx, int32_t n_lo, int32_t n_hi, int64_t samples, uint64_t seed, const char *out_path) {
    if (!ctx) return SSW_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (n_lo < 1 || n_hi > SSW_MAX_N || n_lo > n_hi || samples < 1)
        return ctx->fail(SSW_ERR_ARG, "autotune range [%d,%d] / samples %lld", n_lo, n_hi, (long long)samples);
    if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(SSW_ERR_DEVICE, "hipSetDevice");
    std::vector<double> best_ms(n_hi - n_lo + 1, 0.0);
    for (int n = n_lo; n <= n_hi; ++n) {
        // random automata of size n (n > 20 uses fewer: the frontier work grows with n)
        const int64_t count = n <= 20 ? samples : std::max<int64_t>(1024, samples / 8);
        std::vector<int32_t> a((size_t)count * n), b((size_t)count * n);
        ssw_generate_uniform(seed + (uint64_t)n, n, 0, count, a.data(), b.data());
        std::vector<uint8_t> packed((size_t)count * 2 * n);
        for (int64_t i = 0; i < count; ++i)
            for (int q = 0; q < n; ++q) {
                packed[(size_t)i * 2 * n + q] = (uint8_t)a[(size_t)i * n + q];
                packed[(size_t)i * 2 * n + n + q] = (uint8_t)b[(size_t)i * n + q];
            }
        int rc;
        if ((rc = ctx->ensure(ctx->d_delta, packed.size(), "delta"))) return rc;
        if ((rc = ctx->ensure(ctx->d_len, 4ull * count, "length"))) return rc;
        if (hipMemcpyAsync(ctx->d_delta.p, packed.data(), packed.size(), hipMemcpyHostToDevice, ctx->stream) !=
                hipSuccess ||
            hipStreamSynchronize(ctx->stream) != hipSuccess)
            return ctx->fail(SSW_ERR_DEVICE, "H2D");
        DeviceBatch db;
        db.count = count;
        db.uniform_n = n;
        db.delta = (const uint8_t *)ctx->d_delta.p;
        db.length = (int32_t *)ctx->d_len.p;
        Timed best;
        if ((rc = tune_one(*ctx, n, db, &best))) return rc;
        ctx->tune[n] = best.t;
        best_ms[n - n_lo] = best.ms;
    }
    if (out_path) {
        FILE *f = fopen(out_path, "w");
        if (!f) return ctx->fail(SSW_ERR_ARG, "cannot write %s", out_path);
        write_json(f, *ctx, n_lo, n_hi, samples, best_ms);
        fclose(f);
    }
    return SSW_OK;
}

int ssw_autotune_batch(ssw_context *ctx, const ssw_batch *b, int32_t memory) {
    if (!ctx || !b) return SSW_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (memory != SSW_DEVICE || b->uniform_n < 1 || b->uniform_n > SSW_MAX_N || b->count < 1 || b->witness)
        return ctx->fail(SSW_ERR_ARG, "autotune_batch needs a device batch of one n without witness");
    if (hipSetDevice(ctx->device) != hipSuccess) return ctx->fail(SSW_ERR_DEVICE, "hipSetDevice");
    DeviceBatch db;
    db.count = b->count;
    db.uniform_n = b->uniform_n;
    db.offset = b->offset;
    db.delta = b->delta;
    db.length = b->length;
    db.sync = b->sync;
    db.expanded = b->expanded;
    db.bound = b->bound;
    Timed best;
    return tune_one(*ctx, b->uniform_n, db, &best);
}

int ssw_load_tuning(ssw_context *ctx, const char *path) {
    if (!ctx || !path) return SSW_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    FILE *f = fopen(path, "r");
    if (!f) return ctx->fail(SSW_ERR_ARG, "cannot read %s", path);
    std::string text;
    char buf[4096];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, k);
    fclose(f);
    const size_t root = text.find("\"tuning\"");
    if (root == std::string::npos) return ctx->fail(SSW_ERR_ARG, "%s: no \"tuning\" object", path);
    int loaded = 0;
    size_t pos = text.find('{', root);
    while (pos != std::string::npos) {
        const size_t q1 = text.find('"', pos + 1);
        if (q1 == std::string::npos) break;
        const size_t q2 = text.find('"', q1 + 1);
        const size_t ob = text.find('{', q2);
        const size_t cb = text.find('}', ob);
        if (q2 == std::string::npos || ob == std::string::npos || cb == std::string::npos) break;
        const int n = atoi(text.substr(q1 + 1, q2 - q1 - 1).c_str());
        const std::string obj = text.substr(ob, cb - ob + 1);
        ssw_tuning t = (n >= 1 && n <= SSW_MAX_N) ? ctx->tune[n] : ssw_tuning{64, 0, 2, 2048, 1};
        read_int(obj, "group_lanes", &t.group_lanes);
        read_int(obj, "block_threads", &t.block_threads);
        read_int(obj, "chunk", &t.chunk);
        read_int(obj, "arena_lds", &t.arena_lds);
        read_int(obj, "lane_entries", &t.lane_entries);
        read_int(obj, "spill_arena_lds", &t.spill_arena_lds);
        read_int(obj, "spill_lane_entries", &t.spill_lane_entries);
        read_int(obj, "spill_group_lanes", &t.spill_group_lanes);
        read_int(obj, "narrow_bits", &t.narrow_bits);
        if (n >= 1 && n <= SSW_MAX_N && valid_config(t)) {
            ctx->tune[n] = t;
            ++loaded;
        }
        pos = text.find(',', cb);
        if (pos == std::string::npos) break;
        if (text.find('"', pos) == std::string::npos) break;
    }
    return loaded ? SSW_OK : ctx->fail(SSW_ERR_ARG, "%s: no tuning entries", path);
}

}  // extern "C"
