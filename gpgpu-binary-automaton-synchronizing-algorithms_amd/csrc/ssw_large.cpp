// ssw_large.cpp -- host driver of the device-wide level loop (ssw_large.hip).
//
// The engine owns up to kLargeMaxInstances instance slots.  Each admitted
// automaton gets a visited bitmap (2^n bits), a frontier ring and a table in
// HBM; a captured graph of `levels` [large_step, large_plan] pairs is replayed
// until every instance has finished, and between replays the host grows the
// rings of paused instances, produces witnesses and recycles finished slots.
#include "ssw_host.h"

#include <algorithm>
#include <cstdlib>
#include <cstring>

namespace ssw {
hipError_t large_launch_init(LargeInstance *inst, int slot, const uint8_t *delta, const uint64_t *offset,
                             int64_t index, hipStream_t st);
hipError_t large_launch_plan(LargeState *s, LargeInstance *inst, int advance, hipStream_t st);
hipError_t large_launch_step(LargeState *s, LargeInstance *inst, int grid, hipStream_t st);
hipError_t large_launch_regrow(const uint32_t *src, uint64_t src_mask, uint32_t *dst, uint64_t dst_mask,
                               uint64_t from, uint64_t to, hipStream_t st);
hipError_t large_launch_back(const LargeInstance *inst, int slot, int j, int L, const uint32_t *src,
                             uint32_t *dst, int grid, hipStream_t st);
hipError_t large_launch_fwd(const LargeInstance *inst, int slot, int L, const uint32_t *g0, const uint32_t *g1,
                            uint8_t *out, int stride, hipStream_t st);
hipError_t fill_u32(uint32_t *p, uint64_t words, uint32_t value, hipStream_t st);
hipError_t reset_counters(uint32_t *deferred, unsigned long long *violation, unsigned long long *progress,
                          unsigned long long *launch_stats, uint32_t stat_words, hipStream_t st);

#define LG_TRY(expr, what)                                  \
    do {                                                    \
        hipError_t _e = (expr);                             \
        if (_e != hipSuccess) return ctx.hip_fail(_e, what); \
    } while (0)

static uint64_t pow2_at_least(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return p;
}

static size_t state_bytes() {
    return (sizeof(LargeState) + 255) / 256 * 256 + (size_t)kLargeMaxInstances * sizeof(LargeInstance);
}

LargeState *LargeEngine::dstate() const { return (LargeState *)state.p; }
LargeInstance *LargeEngine::dinst() const {
    return (LargeInstance *)((uint8_t *)state.p + (sizeof(LargeState) + 255) / 256 * 256);
}

void LargeEngine::release() {
    for (auto &s : slots) free_slot_buffers(s);
    slots.clear();
    drop_spare();
    for (auto &kv : graphs) hipGraphExecDestroy(kv.second);
    graphs.clear();
    if (state.p) hipFree(state.p);
    state = DevBuf();
}

hipError_t LargeEngine::take(void **p, uint64_t bytes) {
    auto it = spare.find(bytes);
    if (it != spare.end()) {
        *p = it->second;
        spare.erase(it);
        spare_bytes -= bytes;
        return hipSuccess;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess && !spare.empty()) {
        (void)hipGetLastError();
        drop_spare();
        e = hipMalloc(p, bytes);
    }
    return e;
}

void LargeEngine::give(void *p, uint64_t bytes) {
    if (!p) return;
    if (spare_bytes + bytes <= spare_limit) {
        spare.emplace(bytes, p);
        spare_bytes += bytes;
    } else {
        hipFree(p);
    }
}

void LargeEngine::drop_spare() {
    for (auto &kv : spare) hipFree(kv.second);
    spare.clear();
    spare_bytes = 0;
}

void LargeEngine::free_slot_buffers(LargeSlot &s) {
    give(s.ring, 4ull * s.ring_cap);
    give(s.bitmap, std::max<uint64_t>(256, (1ull << s.n) / 8));
    give(s.table, 8192);
    give(s.lvl, 8ull * s.lvl_cap);
    s.ring = nullptr;
    s.bitmap = nullptr;
    s.table = nullptr;
    s.lvl = nullptr;
    bytes_in_use -= s.bytes;
    s.bytes = 0;
}

int LargeEngine::setup(ssw_context &ctx) {
    if (!state.p) {
        int rc = ctx.ensure(state, state_bytes(), "large state");
        if (rc) return rc;
        LG_TRY(hipMemsetAsync(state.p, 0, state_bytes(), ctx.stream), "memset");
        slots.assign(kLargeMaxInstances, LargeSlot());
        hinst.assign(kLargeMaxInstances, LargeInstance());
        num = 0;
        // workgroups of large_step per CU (SSW_LARGE_GRID_PER_CU overrides the
        // default 1: more workgroups only add contention on the bitmaps)
        int per_cu = 1;
        if (const char *e = getenv("SSW_LARGE_GRID_PER_CU")) per_cu = std::max(1, std::min(32, atoi(e)));
        grid = std::max(1, ctx.cus) * per_cu;
        spare_limit = ctx.budget / 4;
        if (const char *e = getenv("SSW_LARGE_LEVELS")) tick_levels = std::max(1, std::min(256, atoi(e)));
    }
    return SSW_OK;
}

int LargeEngine::set_outputs(ssw_context &ctx, int32_t *length, uint8_t *sync, int64_t *expanded,
                             unsigned long long *violation, unsigned long long *progress) {
    LargeState &hs = hstate;
    std::memset(&hs, 0, sizeof hs);
    // only the output pointers are written; the rest is owned by large_plan
    const size_t off = offsetof(LargeState, length);
    hs.length = length;
    hs.sync = sync;
    hs.expanded = expanded;
    hs.violation = violation;
    hs.progress = progress;
    LG_TRY(hipMemcpyAsync((uint8_t *)state.p + off, (uint8_t *)&hs + off, sizeof(LargeState) - off,
                          hipMemcpyHostToDevice, ctx.stream),
           "H2D outputs");
    return SSW_OK;
}

// The first frontier ring of an instance: every subset (plus the root's slot
// in append mode, which keeps all levels) up to 2^16 entries; larger rings
// grow between graph replays.
uint64_t LargeEngine::first_ring_cap(int n, bool append) {
    const uint64_t all = 1ull << n;
    return append ? pow2_at_least(std::min<uint64_t>(all, 1ull << 16) + 1)
                  : pow2_at_least(std::min<uint64_t>(all, 1ull << 16));
}

// Device bytes an instance holds when admitted: bitmap, first ring, tables
// and (append mode) the level-start table.  Admission and the budget checks
// made before it (the server's submit) use this one figure.
uint64_t LargeEngine::admit_bytes(int n, bool append) {
    const uint64_t bm = std::max<uint64_t>(256, (1ull << n) / 8);
    return bm + 4 * first_ring_cap(n, append) + 8192 + (append ? 8ull * ((n - 1) * (n - 1) + 8) : 0);
}

int LargeEngine::admit(ssw_context &ctx, int64_t index, int n, const uint8_t *delta, const uint64_t *offset,
                       int32_t bound, bool report, bool append, uint64_t budget, int *slot_out) {
    *slot_out = -1;
    int slot = -1;
    for (int i = 0; i < kLargeMaxInstances; ++i)
        if (!slots[i].used) {
            slot = i;
            break;
        }
    if (slot < 0) return SSW_OK;  // engine full
    const uint64_t all = 1ull << n;
    const uint64_t ring_cap = first_ring_cap(n, append);
    const uint64_t need = admit_bytes(n, append);
    if (bytes_in_use + need > budget && bytes_in_use > 0) return SSW_OK;  // wait for memory
    if (need > budget) return ctx.fail(SSW_ERR_BUDGET, "n=%d needs %llu bytes, budget %llu", n,
                                       (unsigned long long)need, (unsigned long long)budget);
    LargeSlot &S = slots[slot];
    S = LargeSlot();
    const uint64_t bm_bytes = std::max<uint64_t>(256, all / 8);
    S.n = n;
    S.ring_cap = ring_cap;
    S.lvl_cap = append ? (uint32_t)((n - 1) * (n - 1) + 8) : 0;
    LG_TRY(take((void **)&S.bitmap, bm_bytes), "bitmap");
    LG_TRY(take((void **)&S.ring, 4 * ring_cap), "ring");
    LG_TRY(take((void **)&S.table, 8192), "table");
    if (append) LG_TRY(take((void **)&S.lvl, 8ull * S.lvl_cap), "levels");
    S.bytes = need;
    bytes_in_use += need;
    S.used = true;
    S.index = index;
    S.append = append;
    S.harvested = false;
    LG_TRY(fill_u32(S.bitmap, bm_bytes / 4, 0, ctx.stream), "clear bitmap");
    LargeInstance I;
    std::memset(&I, 0, sizeof I);
    I.ring = S.ring;
    I.bitmap = S.bitmap;
    I.table = S.table;
    I.lvl = S.lvl;
    I.ring_mask = ring_cap - 1;
    I.index = index;
    I.n = n;
    I.bound = bound;
    I.lvl_cap = S.lvl_cap;
    I.append = append ? 1u : 0u;
    I.report = report ? 1 : 0;
    I.status = kLIdle;
    hinst[slot] = I;
    LG_TRY(hipMemcpyAsync(dinst() + slot, &hinst[slot], sizeof I, hipMemcpyHostToDevice, ctx.stream),
           "H2D instance");
    LG_TRY(hipStreamSynchronize(ctx.stream), "sync");
    LG_TRY(large_launch_init(dinst(), slot, delta, offset, index, ctx.stream), "large_init");
    hinst[slot].status = kLRun;  // what large_init leaves on the device
    if (slot + 1 > num) {
        num = slot + 1;
        LG_TRY(hipMemcpyAsync(&dstate()->num, &num, sizeof num, hipMemcpyHostToDevice, ctx.stream), "H2D num");
    }
    *slot_out = slot;
    return SSW_OK;
}

// The captured chain of `levels` [large_step, large_plan] pairs, built once
// per level count and kept.
int LargeEngine::build_graph(ssw_context &ctx, int levels) {
    if (graphs.count(levels)) return SSW_OK;
    hipGraph_t g = nullptr;
    LG_TRY(hipStreamBeginCapture(ctx.stream, hipStreamCaptureModeThreadLocal), "begin capture");
    for (int k = 0; k < levels; ++k) {
        hipError_t e = large_launch_step(dstate(), dinst(), grid, ctx.stream);
        if (e == hipSuccess) e = large_launch_plan(dstate(), dinst(), 1, ctx.stream);
        if (e != hipSuccess) {
            hipStreamEndCapture(ctx.stream, &g);
            if (g) hipGraphDestroy(g);
            return ctx.hip_fail(e, "capture");
        }
    }
    LG_TRY(hipStreamEndCapture(ctx.stream, &g), "end capture");
    hipGraphExec_t exec = nullptr;
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return ctx.hip_fail(e, "graph instantiate");
    graphs[levels] = exec;
    return SSW_OK;
}

// Runs `levels` levels of every running instance, then grows paused rings.
int LargeEngine::tick(ssw_context &ctx, int levels) {
    int rc = build_graph(ctx, levels);
    if (rc) return rc;
    LG_TRY(large_launch_plan(dstate(), dinst(), 0, ctx.stream), "plan");
    LG_TRY(hipGraphLaunch(graphs[levels], ctx.stream), "graph launch");
    ctx.total_launches += 2 * levels;
    return read_back(ctx);
}

int LargeEngine::read_back(ssw_context &ctx) {
    if (num == 0) return SSW_OK;
    LG_TRY(hipMemcpyAsync(hinst.data(), dinst(), sizeof(LargeInstance) * num, hipMemcpyDeviceToHost, ctx.stream),
           "D2H instances");
    LG_TRY(hipStreamSynchronize(ctx.stream), "sync");
    for (int i = 0; i < num; ++i) {
        LargeSlot &S = slots[i];
        LargeInstance &I = hinst[i];
        if (!S.used || I.status != kLPause) continue;
        // grow the ring (or the level table) so the next level fits; the
        // same bound as large_plan's: at most 2 * width new subsets, and no
        // more than the 2^n distinct ones in the ring
        const uint64_t width = I.hi - I.lo;
        const uint64_t all = 1ull << S.n;
        const uint64_t need = std::min<uint64_t>(I.append ? I.hi + 2 * width : 3 * width, all);
        if (need > S.ring_cap) {
            uint64_t cap = pow2_at_least(std::max<uint64_t>(2 * need, 2 * S.ring_cap));
            if (!I.append) cap = std::min<uint64_t>(cap, pow2_at_least(all));
            else cap = std::min<uint64_t>(cap, pow2_at_least(all + 1));
            if (cap < need)
                return ctx.fail(SSW_ERR_INTERNAL, "device-wide path: ring of %llu entries for a level needing %llu",
                                (unsigned long long)cap, (unsigned long long)need);
            uint32_t *nr = nullptr;
            LG_TRY(take((void **)&nr, 4 * cap), "grow ring");
            const uint64_t from = I.append ? 0 : I.lo;
            LG_TRY(large_launch_regrow(S.ring, S.ring_cap - 1, nr, cap - 1, from, I.hi, ctx.stream), "regrow");
            LG_TRY(hipStreamSynchronize(ctx.stream), "sync");
            give(S.ring, 4 * S.ring_cap);
            bytes_in_use += 4 * (cap - S.ring_cap);
            S.bytes += 4 * (cap - S.ring_cap);
            S.ring = nr;
            S.ring_cap = cap;
            I.ring = nr;
            I.ring_mask = cap - 1;
        }
        if (I.append && (uint32_t)I.level + 2 >= I.lvl_cap) {
            // more levels than (n-1)^2 + 8: a Cerny-bound violation; keep going
            const uint32_t cap = I.lvl_cap * 2;
            uint64_t *nl = nullptr;
            LG_TRY(take((void **)&nl, 8ull * cap), "grow levels");
            LG_TRY(hipMemcpyAsync(nl, S.lvl, 8ull * I.lvl_cap, hipMemcpyDeviceToDevice, ctx.stream), "copy levels");
            LG_TRY(hipStreamSynchronize(ctx.stream), "sync");
            give(S.lvl, 8ull * S.lvl_cap);
            S.lvl = nl;
            S.lvl_cap = cap;
            I.lvl = nl;
            I.lvl_cap = cap;
        }
        I.status = kLRun;
        LG_TRY(hipMemcpyAsync(dinst() + i, &I, sizeof I, hipMemcpyHostToDevice, ctx.stream), "H2D instance");
        LG_TRY(hipStreamSynchronize(ctx.stream), "sync");
    }
    return SSW_OK;
}

int LargeEngine::running() const {
    int r = 0;
    for (int i = 0; i < num; ++i)
        if (slots[i].used && (hinst[i].status == kLRun || hinst[i].status == kLPause)) ++r;
    return r;
}

// Witness of a finished, found instance in append mode.
int LargeEngine::witness(ssw_context &ctx, int slot, uint8_t *out, int stride) {
    LargeSlot &S = slots[slot];
    const LargeInstance &I = hinst[slot];
    const int L = I.length;
    const uint64_t words = std::max<uint64_t>(1, (1ull << S.n) / 32);
    uint32_t *good = nullptr;
    LG_TRY(take((void **)&good, 8 * words), "witness bitmaps");
    LG_TRY(fill_u32(good, 2 * words, 0, ctx.stream), "clear");
    uint32_t *g0 = good, *g1 = good + words;
    for (int j = L - 1; j >= 0; --j) {
        uint32_t *dst = (j & 1) ? g1 : g0;
        const uint32_t *src = (j & 1) ? g0 : g1;
        LG_TRY(large_launch_back(dinst(), slot, j, L, src, dst, grid, ctx.stream), "large_back");
    }
    LG_TRY(large_launch_fwd(dinst(), slot, L, g0, g1, out, stride, ctx.stream), "large_fwd");
    LG_TRY(hipStreamSynchronize(ctx.stream), "sync");
    give(good, 8 * words);
    ctx.total_launches += L + 1;
    return SSW_OK;
}

void LargeEngine::retire(int slot) {
    LargeSlot &S = slots[slot];
    free_slot_buffers(S);
    S.used = false;
    LargeInstance &I = hinst[slot];
    I.status = kLIdle;
}

int LargeEngine::clear_slot(ssw_context &ctx, int slot) {
    std::memset(&hinst[slot], 0, sizeof(LargeInstance));
    hinst[slot].status = kLIdle;
    LG_TRY(hipMemcpyAsync(dinst() + slot, &hinst[slot], sizeof(LargeInstance), hipMemcpyHostToDevice, ctx.stream),
           "H2D instance");
    LG_TRY(hipStreamSynchronize(ctx.stream), "sync");
    return SSW_OK;
}

// Solves automata `idx` of the batch concurrently (as many at a time as
// slots and memory allow).  Results go straight into the batch outputs.
int LargeEngine::solve_indices(ssw_context &ctx, const DeviceBatch &db, const uint8_t *host_n,
                               const std::vector<uint32_t> &idx) {
    int rc = setup(ctx);
    if (rc) return rc;
    uint8_t *c = (uint8_t *)ctx.ctrs.p;
    rc = set_outputs(ctx, db.length, db.sync, db.expanded, (unsigned long long *)(c + kCtrViolation),
                     (unsigned long long *)ctx.prog.p);
    if (rc) return rc;
    const bool append = db.witness != nullptr;
    const uint64_t budget = ctx.budget / 2;
    size_t next = 0;
    std::vector<int> slot_of(kLargeMaxInstances, -1);
    while (next < idx.size() || running() > 0) {
        // admit
        while (next < idx.size()) {
            const uint32_t i = idx[next];
            const int n = db.uniform_n > 0 ? db.uniform_n : host_n[i];
            int slot = -1;
            rc = admit(ctx, i, n, db.delta, db.offset, db.bound, true, append, budget, &slot);
            if (rc) return rc;
            if (slot < 0) break;
            ++next;
        }
        if (running() == 0 && next < idx.size())
            return ctx.fail(SSW_ERR_BUDGET, "cannot admit automaton %u to the device-wide path", idx[next]);
        rc = tick(ctx, tick_levels);
        if (rc) return rc;
        for (int s = 0; s < num; ++s) {
            if (!slots[s].used) continue;
            const LargeInstance &I = hinst[s];
            if (I.status == kLError) {
                // read before the slot is cleared
                const int64_t bad = I.index;
                const int32_t why = I.length;
                retire(s);
                rc = clear_slot(ctx, s);
                if (rc) return rc;
                if (why == -3) {
                    // invalid transitions: report like the batched kernel
                    // (length SSW_LEN_INVALID, flag 0, nothing expanded)
                    static const int32_t v = -3;
                    static const uint8_t no_sync = 0;
                    static const int64_t none = 0;
                    LG_TRY(hipMemcpyAsync(db.length + bad, &v, 4, hipMemcpyHostToDevice, ctx.stream), "H2D");
                    if (db.sync)
                        LG_TRY(hipMemcpyAsync(db.sync + bad, &no_sync, 1, hipMemcpyHostToDevice, ctx.stream), "H2D");
                    if (db.expanded)
                        LG_TRY(hipMemcpyAsync(db.expanded + bad, &none, 8, hipMemcpyHostToDevice, ctx.stream),
                               "H2D");
                    if (append)
                        LG_TRY(hipMemsetAsync(db.witness + (uint64_t)bad * db.witness_stride, 0, db.witness_stride,
                                              ctx.stream),
                               "clear witness");
                    continue;
                }
                return ctx.fail(SSW_ERR_INTERNAL, "device-wide path: frontier overflow on automaton %lld",
                                (long long)bad);
            }
            if (I.status == kLFound || I.status == kLNoSync) {
                if (append && I.status == kLFound) {
                    rc = witness(ctx, s, db.witness + (uint64_t)I.index * db.witness_stride, db.witness_stride);
                    if (rc) return rc;
                } else if (append) {
                    LG_TRY(hipMemsetAsync(db.witness + (uint64_t)I.index * db.witness_stride, 0,
                                          db.witness_stride, ctx.stream),
                           "clear witness");
                }
                retire(s);
                rc = clear_slot(ctx, s);
                if (rc) return rc;
            }
        }
    }
    LG_TRY(hipStreamSynchronize(ctx.stream), "sync");
    return SSW_OK;
}

int LargeEngine::solve_all(ssw_context &ctx, const DeviceBatch &db, const uint8_t *host_n,
                           int64_t *violation_out) {
    hipError_t e;
    if ((e = fill_u32((uint32_t *)((uint8_t *)ctx.ctrs.p + kCtrViolation), 2, 0xFFFFFFFFu, ctx.stream)) !=
        hipSuccess)
        return ctx.hip_fail(e, "fill");
    // automata done in this solve <- 0
    if ((e = reset_counters(nullptr, nullptr, (unsigned long long *)ctx.prog.p, nullptr, 0, ctx.stream)) != hipSuccess)
        return ctx.hip_fail(e, "fill");
    std::vector<uint32_t> idx(db.count);
    for (int64_t i = 0; i < db.count; ++i) idx[i] = (uint32_t)i;
    LG_TRY(hipEventRecord(ctx.ev0, ctx.stream), "event");
    const uint64_t before = ctx.total_launches.load();
    int rc = solve_indices(ctx, db, host_n, idx);
    if (rc) return rc;
    LG_TRY(hipEventRecord(ctx.ev1, ctx.stream), "event");
    LG_TRY(hipEventSynchronize(ctx.ev1), "event sync");
    float ms = 0;
    LG_TRY(hipEventElapsedTime(&ms, ctx.ev0, ctx.ev1), "elapsed");
    ctx.last_ms = ms;
    ctx.last_launches = (int64_t)(ctx.total_launches.load() - before);
    unsigned long long viol = ~0ull;
    LG_TRY(hipMemcpyAsync(&viol, (uint8_t *)ctx.ctrs.p + kCtrViolation, 8, hipMemcpyDeviceToHost, ctx.stream), "D2H");
    LG_TRY(hipStreamSynchronize(ctx.stream), "sync");
    if (violation_out) *violation_out = viol == ~0ull ? -1 : (int64_t)viol;
    return SSW_OK;
}

}  // namespace ssw
