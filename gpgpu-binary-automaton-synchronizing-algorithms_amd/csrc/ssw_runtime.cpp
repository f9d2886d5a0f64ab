// ssw_runtime.cpp -- the C ABI (include/ssw.h): contexts, buffers, dispatch of
// the batched / plain / device-wide solvers, tuning table, progress counters.
//
// The launcher replaces SlimGPUQueue.ComputeManyWithAction
// (GPGPU/GPGPU/GPU/SlimGPUQueue.cs:32-208): instead of allocating and freeing
// device arrays and streams on every call, a context keeps one stream, grows
// its device buffers monotonically and keeps per-group scratch ("slots")
// resident and clean between calls.
#include "ssw.h"
#include "ssw_internal.h"
#include "ssw_host.h"

#include <algorithm>
#include <cmath>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace ssw {
hipError_t launch_batch(int variant, int group_lanes, const BatchParams &P, int grid, int block,
                        size_t lds, hipStream_t st);
hipError_t batch_occupancy(int variant, int group_lanes, int block, size_t lds, int *blocks);
size_t batch_scratch_bytes(int group_lanes);
size_t batch_group_lds(int variant, uint32_t bitmap_words, uint32_t arena_cap);
size_t batch_block_lds(int variant, int group_lanes, int groups, uint32_t bitmap_words, uint32_t arena_cap);
hipError_t fill_u32(uint32_t *p, uint64_t words, uint32_t value, hipStream_t st);
hipError_t count_sizes(const uint8_t *n_per, int64_t count, unsigned long long *hist, hipStream_t st);
#ifdef SSW_CHECKED
hipError_t checked_take_batch(hipStream_t s, unsigned long long *out);
hipError_t checked_take_large(hipStream_t s, unsigned long long *out);
#endif
hipError_t init_queues(uint32_t *queues, int64_t count, hipStream_t st, bool empty);
hipError_t init_queues3(uint32_t *q0, uint32_t *q1, uint32_t *q2, int64_t count, uint32_t empty_mask,
                        hipStream_t st);
hipError_t reset_counters(uint32_t *deferred, unsigned long long *violation, unsigned long long *progress,
                          unsigned long long *launch_stats, uint32_t stat_words, hipStream_t st);
hipError_t gather_listed(const uint32_t *queue, const uint32_t *list, uint32_t items, uint32_t count,
                         const int32_t *length, const uint8_t *sync, const int64_t *expanded, unsigned long long *out,
                         hipStream_t st);
hipError_t noalloc_scratch(int device, void **out);
std::mutex &noalloc_call_lock(int device);
}  // namespace ssw

using namespace ssw;

static_assert(SSW_NOALLOC_SCRATCH_BYTES == kNoAllocScratchBytes && SSW_NOALLOC_MAX_COUNT == kNoAllocMaxCount &&
                  SSW_NOALLOC_MAX_N == kNoAllocMaxN,
              "allocation-free limits in ssw.h");

static thread_local std::string g_create_error;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------

int ssw_context::fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
}

int ssw_context::hip_fail(hipError_t e, const char *what) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return fail(SSW_ERR_NOMEM, "%s: %s", what, hipGetErrorString(e));
    return fail(SSW_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(expr, what)                          \
    do {                                             \
        hipError_t _e = (expr);                      \
        if (_e != hipSuccess) return hip_fail(_e, what); \
    } while (0)

int ssw_context::ensure(DevBuf &b, size_t bytes, const char *what) {
    if (b.bytes >= bytes && b.p) return SSW_OK;
    if (b.p) {
        hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return hip_fail(e, what);
    }
    b.bytes = want;
    return SSW_OK;
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------

int ssw_context::open(int dev, uint64_t mem_budget) {
    device = dev;
    HIP_TRY(hipSetDevice(dev), "hipSetDevice");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    cus = prop.multiProcessorCount;
    lds_per_cu = prop.maxSharedMemoryPerMultiProcessor;
    lds_per_block = prop.sharedMemPerBlock;
    HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    HIP_TRY(hipStreamCreateWithFlags(&monitor, hipStreamNonBlocking), "hipStreamCreate");
    HIP_TRY(hipEventCreate(&ev0), "hipEventCreate");
    HIP_TRY(hipEventCreate(&ev1), "hipEventCreate");
    HIP_TRY(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreate");
    for (int k = 0; k < kClassStreams; ++k) {
        HIP_TRY(hipStreamCreateWithFlags(&class_streams[k], hipStreamNonBlocking), "hipStreamCreate");
        HIP_TRY(hipEventCreateWithFlags(&ev_join[k], hipEventDisableTiming), "hipEventCreate");
    }
    for (hipEvent_t &e : pipe_ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    size_t free_b = 0, total_b = 0;
    HIP_TRY(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    budget = mem_budget ? mem_budget : (uint64_t)(0.8 * (double)free_b);
    int rc = ensure(ctrs, kCtrBytes, "counters");
    if (rc) return rc;
    rc = ensure(queues, kQueueBytes, "work queues");
    if (rc) return rc;
    rc = ensure(lstats, 8u * kLaunchStatWords, "launch statistics");
    if (rc) return rc;
    rc = ensure(prog, kProgressBytes, "progress counters");
    if (rc) return rc;
    for (int k = 0; k < 2 * kMaxLaunchRecords; ++k) HIP_TRY(hipEventCreate(&lev[k]), "hipEventCreate");
    HIP_TRY(hipMemsetAsync(lstats.p, 0, 8u * kLaunchStatWords, stream), "hipMemset");
    HIP_TRY(hipMemsetAsync(prog.p, 0, kProgressBytes, stream), "hipMemset");
    HIP_TRY(hipMemsetAsync(ctrs.p, 0, kCtrBytes, stream), "hipMemset");
    HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    // the allocation-free entry point's scratch on this device (its own
    // address per GPU); a failure here is reported by that entry point only
    noalloc_status = noalloc_scratch(dev, &noalloc_base);
    default_tuning();
    return SSW_OK;
}

void ssw_context::close() {
    stop_progress_thread();
    if (device >= 0) hipSetDevice(device);
    for (DevBuf *b : {&ctrs, &queues, &deferred, &order, &spill, &lstats, &prog, &sizes, &d_delta, &d_nper, &d_off, &d_len,
                      &d_sync, &d_exp, &d_wit, &d_delta2, &d_len2, &d_sync2, &d_exp2, &pipe_q, &pipe_list})
        if (b->p) hipFree(b->p);
    for (int k = 0; k < 2 * kMaxLaunchRecords; ++k)
        if (lev[k]) hipEventDestroy(lev[k]);
    for (auto &kv : slot_cache)
        if (kv.second.p) hipFree(kv.second.p);
    slot_cache.clear();
    large.release();
    if (ev0) hipEventDestroy(ev0);
    if (ev1) hipEventDestroy(ev1);
    if (ev_fork) hipEventDestroy(ev_fork);
    for (int k = 0; k < kClassStreams; ++k) {
        if (class_streams[k]) hipStreamDestroy(class_streams[k]);
        if (ev_join[k]) hipEventDestroy(ev_join[k]);
    }
    if (stream) hipStreamDestroy(stream);
    if (monitor) hipStreamDestroy(monitor);
    for (hipEvent_t &e : pipe_ev)
        if (e) hipEventDestroy(e);
    if (h_stage) hipHostFree(h_stage);
    h_stage = nullptr;
    if (h_out) hipHostFree(h_out);
    h_out = nullptr;
}

void ssw_context::default_tuning() {
    for (int n = 0; n <= SSW_MAX_N; ++n) {
        ssw_tuning t;
        if (n <= 10) {
            t.group_lanes = 16; t.block_threads = 256; t.chunk = 16; t.arena_lds = 1024; t.lane_entries = 1;
        } else if (n <= 13) {
            t.group_lanes = 32; t.block_threads = 256; t.chunk = 8; t.arena_lds = 2048; t.lane_entries = 1;
        } else if (n <= 16) {
            t.group_lanes = 64; t.block_threads = 0; t.chunk = 2; t.arena_lds = 2048; t.lane_entries = 1;
        } else {
            t.group_lanes = 64; t.block_threads = 256; t.chunk = 2; t.arena_lds = 2048; t.lane_entries = 1;
        }
        tune[n] = t;
    }
}

// ---------------------------------------------------------------------------
// batched solve on device buffers
// ---------------------------------------------------------------------------

// Per-group scratch in HBM for one launch configuration.
int ssw_context::get_slots(const SlotKey &key, uint64_t bytes, DevBuf **out, hipStream_t st) {
    auto it = slot_cache.find(key);
    if (it != slot_cache.end() && it->second.bytes >= bytes) {
        slot_gen[key] = call_gen;
        *out = &it->second;
        return SSW_OK;
    }
    if (it != slot_cache.end()) {
        hipFree(it->second.p);
        slot_cache.erase(it);
    }
    uint64_t used = 0;
    for (auto &kv : slot_cache) used += kv.second.bytes;
    // evict configurations not used by this call rather than exceed the budget
    for (auto v = slot_cache.begin(); v != slot_cache.end() && used + bytes > budget;) {
        if (slot_gen[v->first] == call_gen) {
            ++v;
            continue;
        }
        used -= v->second.bytes;
        hipFree(v->second.p);
        slot_gen.erase(v->first);
        v = slot_cache.erase(v);
    }
    DevBuf b;
    int rc = ensure(b, bytes, "slot scratch");
    if (rc) return rc;
    // bitmaps must start clean; the kernels keep them clean afterwards
    HIP_TRY(hipMemsetAsync(b.p, 0, b.bytes, st), "hipMemset slots");
    slot_cache[key] = b;
    slot_gen[key] = call_gen;
    *out = &slot_cache[key];
    return SSW_OK;
}

static inline uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// host calls cut into chunks: few large ones, since every solve pays its
// launches' tails (the longest searches of the chunk), so a chunk below
// ~2^18 automata costs more in tails than the host work and copies it hides
static constexpr int64_t kMinHostChunk = 1 << 18;
// result records of the host pipeline's last launch copied back with their count
constexpr uint64_t kPatchFirst = 16384;


static inline uint32_t next_pow2(uint64_t x) {
    uint64_t p = 1;
    while (p < x) p <<= 1;
    return (uint32_t)p;
}

// One launch class.  With ring spill on (n <= 16 path, spill_arena_lds > 0)
// it is a chain of launches on the class's stream, each over the automata the
// previous one listed:
//   1. the tuned configuration; automata whose frontier outgrows the small LDS
//      ring are listed instead of moved to HBM.  With narrow_bits, states are
//      relabelled so that a(Q) | b(Q) takes the lowest labels and the visited
//      bitmap has 2^narrow_bits bits; automata with |a(Q) | b(Q)| above that
//      are listed at start,
//   2. (narrow_bits only) the same ring with the full 2^n-bit bitmap,
//   3. the larger ring (spill_arena_lds), byte tables.
int ssw_context::launch_class(const DeviceBatch &db, const ClassSpec &cls, bool plain) {
    if (cls.static_slots) return launch_class_static(db, cls);
    ssw_tuning t = plain ? ssw_tuning{256, 256, 1, 8192, 1, 0, 0, 0, 0} : tune[cls.nmax];
    const bool small = cls.lds && cls.nmax <= kMaxLdsN && (db.witness == nullptr || !plain);
    if (small && db.witness != nullptr) {
        // witness words on the n <= 16 path: one wavefront per automaton,
        // one entry per lane in the first launches and four in the last
        t.group_lanes = 64;
        t.block_threads = 0;
        t.lane_entries = 1;
        t.spill_group_lanes = 0;
        t.spill_lane_entries = 4;
    }
    if (!plain && !small && db.witness == nullptr && cls.nmax <= SSW_MAX_N && t.narrow_bits == 16 &&
        cls.spill_queue[0]) {
        // n > 16: automata with |a(Q)|, |b(Q)| <= 15 run in the n <= 16
        // solver with the two-family encoding (LDS bitmap, 16-bit ring; the
        // first launch configuration of n = 16); the others, and frontiers
        // that outgrow its ring, go on to this class's solver from a spill
        // queue
        hipStream_t st = cls.stream ? cls.stream : stream;
        HIP_TRY(init_queues(cls.spill_queue[0], cls.items, st, true), "init queues");
        ssw_tuning tp = tune[kMaxLdsN];
        tp.group_lanes = 64;
        tp.lane_entries = 1;
        // a ring of 2048 entries: the frontiers of n > 16 are wider than
        // those of n = 16 and a search that outgrows the ring starts over in
        // the slower solver
        tp.arena_lds = std::max(tp.arena_lds, 2048);
        tp.block_threads = 0;
        PassIO pre;
        pre.kind = kPassSpillFrom;
        pre.out_list = cls.spill_list[0];
        pre.out_queue = cls.spill_queue[0];
        pre.narrow_bits = 16;
        pre.force_small = true;
        pre.fam = true;
        int rc = launch_pass(db, cls, tp, false, pre);
        if (rc) return rc;
        PassIO rest;
        rest.kind = kPassSpillTo;
        rest.in_list = cls.spill_list[0];
        rest.in_queue = cls.spill_queue[0];
        // what is left are the long searches (a letter close to a
        // permutation): four frontier entries per lane and step, and with
        // |a(Q)|, |b(Q)| <= 18 the two-family encoding with its bitmap in LDS
        rest.general_entries = 4;
        rest.general_fam = true;
        return launch_pass(db, cls, t, false, rest);
    }
    if (plain || !small || t.spill_arena_lds <= 0 || !cls.spill_queue[0])
        return launch_pass(db, cls, t, plain, PassIO{});
    const int nb = (t.narrow_bits >= 8 && t.narrow_bits < cls.nmax) ? t.narrow_bits : 0;
    hipStream_t st = cls.stream ? cls.stream : stream;
    if (cls.defer_last) {
        // the first two launches only; what the last one takes goes to the
        // parts' shared list
        if (!nb) return fail(SSW_ERR_INTERNAL, "deferred last launch without the narrow chain (n=%d)", cls.nmax);
        HIP_TRY(init_queues3(cls.queues, cls.spill_queue[0], nullptr, cls.items, 2u, st), "init queues");
        PassIO first;
        first.queue_ready = true;
        first.kind = kPassSpillFrom;
        first.out_list = cls.spill_list[0];
        first.out_queue = cls.spill_queue[0];
        first.narrow_bits = nb;
        first.skip_list = cls.spill_list[1];
        first.skip_queue = cls.spill_queue[1];
        first.skip_base = cls.list1_base;
        int rc = launch_pass(db, cls, t, false, first);
        if (rc) return rc;
        PassIO mid;
        mid.kind = kPassSpillFrom;
        mid.in_list = cls.spill_list[0];
        mid.in_queue = cls.spill_queue[0];
        mid.out_list = cls.spill_list[1];
        mid.out_queue = cls.spill_queue[1];
        mid.out_base = cls.list1_base;
        ssw_tuning tm = t;
        tm.chunk = std::min(t.chunk, 4);
        return launch_pass(db, cls, tm, false, mid);
    }
    // one launch sets up the class's queue (full) and the later launches'
    // (the first launch's segments, empty)
    HIP_TRY(init_queues3(cls.queues, cls.spill_queue[0], nb ? cls.spill_queue[1] : nullptr, cls.items, 6u, st),
            "init queues");
    PassIO first;
    first.queue_ready = true;
    first.kind = kPassSpillFrom;
    first.out_list = cls.spill_list[0];
    first.out_queue = cls.spill_queue[0];
    first.narrow_bits = nb;
    if (nb) {
        // automata with a permutation letter go straight to the large ring
        first.skip_list = cls.spill_list[1];
        first.skip_queue = cls.spill_queue[1];
    }
    int rc = launch_pass(db, cls, t, false, first);
    if (rc) return rc;
    PassIO last;
    last.kind = kPassSpillTo;
    last.in_list = cls.spill_list[0];
    last.in_queue = cls.spill_queue[0];
    last.byte_tab = true;
    if (nb) {
        PassIO mid;
        mid.kind = kPassSpillFrom;
        mid.in_list = cls.spill_list[0];
        mid.in_queue = cls.spill_queue[0];
        mid.out_list = cls.spill_list[1];
        mid.out_queue = cls.spill_queue[1];
        // few automata per claim: the listed automata are fewer and their
        // searches longer and more uneven (C_n among them) than the first
        // launch's, so 16-automaton claims leave most groups idle
        ssw_tuning tm = t;
        tm.chunk = std::min(t.chunk, 4);
        rc = launch_pass(db, cls, tm, false, mid);
        if (rc) return rc;
        last.in_list = cls.spill_list[1];
        last.in_queue = cls.spill_queue[1];
    }
    ssw_tuning t2 = t;
    if (t.spill_group_lanes > 0) t2.group_lanes = t.spill_group_lanes;
    t2.block_threads = t2.group_lanes > 64 ? t2.group_lanes : 0;
    t2.chunk = 1;
    t2.arena_lds = t.spill_arena_lds;
    t2.lane_entries = t.spill_lane_entries > 0 ? t.spill_lane_entries : t.lane_entries;
    return launch_pass(db, cls, t2, false, last);
}

// The allocation-free chain of a uniform class of n <= 16: the tuned first
// launch (and with narrow bitmaps the full-bitmap one), then the larger ring
// with byte tables, all of them handing a frontier that outgrows their LDS
// ring to the next launch; the last launch (kNoAllocLastRing entries, one
// wavefront per automaton) is the only one that may move a frontier to an
// HBM ring, and its grid is capped so those rings fit the static scratch.
int ssw_context::launch_class_static(const DeviceBatch &db, const ClassSpec &cls) {
    ssw_tuning t = tune[cls.nmax];
    if (t.spill_arena_lds <= 0) t.spill_arena_lds = 2048;
    if (t.spill_lane_entries <= 0) t.spill_lane_entries = 4;
    const int nb = (t.narrow_bits >= 8 && t.narrow_bits < cls.nmax) ? t.narrow_bits : 0;
    hipStream_t st = cls.stream ? cls.stream : stream;
    HIP_TRY(init_queues3(cls.queues, cls.spill_queue[0], nb ? cls.spill_queue[1] : nullptr, cls.items, 6u, st),
            "init queues");
    HIP_TRY(init_queues(cls.spill_queue[2], cls.items, st, true), "init queues");
    PassIO first;
    first.queue_ready = true;
    first.kind = kPassSpillFrom;
    first.out_list = cls.spill_list[0];
    first.out_queue = cls.spill_queue[0];
    first.narrow_bits = nb;
    if (nb) {
        first.skip_list = cls.spill_list[1];
        first.skip_queue = cls.spill_queue[1];
    }
    int rc = launch_pass(db, cls, t, false, first);
    if (rc) return rc;
    PassIO wide;  // the larger ring, byte tables, spilling to the last launch
    wide.kind = kPassSpillFrom;
    wide.in_list = cls.spill_list[0];
    wide.in_queue = cls.spill_queue[0];
    wide.out_list = cls.spill_list[2];
    wide.out_queue = cls.spill_queue[2];
    wide.byte_tab = true;
    if (nb) {
        PassIO mid;
        mid.kind = kPassSpillFrom;
        mid.in_list = cls.spill_list[0];
        mid.in_queue = cls.spill_queue[0];
        mid.out_list = cls.spill_list[1];
        mid.out_queue = cls.spill_queue[1];
        ssw_tuning tm = t;
        tm.chunk = std::min(t.chunk, 4);
        rc = launch_pass(db, cls, tm, false, mid);
        if (rc) return rc;
        wide.in_list = cls.spill_list[1];
        wide.in_queue = cls.spill_queue[1];
    }
    ssw_tuning t2 = t;
    t2.group_lanes = 64;
    t2.block_threads = 0;
    t2.chunk = 1;
    t2.arena_lds = t.spill_arena_lds;
    t2.lane_entries = t.spill_lane_entries;
    rc = launch_pass(db, cls, t2, false, wide);
    if (rc) return rc;
    PassIO last;
    last.kind = kPassSpillTo;
    last.in_list = cls.spill_list[2];
    last.in_queue = cls.spill_queue[2];
    last.byte_tab = true;
    ssw_tuning t3 = t2;
    t3.arena_lds = (int32_t)kNoAllocLastRing;
    t3.lane_entries = 4;
    return launch_pass(db, cls, t3, false, last);
}

int ssw_context::launch_pass(const DeviceBatch &db, const ClassSpec &cls, const ssw_tuning &t, bool plain,
                             const PassIO &io) {
    const int G = t.group_lanes;
    int block = G > 64 ? G : std::max(64, std::min(256, t.block_threads ? t.block_threads : 256));
    const bool use_lds = cls.lds || io.force_small;
    const int nmax = cls.nmax;
    const bool want_wit = db.witness != nullptr;
    // the small path: bitmap in LDS, 16-bit ring frontier (with witness
    // words: one wavefront per automaton, not the plain path)
    const bool small = io.force_small || (use_lds && nmax <= kMaxLdsN && (!want_wit || (!plain && G == 64)));
    // the last launch of a spill chain takes the wide searches: byte-sliced
    // image tables (two lookups per subset instead of four), and with one
    // wavefront per automaton up to 8 entries per lane
    const bool byte_tab = io.byte_tab;
    const bool wit_small = small && want_wit;
    const int T = wit_small ? (byte_tab ? 4 : 1)
                  : (t.lane_entries >= 8 && byte_tab && G == 64) ? 8
                  : (t.lane_entries >= 4 ? 4 : (t.lane_entries >= 2 ? 2 : 1));
    const bool general4 = !small && io.general_entries >= 4 && !plain && (G == 64 || G == 256);
    const int variant = !small ? (general4 ? (io.general_fam && !want_wit ? 12 : 11) : 0)
                        : io.fam ? 10
                        : wit_small ? (byte_tab ? 9 : 8)
                        : (T == 8 ? 7 : (T == 1 ? 1 : (T == 2 ? 2 : 3)) + (byte_tab ? 3 : 0));

    BatchParams P;
    std::memset(&P, 0, sizeof P);
    P.count = plain ? db.count : (int64_t)cls.items;  // an upper bound in the spill-to pass
    P.uniform_n = db.uniform_n;
    P.bound = db.bound;
    P.n_per = db.n_per;
    P.offset = db.offset;
    P.delta = db.delta;
    P.order = io.in_list ? io.in_list : cls.order;
    P.spill = io.out_list ? 1u : 0u;
    P.spill_queue = io.out_queue;
    P.spill_list = io.out_list;
    P.skip_queue = io.skip_queue;
    P.skip_list = io.skip_list;
    P.skip_base = io.skip_base;
    P.spill_base = io.out_base;
    P.narrow_bits = small ? io.narrow_bits : 0;
    // only Verify reads nothing but the violation, so only it may skip searches
    P.prefilter = (host_no_length && verify_prefilter) ? 1u : 0u;
    P.length = db.length;
    P.sync = db.sync;
    P.expanded = db.expanded;
    P.witness = db.witness;
    P.witness_stride = db.witness_stride;
    P.want_witness = want_wit ? 1u : 0u;
    uint8_t *c = (uint8_t *)ctrs.p;
    P.queues = io.in_queue ? io.in_queue : cls.queues;
    hipStream_t st = cls.stream ? cls.stream : stream;
    P.deferred_count = (unsigned int *)(c + kCtrDeferred);
    P.violation = (unsigned long long *)(c + kCtrViolation);
    P.progress = (unsigned long long *)prog.p;
    P.deferred_list = (uint32_t *)deferred.p;
    P.chunk = (uint32_t)std::max(1, t.chunk);
    P.lds_n = use_lds ? nmax : 0;
    const uint32_t bm_words = small ? SmallLds::bitmap_bytes(P.narrow_bits > 0 ? P.narrow_bits : nmax) / 4u
                              : variant == 12 ? kFamBitmapWords
                              : (use_lds ? (uint32_t)align_up(nmax >= 5 ? (1u << (nmax - 5)) : 1u, 4) : 4u);
    P.lds_bitmap_words = bm_words;
    const uint64_t all_subsets = 1ull << nmax;
    uint32_t arena_lds;
    if (small) {
        // a ring holding every subset of Q never has to leave LDS; a spilling
        // pass only needs room for one narrow step; groups of several
        // wavefronts check whole levels against the ring instead of steps
        const int min_ring = G > 64 ? 256 : (io.out_list ? 4 * G : 4 * G * T);
        arena_lds = next_pow2((uint64_t)std::max(t.arena_lds, min_ring));
        arena_lds = std::min(arena_lds, next_pow2(all_subsets + 2ull * G * T));
    } else {
        arena_lds = (uint32_t)align_up((uint64_t)std::max(4, t.arena_lds), 4);
        if (arena_lds > all_subsets + 2u * G) arena_lds = (uint32_t)align_up(all_subsets + 2u * G, 4);
        // with the 64 KiB bitmap: two 256-lane groups per CU (the arena
        // continues in HBM; its first words hold the encoding's set-up)
        if (variant == 12) arena_lds = std::max(256u, std::min(arena_lds, 2048u));
    }
    P.lds_arena_cap = arena_lds;
    const size_t per_group = batch_group_lds(variant, bm_words, arena_lds);
    P.lds_group_bytes = (uint32_t)per_group;
    // workgroup size: the tuned one, or (block_threads = 0) the one that keeps
    // the most groups resident per CU under the LDS budget, the smallest of
    // those (a small batch then spreads over more CUs)
    int per_cu = 0;
    if (!plain && G <= 64 && t.block_threads == 0) {
        int best_groups = 0;
        for (int cand = 64; cand <= 256; cand += 64) {
            if (cand % G) continue;
            const size_t l = batch_block_lds(variant, G, cand / G, bm_words, arena_lds);
            if (l > lds_per_block) continue;
            int occ = 0;
            HIP_TRY(occupancy(variant, G, cand, l, &occ), "occupancy");
            if (occ * (cand / G) > best_groups && occ > 0) {
                best_groups = occ * (cand / G);
                block = cand;
                per_cu = occ;
            }
        }
    } else {
        // fewer groups per workgroup when the tuned block does not fit the LDS limit
        while (G <= 64 && block > 64 && batch_block_lds(variant, G, block / G, bm_words, arena_lds) > lds_per_block)
            block -= 64;
    }
    const int groups_per_block = block / G;
    const size_t lds = batch_block_lds(variant, G, groups_per_block, bm_words, arena_lds);
    if (lds > lds_per_block)
        return fail(SSW_ERR_ARG, "tuning for n=%d needs %zu bytes of LDS per block (limit %zu)", nmax, lds,
                    lds_per_block);

    uint64_t off = 0;
    if (small) {
        // 16-bit HBM ring for frontiers that outgrow the LDS ring
        P.garena_off = 0;
        P.garena_cap = 65536;
        off = align_up(2ull * 65536, 256);
        if (wit_small) {
            // discovery order (u16, and 64 dummy entries the search's lanes
            // without a new subset write), a good flag per position (u8) and
            // level starts (u32) of up to 65536 subsets
            P.hist_off = off;
            off += 2ull * (65536 + 64);
            P.good_off = off;
            off += 65536;
            P.lvl_off = off;
            off = align_up(off + 4ull * (65536 + 4), 256);
            // the choice bitmap of the witness passes, as wide as the LDS
            // bitmap (zeroed with the slot, left clean by every word)
            P.gbitmap_off = off;
            off = align_up(off + 4ull * bm_words, 256);
            // the long searches' launch also keeps each expanded subset's
            // images (u32 per history position), which its backward pass
            // reads instead of looking them up again
            P.himg_off = off;
            if (T >= 4) off = align_up(off + 4ull * (65536 + 64), 256);
        } else {
            P.hist_off = P.gbitmap_off = P.good_off = P.lvl_off = off;
        }
    } else {
        // every subset fits the HBM arena when n <= 20 (with the step's
        // margin of two images per entry taken); larger automata whose
        // frontier outgrows it are deferred to the device-wide path
        uint64_t garena_cap;
        const uint64_t margin = 2ull * G * (variant >= 11 ? 4u : 1u);
        if (use_lds || nmax <= 20)
            garena_cap = all_subsets + margin > arena_lds ? all_subsets + margin - arena_lds : 0;
        else
            garena_cap = plain ? (1ull << 22) : (1ull << 18);
        const uint64_t gbm_words = use_lds ? 0 : std::max<uint64_t>(1, all_subsets / 32);
        const uint64_t good_words = want_wit ? std::max<uint64_t>(1, all_subsets / 32) : 0;
        const uint32_t lvl_cap = want_wit ? (uint32_t)((nmax - 1) * (nmax - 1) + 8) : 0;
        P.garena_off = off;
        P.garena_cap = garena_cap;
        off = align_up(off + 4 * garena_cap, 256);
        P.gbitmap_off = off;
        P.gbitmap_words = gbm_words;
        off = align_up(off + 4 * gbm_words, 256);
        P.good_off = off;
        P.good_words = good_words;
        off = align_up(off + 8 * good_words, 256);
        P.lvl_off = off;
        P.lvl_cap = lvl_cap;
        off = align_up(off + 4ull * lvl_cap, 256);
    }
    P.slot_bytes = std::max<uint64_t>(off, 256);

    int grid = 1;
    if (!plain) {
        if (per_cu == 0) HIP_TRY(occupancy(variant, G, block, lds, &per_cu), "occupancy");
        if (per_cu < 1) return fail(SSW_ERR_ARG, "kernel does not fit a CU (lds %zu)", lds);
        grid = per_cu * cus;
        const int64_t need = (P.count + groups_per_block - 1) / groups_per_block;
        if (grid > need) grid = (int)std::max<int64_t>(1, need);
        // stay inside the memory budget
        const uint64_t max_slots = std::max<uint64_t>(1, budget / 2 / P.slot_bytes);
        while ((uint64_t)grid * groups_per_block > max_slots && grid > 1) grid = grid / 2;
    }
    if (cls.static_slots) {
        // allocation-free chain: a spilling launch never leaves its LDS ring
        // (no slot memory: every group aliases one unused slot); the last
        // launch's groups get the HBM rings of the static scratch
        if (!small || want_wit) return fail(SSW_ERR_INTERNAL, "allocation-free chain on the general kernel");
        if (io.out_list) {
            P.slot_bytes = 0;
        } else {
            const uint64_t max_groups = cls.static_slot_bytes / P.slot_bytes;
            if (max_groups < (uint64_t)groups_per_block)
                return fail(SSW_ERR_INTERNAL, "static scratch holds %llu rings of %llu bytes",
                            (unsigned long long)max_groups, (unsigned long long)P.slot_bytes);
            grid = (int)std::min<uint64_t>((uint64_t)grid, max_groups / (uint64_t)groups_per_block);
        }
        P.slot_base = cls.static_slots;
    } else {
        const uint64_t slots = (uint64_t)grid * groups_per_block;
        SlotKey key{G, block, P.lds_n, nmax, (int)want_wit, (int)arena_lds,
                    (plain ? 1 : 0) + 2 * variant + 32 * io.kind + 128 * io.narrow_bits};
        DevBuf *sb = nullptr;
        int rc = get_slots(key, slots * P.slot_bytes, &sb, st);
        if (rc) return rc;
        P.slot_base = (uint8_t *)sb->p;
    }
    last_grid = grid;
    last_block = block;
    last_lds = lds;
    last_variant = variant;

    if (!plain) {
        if (!io.in_queue && !io.queue_ready) HIP_TRY(init_queues(P.queues, P.count, st, false), "init queues");
        // the launch's statistics: its counters and the events around it on its stream
        const int rec = records.size() < (size_t)kMaxLaunchRecords ? (int)records.size() : -1;
        if (rec >= 0) {
            LaunchRecord r;
            r.n = nmax;
            r.variant = variant;
            r.G = G;
            r.block = block;
            r.grid = grid;
            r.lds = (int)lds;
            r.ring = (int)arena_lds;
            r.T = small ? T : (variant >= 11 ? 4 : 1);
            r.bitmap_bits = (use_lds || variant == 12) ? (int)(32u * bm_words) : 0;
            r.pass = io.kind == kPassAll && !io.out_list ? 0 : (io.in_list ? 2 : 1);
            r.byte_tab = byte_tab;
            r.witness = want_wit;
            records.push_back(r);
            P.launch_stats = (unsigned long long *)lstats.p + (size_t)rec * kCounterSlots * kStatWords;
            HIP_TRY(hipEventRecord(lev[2 * rec], st), "hipEventRecord");
        }
        HIP_TRY(launch_batch(variant, G, P, grid, block, lds, st), "batch kernel launch");
        if (rec >= 0) HIP_TRY(hipEventRecord(lev[2 * rec + 1], st), "hipEventRecord");
        ++launches;
        ++total_launches;
    } else {
        for (int64_t i = 0; i < db.count; ++i) {
            if (cls.host_n && cls.host_n[i] != cls.plain_n) continue;
            P.single_item = 1;
            P.first_item = i;
            HIP_TRY(launch_batch(variant, G, P, 1, block, lds, st), "plain kernel launch");
            ++launches;
            ++total_launches;
        }
    }
    return SSW_OK;
}

// Splits the batch into launch classes: (thread-group width, bitmap in LDS).
int ssw_context::solve_device(const DeviceBatch &db, const uint8_t *host_n, bool plain,
                              int64_t *violation_out, const int64_t *size_hist, void *head_out,
                              const ClassSpec *defer) {
    // the sizes one by one, copied from the device when a route needs them
    std::vector<uint8_t> fetched_n;
    auto need_host_n = [&]() -> int {
        if (host_n || db.uniform_n > 0) return SSW_OK;
        fetched_n.resize((size_t)db.count);
        HIP_TRY(hipMemcpyAsync(fetched_n.data(), db.n_per, (size_t)db.count, hipMemcpyDeviceToHost, stream),
                "D2H n_per");
        HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
        host_n = fetched_n.data();
        return SSW_OK;
    };
    launches = 0;
    ++call_gen;
    int rc = ensure(deferred, 4ull * std::max<int64_t>(1, db.count), "deferred list");
    if (rc) return rc;
    // violation <- UINT64_MAX, deferred count <- 0, done in this solve <- 0,
    // launch statistics <- 0
    records.clear();
    HIP_TRY(reset_counters((uint32_t *)((uint8_t *)ctrs.p + kCtrDeferred),
                           (unsigned long long *)((uint8_t *)ctrs.p + kCtrViolation),
                           (unsigned long long *)prog.p, (unsigned long long *)lstats.p, kLaunchStatWords, stream),
            "reset counters");
    HIP_TRY(hipEventRecord(ev0, stream), "hipEventRecord");

    int nmin = db.uniform_n, nmaxall = db.uniform_n;
    std::vector<int64_t> hist(SSW_MAX_N + 1, 0);
    if (db.uniform_n > 0) {
        hist[db.uniform_n] = db.count;
    } else if (!host_n && size_hist) {
        nmin = SSW_MAX_N;
        nmaxall = 1;
        for (int n = 1; n <= SSW_MAX_N; ++n) {
            hist[n] = size_hist[n];
            if (hist[n]) {
                nmin = std::min(nmin, n);
                nmaxall = std::max(nmaxall, n);
            }
        }
    } else {
        if ((rc = need_host_n())) return rc;
        nmin = SSW_MAX_N;
        nmaxall = 1;
        for (int64_t i = 0; i < db.count; ++i) {
            const int n = host_n[i];
            if (n < 1 || n > SSW_MAX_N) return fail(SSW_ERR_ARG, "automaton %lld: n=%d out of [1,30]", (long long)i, n);
            hist[n]++;
            nmin = std::min(nmin, n);
            nmaxall = std::max(nmaxall, n);
        }
    }

    std::vector<uint32_t> host_order;
    if (plain) {
        if ((rc = need_host_n())) return rc;
        // one launch per automaton, with the LDS bitmap whenever n <= 16
        for (int n = 1; n <= SSW_MAX_N; ++n) {
            if (!hist[n]) continue;
            ClassSpec cls;
            cls.nmax = n;
            cls.lds = n <= kMaxLdsN;
            cls.items = hist[n];
            cls.order = nullptr;
            cls.host_n = db.uniform_n > 0 ? nullptr : host_n;
            cls.plain_n = n;
            rc = launch_class(db, cls, true);
            if (rc) return rc;
        }
    } else {
        // Launch classes: one per n for the LDS-bitmap sizes (each launch
        // sizes its LDS for its own n, so small automata keep many groups
        // resident), one per thread-group width for the HBM-bitmap sizes.
        // All classes run concurrently on their own streams, each with its
        // own device work queue; inside a class the work order puts the
        // largest n first so the long searches start early.
        // A batch of mixed sizes above 16 without witness words is one class
        // instead: the two-family launch takes every size at once (its
        // tables are built per automaton) and the general solver what it
        // hands on, so the sizes do not queue behind each other in launches
        // that each fill the GPU.
        std::map<std::tuple<int, int, int>, std::vector<int>> classes;  // (G or 0, lds, n or 0) -> sizes
        const bool merged = db.uniform_n <= 0 && db.witness == nullptr && nmaxall > kMaxLdsN &&
                            tune[nmaxall].narrow_bits == 16;
        for (int n = 1; n <= SSW_MAX_N; ++n) {
            if (!hist[n]) continue;
            const bool lds = n <= kMaxLdsN;
            if (merged)
                classes[std::make_tuple(0, 0, 0)].push_back(n);
            else
                classes[std::make_tuple(lds ? 0 : tune[n].group_lanes, lds ? 1 : 0, lds ? n : 0)].push_back(n);
        }
        // one class in index order: a uniform batch, or the merged sizes
        // (the two-family launch's per-automaton cost varies little with n,
        // and its long searches are handed on, so no size order is needed)
        const bool single_class = classes.size() == 1 && (db.uniform_n > 0 || merged);
        std::vector<int64_t> class_start;
        if (!single_class) {
            if ((rc = need_host_n())) return rc;
            rc = ensure(order, 4ull * db.count, "order");
            if (rc) return rc;
            // counting sort of the automata into class order, n descending
            std::vector<int64_t> next(SSW_MAX_N + 1, 0);
            int64_t pos = 0;
            for (auto &kv : classes) {
                class_start.push_back(pos);
                std::vector<int> ns = kv.second;
                std::sort(ns.rbegin(), ns.rend());
                for (int n : ns) {
                    next[n] = pos;
                    pos += hist[n];
                }
            }
            host_order.resize((size_t)db.count);
            for (int64_t i = 0; i < db.count; ++i)
                host_order[(size_t)next[db.uniform_n > 0 ? db.uniform_n : host_n[i]]++] = (uint32_t)i;
            HIP_TRY(hipMemcpyAsync(order.p, host_order.data(), 4ull * db.count, hipMemcpyHostToDevice, stream),
                    "upload order");
        } else {
            class_start.push_back(0);
        }
        rc = ensure(queues, kQueueBytes * kClassQueues * classes.size(), "work queues");
        if (rc) return rc;
        // spill lists: two per class, one slot per automaton at its position
        // in class order
        rc = ensure(spill, 8ull * db.count, "spill lists");
        if (rc) return rc;
        // Streams: the classes are enqueued longest first, each on the stream
        // with the least estimated work so far (at most kConcurrentClasses
        // streams: the process has few hardware queues, and a long class
        // queued behind short ones on a shared queue would start late).
        // Estimated work: automata x 1.45^(n - 12), doubled for HBM bitmaps.
        std::vector<std::pair<double, int>> by_cost;
        {
            int k = 0;
            for (auto &kv : classes) {
                double c = 0;
                for (int n : kv.second) c += (double)hist[n] * std::pow(1.45, n - 12);
                if (std::get<1>(kv.first) == 0) c *= 2.0;
                by_cost.push_back({c, k++});
            }
            std::sort(by_cost.begin(), by_cost.end(), [](const std::pair<double, int> &x, const std::pair<double, int> &y) {
                return x.first > y.first;
            });
        }
        const int nstreams = std::min<int>((int)classes.size(), kConcurrentClasses);
        std::vector<double> load((size_t)nstreams, 0.0);
        HIP_TRY(hipEventRecord(ev_fork, stream), "hipEventRecord");
        for (int k = 0; k < nstreams; ++k) HIP_TRY(hipStreamWaitEvent(class_streams[k], ev_fork, 0), "wait fork");
        std::vector<std::pair<std::tuple<int, int, int>, std::vector<int>>> cl(classes.begin(), classes.end());
        for (const auto &bc : by_cost) {
            const int k = bc.second;
            const auto &kv = cl[(size_t)k];
            const int s = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            load[(size_t)s] += bc.first;
            ClassSpec cls;
            cls.lds = std::get<1>(kv.first) != 0;
            cls.nmax = *std::max_element(kv.second.begin(), kv.second.end());
            cls.items = 0;
            for (int n : kv.second) cls.items += hist[n];
            cls.order = single_class ? nullptr : (const uint32_t *)order.p + class_start[k];
            cls.host_n = nullptr;
            cls.plain_n = 0;
            cls.stream = nstreams > 1 ? class_streams[s] : stream;
            cls.queues = (uint32_t *)((uint8_t *)queues.p + kQueueBytes * kClassQueues * k);
            for (int j = 0; j < 2; ++j) {
                cls.spill_queue[j] = (uint32_t *)((uint8_t *)cls.queues + kQueueBytes * (j + 1));
                cls.spill_list[j] = (uint32_t *)spill.p + j * db.count + class_start[k];
            }
            if (defer && single_class && db.uniform_n > 0) {
                // a part of a larger batch whose last launch runs once for
                // all parts (solve_last_launch)
                cls.defer_last = true;
                cls.spill_queue[1] = defer->spill_queue[1];
                cls.spill_list[1] = defer->spill_list[1];
                cls.list1_base = defer->list1_base;
            }
            rc = launch_class(db, cls, false);
            if (rc) return rc;
        }
        if (nstreams > 1)
            for (int j = 0; j < nstreams; ++j) {
                HIP_TRY(hipEventRecord(ev_join[j], class_streams[j]), "hipEventRecord");
                HIP_TRY(hipStreamWaitEvent(stream, ev_join[j], 0), "wait join");
            }
    }
    HIP_TRY(hipEventRecord(ev1, stream), "hipEventRecord");

    // the deferred count and the violation in one read: automata whose
    // frontier outgrew a group's arena go on to the device-wide path
    static_assert(kCtrDeferred + 4 <= kCtrViolation && kCtrViolation + 8 <= 16, "counter layout");
    if (head_out) {
        HIP_TRY(hipMemcpyAsync(head_out, ctrs.p, 16, hipMemcpyDeviceToHost, stream), "D2H");
        last_launches = launches;
        return SSW_OK;
    }
    alignas(8) uint8_t head[16];
    HIP_TRY(hipMemcpyAsync(head, ctrs.p, sizeof head, hipMemcpyDeviceToHost, stream), "D2H");
    HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    uint32_t ndef = 0;
    unsigned long long viol = ~0ull;
    std::memcpy(&ndef, head + kCtrDeferred, 4);
    std::memcpy(&viol, head + kCtrViolation, 8);
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
    last_ms = ms;
    last_launches = launches;
    if (ndef) {
        std::vector<uint32_t> idx(ndef);
        HIP_TRY(hipMemcpyAsync(idx.data(), deferred.p, 4ull * ndef, hipMemcpyDeviceToHost, stream), "D2H deferred");
        HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
        deferred_total += ndef;
        if ((rc = need_host_n())) return rc;
        rc = large.solve_indices(*this, db, host_n, idx);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(&viol, (uint8_t *)ctrs.p + kCtrViolation, 8, hipMemcpyDeviceToHost, stream), "D2H");
        HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
    if (violation_out) *violation_out = viol == ~0ull ? -1 : (int64_t)viol;
    return SSW_OK;
}

// The last launch of a uniform batch of n <= 16 solved in parts whose
// chains stopped before it: it searches the automata listed in list1
// (indices into the whole batch, appended by the parts' launches through
// queue1) with the larger ring and byte tables, on the context's stream; the
// 16 counter bytes (deferred count, violation) are copied to head_out
// without waiting.
int ssw_context::solve_last_launch(const DeviceBatch &db, uint32_t *queue1, uint32_t *list1, int64_t items,
                                   void *head_out) {
    const int n = db.uniform_n;
    if (n < 1 || n > kMaxLdsN) return fail(SSW_ERR_INTERNAL, "last launch of n=%d on its own", n);
    launches = 0;
    ++call_gen;
    records.clear();
    HIP_TRY(reset_counters((uint32_t *)((uint8_t *)ctrs.p + kCtrDeferred),
                           (unsigned long long *)((uint8_t *)ctrs.p + kCtrViolation),
                           (unsigned long long *)prog.p, (unsigned long long *)lstats.p, kLaunchStatWords, stream),
            "reset counters");
    HIP_TRY(hipEventRecord(ev0, stream), "hipEventRecord");
    ClassSpec cls;
    cls.nmax = n;
    cls.lds = true;
    cls.items = items;
    cls.stream = stream;
    cls.queues = queue1;
    const ssw_tuning &t = tune[n];
    ssw_tuning t2 = t;
    if (t.spill_group_lanes > 0) t2.group_lanes = t.spill_group_lanes;
    t2.block_threads = t2.group_lanes > 64 ? t2.group_lanes : 0;
    t2.chunk = 1;
    t2.arena_lds = t.spill_arena_lds;
    t2.lane_entries = t.spill_lane_entries > 0 ? t.spill_lane_entries : t.lane_entries;
    PassIO last;
    last.kind = kPassSpillTo;
    last.in_list = list1;
    last.in_queue = queue1;
    last.byte_tab = true;
    const int rc = launch_pass(db, cls, t2, false, last);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(ev1, stream), "hipEventRecord");
    HIP_TRY(hipMemcpyAsync(head_out, ctrs.p, 16, hipMemcpyDeviceToHost, stream), "D2H");
    last_launches = launches;
    return SSW_OK;
}

// Blocks of batch_kernel resident per CU for a launch shape, remembered per
// shape (the runtime query costs more than a map lookup on every solve).
hipError_t ssw_context::occupancy(int variant, int G, int block, size_t lds, int *out) {
    const auto key = std::make_tuple(variant, G, block, lds);
    auto it = occ_cache.find(key);
    if (it != occ_cache.end()) {
        *out = it->second;
        return hipSuccess;
    }
    const hipError_t e = batch_occupancy(variant, G, block, lds, out);
    if (e == hipSuccess) occ_cache[key] = *out;
    return e;
}

// ---------------------------------------------------------------------------
// host-buffer front ends
// ---------------------------------------------------------------------------

// Runs fn(lo, hi) over [0, count) on up to 16 host threads (one when the
// range is short).
void ssw_context::parallel_ranges(int64_t count, int64_t per_thread,
                                  const std::function<void(int64_t, int64_t)> &fn) {
    const int threads = (int)std::max<int64_t>(1, std::min<int64_t>(pool.size(), count / per_thread));
    if (threads == 1 || !pool.idle()) {
        fn((int64_t)0, count);
        return;
    }
    pool.run([&fn, count, threads](int t) { fn(count * t / threads, count * (t + 1) / threads); }, threads);
    pool.wait();
}

int ssw_context::ensure_stage(size_t bytes) {
    if (h_stage && h_stage_bytes >= bytes) return SSW_OK;
    if (h_stage) hipHostFree(h_stage);
    h_stage = nullptr;
    h_stage_bytes = 0;
    const size_t want = std::max<size_t>(bytes, 1u << 20);
    HIP_TRY(hipHostMalloc(&h_stage, want, hipHostMallocDefault), "hipHostMalloc");
    h_stage_bytes = want;
    return SSW_OK;
}

int ssw_context::solve_host_packed(const ssw_batch *b, int mode, int64_t *violation_out) {
    const int64_t count = b->count;
    const int64_t nchunks = std::min<int64_t>(4, count / kMinHostChunk);
    if (!host_checked && mode == kModeBatch && b->uniform_n > 0 && b->uniform_n <= kMaxLdsN && !b->offset &&
        !b->witness && nchunks > 1) {
        // the caller's rows (pageable memory) copied into pinned staging and
        // checked by host threads chunk by chunk, under the transfers and
        // kernels of the chunks before (host_pipeline)
        const int32_t n = b->uniform_n;
        int rc = ensure_stage((size_t)count * 2 * n);
        if (rc) return rc;
        uint8_t *packed = (uint8_t *)h_stage;
        return host_pipeline(
            n, count, nchunks,
            [&](int64_t lo, int64_t hi) -> int64_t {
                const size_t row = 2u * (size_t)n;
                const uint8_t *src = b->delta + (size_t)lo * row;
                std::memcpy(packed + (size_t)lo * row, src, (size_t)(hi - lo) * row);
                uint8_t most = 0;
                for (size_t q = 0, e = (size_t)(hi - lo) * row; q < e; ++q) most = std::max(most, src[q]);
                if (most < (uint8_t)n) return -1;
                for (int64_t i = lo; i < hi; ++i)
                    for (size_t q = 0; q < row; ++q)
                        if (b->delta[(size_t)i * row + q] >= n) return i;
                return -1;
            },
            [&](int64_t first) { return fail(SSW_ERR_ARG, "automaton %lld: a transition >= n=%d", (long long)first, n); },
            host_no_length ? nullptr : b->length, b->sync, b->expanded, b->bound, violation_out);
    }
    // Large batches solved here (witness words, n > 16, mixed sizes) go
    // through pinned memory both ways: the pool copies uniform rows into the
    // staging buffer while checking them, and the results come back into
    // pinned memory and are copied out by the pool (the driver's copies from
    // and to pageable memory run at a fraction of the pinned rate)
    const bool bounce = mode == kModeBatch && count >= kMinHostChunk;
    // (the int32 entry points pass rows already in the staging)
    const bool staged = bounce && !host_checked && b->uniform_n > 0 && !b->offset;
    // extent of the packed transitions and validation (first bad automaton)
    uint64_t extent = 0;
    if (host_checked && b->uniform_n > 0 && !b->offset) {
        extent = (uint64_t)count * 2u * (uint64_t)b->uniform_n;
    } else if (staged) {
        const int32_t n = b->uniform_n;
        if (n < 1 || n > SSW_MAX_N) return fail(SSW_ERR_ARG, "n=%d out of [1,30]", n);
        const size_t row = 2u * (size_t)n;
        extent = (uint64_t)count * row;
        int rc = ensure_stage((size_t)extent);
        if (rc) return rc;
        uint8_t *packed = (uint8_t *)h_stage;
        std::mutex m;
        int64_t bad = -1;
        parallel_ranges(count, 1 << 16, [&](int64_t lo, int64_t hi) {
            const uint8_t *src = b->delta + (size_t)lo * row;
            std::memcpy(packed + (size_t)lo * row, src, (size_t)(hi - lo) * row);
            uint8_t most = 0;
            for (size_t q = 0, e = (size_t)(hi - lo) * row; q < e; ++q) most = std::max(most, src[q]);
            if (most < (uint8_t)n) return;
            for (int64_t i = lo; i < hi; ++i)
                for (size_t q = 0; q < row; ++q)
                    if (src[(size_t)(i - lo) * row + q] >= n) {
                        std::lock_guard<std::mutex> lk(m);
                        if (bad < 0 || i < bad) bad = i;
                        return;
                    }
        });
        if (bad >= 0) return fail(SSW_ERR_ARG, "automaton %lld: a transition >= n=%d", (long long)bad, n);
    } else {
        std::mutex m;
        int64_t bad = -1;
        parallel_ranges(count, 1 << 16, [&](int64_t lo, int64_t hi) {
            uint64_t ext = 0;
            int64_t first = -1;
            for (int64_t i = lo; i < hi && first < 0; ++i) {
                const int n = b->uniform_n > 0 ? b->uniform_n : b->n_per[i];
                if (n < 1 || n > SSW_MAX_N) {
                    first = i;
                    break;
                }
                const uint64_t o = b->offset ? b->offset[i] : (uint64_t)i * 2u * (uint64_t)n;
                const uint8_t *d = b->delta + o;
                for (int q = 0; q < 2 * n; ++q)
                    if (d[q] >= n) first = i;
                ext = std::max<uint64_t>(ext, o + 2u * (uint64_t)n);
            }
            std::lock_guard<std::mutex> lk(m);
            extent = std::max(extent, ext);
            if (first >= 0 && (bad < 0 || first < bad)) bad = first;
        });
        if (bad >= 0) {
            const int n = b->uniform_n > 0 ? b->uniform_n : b->n_per[bad];
            if (n < 1 || n > SSW_MAX_N)
                return fail(SSW_ERR_ARG, "automaton %lld: n=%d out of [1,30]", (long long)bad, n);
            return fail(SSW_ERR_ARG, "automaton %lld: a transition >= n=%d", (long long)bad, n);
        }
    }
    int rc;
    if ((rc = ensure(d_delta, extent, "delta"))) return rc;
    if ((rc = ensure(d_len, 4ull * count, "length"))) return rc;
    HIP_TRY(hipMemcpyAsync(d_delta.p, staged ? h_stage : (const void *)b->delta, extent, hipMemcpyHostToDevice,
                           stream),
            "H2D");
    DeviceBatch db;
    db.count = count;
    db.uniform_n = b->uniform_n;
    db.bound = b->bound;
    db.delta = (const uint8_t *)d_delta.p;
    db.length = (int32_t *)d_len.p;
    db.witness_stride = b->witness_stride;
    if (b->uniform_n <= 0) {
        if ((rc = ensure(d_nper, count, "n_per"))) return rc;
        HIP_TRY(hipMemcpyAsync(d_nper.p, b->n_per, count, hipMemcpyHostToDevice, stream), "H2D");
        db.n_per = (const uint8_t *)d_nper.p;
    }
    if (b->offset) {
        if ((rc = ensure(d_off, 8ull * count, "offset"))) return rc;
        HIP_TRY(hipMemcpyAsync(d_off.p, b->offset, 8ull * count, hipMemcpyHostToDevice, stream), "H2D");
        db.offset = (const uint64_t *)d_off.p;
    }
    if (b->sync) {
        if ((rc = ensure(d_sync, count, "sync"))) return rc;
        db.sync = (uint8_t *)d_sync.p;
    }
    if (b->expanded) {
        if ((rc = ensure(d_exp, 8ull * count, "expanded"))) return rc;
        db.expanded = (int64_t *)d_exp.p;
    }
    if (b->witness) {
        if ((rc = ensure(d_wit, (uint64_t)b->witness_stride * count, "witness"))) return rc;
        db.witness = (uint8_t *)d_wit.p;
    }
    rc = mode == kModeLarge ? large.solve_all(*this, db, b->uniform_n > 0 ? nullptr : b->n_per, violation_out)
                            : solve_device(db, b->uniform_n > 0 ? nullptr : b->n_per, mode == kModePlain,
                                           violation_out);
    if (rc) return rc;
    if (bounce) {
        // results into pinned memory, then to the caller on the pool
        const uint64_t wrow = b->witness ? (uint64_t)b->witness_stride : 0u;
        const uint64_t o_sync = align_up(4ull * count, 64), o_exp = align_up(o_sync + (uint64_t)count, 64),
                       o_wit = align_up(o_exp + 8ull * count, 64), total = o_wit + wrow * count;
        if ((rc = ensure_out((size_t)total))) return rc;
        uint8_t *h = (uint8_t *)h_out;
        if (!host_no_length)
            HIP_TRY(hipMemcpyAsync(h, db.length, 4ull * count, hipMemcpyDeviceToHost, stream), "D2H");
        if (b->sync) HIP_TRY(hipMemcpyAsync(h + o_sync, db.sync, count, hipMemcpyDeviceToHost, stream), "D2H");
        if (b->expanded)
            HIP_TRY(hipMemcpyAsync(h + o_exp, db.expanded, 8ull * count, hipMemcpyDeviceToHost, stream), "D2H");
        if (b->witness)
            HIP_TRY(hipMemcpyAsync(h + o_wit, db.witness, wrow * count, hipMemcpyDeviceToHost, stream), "D2H");
        HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
        parallel_ranges(count, 1 << 15, [&](int64_t lo, int64_t hi) {
            const size_t k = (size_t)(hi - lo);
            if (!host_no_length) std::memcpy(b->length + lo, h + 4ull * lo, 4 * k);
            if (b->sync) std::memcpy(b->sync + lo, h + o_sync + lo, k);
            if (b->expanded) std::memcpy(b->expanded + lo, h + o_exp + 8ull * lo, 8 * k);
            if (b->witness) std::memcpy(b->witness + wrow * lo, h + o_wit + wrow * lo, wrow * k);
        });
        return SSW_OK;
    }
    if (!host_no_length)
        HIP_TRY(hipMemcpyAsync(b->length, db.length, 4ull * count, hipMemcpyDeviceToHost, stream), "D2H");
    if (b->sync) HIP_TRY(hipMemcpyAsync(b->sync, db.sync, count, hipMemcpyDeviceToHost, stream), "D2H");
    if (b->expanded)
        HIP_TRY(hipMemcpyAsync(b->expanded, db.expanded, 8ull * count, hipMemcpyDeviceToHost, stream), "D2H");
    if (b->witness)
        HIP_TRY(hipMemcpyAsync(b->witness, db.witness, (uint64_t)b->witness_stride * count,
                               hipMemcpyDeviceToHost, stream),
                "D2H");
    HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return SSW_OK;
}

int ssw_context::ensure_out(size_t bytes) {
    if (h_out && h_out_bytes >= bytes) return SSW_OK;
    if (h_out) hipHostFree(h_out);
    h_out = nullptr;
    h_out_bytes = 0;
    const size_t want = std::max<size_t>(bytes, 1u << 20);
    HIP_TRY(hipHostMalloc(&h_out, want, hipHostMallocDefault), "hipHostMalloc");
    h_out_bytes = want;
    return SSW_OK;
}

int ssw_context::solve_entry(const ssw_batch *b, int32_t memory, int mode, int64_t *violation_out) {
    if (!b || b->count < 0 || !b->delta || !b->length) return fail(SSW_ERR_ARG, "null batch field");
    if (b->count > (int64_t)0x7FFFFFFF) return fail(SSW_ERR_ARG, "count above 2^31-1");
    if (b->uniform_n <= 0 && !b->n_per) return fail(SSW_ERR_ARG, "n_per is NULL and uniform_n <= 0");
    if (b->uniform_n > SSW_MAX_N) return fail(SSW_ERR_ARG, "n=%d above %d", b->uniform_n, SSW_MAX_N);
    if (b->witness) {
        int nmax = b->uniform_n;
        if (nmax <= 0) {
            nmax = 1;
            if (memory == SSW_HOST)
                for (int64_t i = 0; i < b->count; ++i) nmax = std::max<int>(nmax, b->n_per[i]);
            else
                nmax = SSW_MAX_N;
        }
        const int need = ((nmax - 1) * (nmax - 1) + 7) / 8;
        if (b->witness_stride < std::max(1, need))
            return fail(SSW_ERR_ARG, "witness_stride %d < %d bytes", b->witness_stride, need);
    }
    if (violation_out) *violation_out = -1;
    if (b->count == 0) return SSW_OK;
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    solve_begin((uint64_t)b->count);
    int rc = solve_run(b, memory, mode, violation_out);
    solve_end();
    const int crc = checked_verdict();
    return rc ? rc : crc;
}

int ssw_context::checked_verdict() {
#ifdef SSW_CHECKED
    // bit 63: set; bits 40..62: source line; bits 0..39: the index or address
    unsigned long long v[2] = {0, 0};
    HIP_TRY(checked_take_batch(stream, &v[0]), "reading the bounds checks (ssw_batch.hip)");
    HIP_TRY(checked_take_large(stream, &v[1]), "reading the bounds checks (ssw_large.hip)");
    for (int k = 0; k < 2; ++k)
        if (v[k])
            return fail(SSW_ERR_INTERNAL, "bounds check failed: %s line %u formed index or LDS address %llu",
                        k == 0 ? "ssw_batch.hip" : "ssw_large.hip", (unsigned)((v[k] >> 40) & 0x7FFFFFu),
                        v[k] & ((1ull << 40) - 1));
#endif
    return SSW_OK;
}

int ssw_context::solve_run(const ssw_batch *b, int32_t memory, int mode, int64_t *violation_out) {
    int rc;
    if (memory == SSW_HOST) {
        rc = solve_host_packed(b, mode, violation_out);
    } else if (memory == SSW_DEVICE) {
        DeviceBatch db;
        db.count = b->count;
        db.uniform_n = b->uniform_n;
        db.bound = b->bound;
        db.n_per = b->n_per;
        db.offset = b->offset;
        db.delta = b->delta;
        db.length = b->length;
        db.sync = b->sync;
        db.expanded = b->expanded;
        db.witness = b->witness;
        db.witness_stride = b->witness_stride;
        std::vector<uint8_t> host_n;
        int64_t size_hist[32] = {};
        bool counted = false;
        if (b->uniform_n <= 0 && mode != kModeLarge && b->count > 0) {
            // the sizes counted on the device: no copy of n_per and no pass
            // over it here unless a size is out of range (then the copy
            // below names the first such automaton)
            rc = ensure(sizes, 32 * sizeof(unsigned long long), "size counts");
            if (rc) return rc;
            HIP_TRY(count_sizes(b->n_per, b->count, (unsigned long long *)sizes.p, stream), "count sizes");
            HIP_TRY(hipMemcpyAsync(size_hist, sizes.p, sizeof size_hist, hipMemcpyDeviceToHost, stream), "D2H sizes");
            HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
            counted = size_hist[0] == 0 && size_hist[31] == 0;
        }
        if (b->uniform_n <= 0 && !counted) {
            host_n.resize(b->count);
            hipError_t e = hipMemcpyAsync(host_n.data(), b->n_per, b->count, hipMemcpyDeviceToHost, stream);
            if (e == hipSuccess) e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return hip_fail(e, "D2H n_per");
            for (int64_t i = 0; i < b->count; ++i)
                if (host_n[i] < 1 || host_n[i] > SSW_MAX_N)
                    return fail(SSW_ERR_ARG, "automaton %lld: n=%d out of [1,30]", (long long)i, (int)host_n[i]);
        }
        const uint8_t *hn = b->uniform_n > 0 || counted ? nullptr : host_n.data();
        rc = mode == kModeLarge ? large.solve_all(*this, db, hn, violation_out)
                                : solve_device(db, hn, mode == kModePlain, violation_out, counted ? size_hist : nullptr);
    } else {
        rc = fail(SSW_ERR_ARG, "memory kind %d", memory);
    }
    return rc;
}

// The allocation-free entry point.  Device buffers only, one uniform n <= 16,
// no witness words, at most kNoAllocMaxCount automata.  Nothing is allocated:
// the counters, stream and events are the context's (set up at creation),
// the work queues, spill lists and the last launch's HBM rings are this GPU's
// static scratch.  Every piece of the scratch a launch reads is written
// earlier in the same call (queues by their set-up kernels, list entries by
// the launch that appends them, a ring's window by the group that moves it
// there), so what an earlier call left behind is never read.
int ssw_context::solve_noalloc(const ssw_batch *b, int64_t *violation_out) {
    if (!b || b->count < 0 || !b->delta || !b->length) return fail(SSW_ERR_ARG, "null batch field");
    if (violation_out) *violation_out = -1;
    if (b->uniform_n < 1 || b->uniform_n > kNoAllocMaxN || b->n_per)
        return fail(SSW_ERR_ARG, "the allocation-free entry point takes one uniform n in [1,%d] (got uniform_n=%d%s)",
                    kNoAllocMaxN, b->uniform_n, b->n_per ? " with n_per" : "");
    if (b->witness) return fail(SSW_ERR_ARG, "the allocation-free entry point computes no witness words");
    if ((uint64_t)b->count > kNoAllocMaxCount)
        return fail(SSW_ERR_ARG, "%lld automata: the allocation-free entry point's scratch holds at most %llu",
                    (long long)b->count, (unsigned long long)kNoAllocMaxCount);
    if (b->count == 0) return SSW_OK;
    if (!noalloc_base)
        return fail(SSW_ERR_DEVICE, "no static scratch on device %d: %s", device, hipGetErrorString(noalloc_status));
    std::lock_guard<std::mutex> gpu_lk(noalloc_call_lock(device));
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    DeviceBatch db;
    db.count = b->count;
    db.uniform_n = b->uniform_n;
    db.bound = b->bound;
    db.offset = b->offset;
    db.delta = b->delta;
    db.length = b->length;
    db.sync = b->sync;
    db.expanded = b->expanded;
    uint8_t *base = (uint8_t *)noalloc_base;
    ClassSpec cls;
    cls.lds = true;
    cls.nmax = b->uniform_n;
    cls.items = b->count;
    cls.stream = stream;
    cls.queues = (uint32_t *)(base + kNoAllocQueueOff);
    for (int j = 0; j < 3; ++j) {
        cls.spill_queue[j] = (uint32_t *)(base + kNoAllocQueueOff + kQueueBytes * (uint64_t)(j + 1));
        cls.spill_list[j] = (uint32_t *)(base + kNoAllocListOff) + (uint64_t)j * kNoAllocMaxCount;
    }
    cls.static_slots = base + kNoAllocSlotOff;
    cls.static_slot_bytes = kNoAllocSlots * kNoAllocSlotBytes;
    solve_begin((uint64_t)b->count);
    launches = 0;
    int rc = SSW_OK;
    records.clear();
    hipError_t e = reset_counters((uint32_t *)((uint8_t *)ctrs.p + kCtrDeferred),
                                  (unsigned long long *)((uint8_t *)ctrs.p + kCtrViolation),
                                  (unsigned long long *)prog.p, (unsigned long long *)lstats.p, kLaunchStatWords,
                                  stream);
    if (e == hipSuccess) e = hipEventRecord(ev0, stream);
    if (e != hipSuccess) rc = hip_fail(e, "set-up");
    if (rc == SSW_OK) rc = launch_class(db, cls, false);
    alignas(8) uint8_t head[16];
    if (rc == SSW_OK) {
        e = hipEventRecord(ev1, stream);
        if (e == hipSuccess) e = hipMemcpyAsync(head, ctrs.p, sizeof head, hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
        if (e != hipSuccess) rc = hip_fail(e, "allocation-free solve");
        last_ms = ms;
        last_launches = launches;
    }
    solve_end();
    if (rc) return rc;
    uint32_t ndef = 0;
    unsigned long long viol = ~0ull;
    std::memcpy(&ndef, head + kCtrDeferred, 4);
    std::memcpy(&viol, head + kCtrViolation, 8);
    if (ndef) return fail(SSW_ERR_INTERNAL, "%u automata deferred on the allocation-free path", ndef);
    if (violation_out) *violation_out = viol == ~0ull ? -1 : (int64_t)viol;
    return SSW_OK;
}

// Packs automata [lo, hi) of the int32 arrays into the staging bytes,
// checking every transition; returns the first bad automaton or -1.  The
// check is a running unsigned maximum, so the copy loop vectorises; only a
// range holding a bad transition is scanned again to find the first one.
static int64_t pack_range(int32_t n, const int32_t *A, const int32_t *B, uint8_t *packed, int64_t lo, int64_t hi) {
    const size_t row = (size_t)n;
    uint32_t most = 0;
    for (int64_t i = lo; i < hi; ++i) {
        uint8_t *d = packed + (size_t)i * 2 * row;
        const int32_t *a = A + (size_t)i * row, *b = B + (size_t)i * row;
        for (size_t q = 0; q < row; ++q) {
            const uint32_t x = (uint32_t)a[q], y = (uint32_t)b[q];
            most = std::max(most, std::max(x, y));
            d[q] = (uint8_t)x;
            d[row + q] = (uint8_t)y;
        }
    }
    if (most < (uint32_t)n) return -1;
    for (int64_t i = lo; i < hi; ++i)
        for (size_t q = 0; q < row; ++q)
            if ((uint32_t)A[i * row + q] >= (uint32_t)n || (uint32_t)B[i * row + q] >= (uint32_t)n) return i;
    return -1;
}

// A uniform batch of n <= 16 from host memory as a pipeline of `nchunks`
// chunks.  Host threads fill the pinned staging buffer chunk by chunk
// (fill(lo, hi) writes the packed rows of automata [lo, hi) and returns the
// first invalid one there, or -1); chunk k's rows are copied on a copy stream
// while chunk k - 1's launches run, into the device set k & 1; its results
// come back into pinned memory on a third stream once its launches are done
// and are copied to the caller while chunk k + 1 runs.  The host waits only
// for filled chunks and finished results; the solves' counters (n <= 16: no
// automaton is deferred) are read once at the end.  An invalid automaton
// drains the work already queued and returns bad_at(index).
// With the three-launch chain (n = 16) the chunks' rows go to their place in
// one whole-batch device buffer and each chunk runs only the first two
// launches, appending what the last launch takes to one shared list; that
// launch runs once after the last chunk and the results come back in one
// copy: the chunked chains each ended in the tail of their longest searches.
// (Solving the whole batch once after pipelined copies instead was measured:
// the 1M n = 16 batch gains little, since its host packing is as long as the
// copies, while the n = 13 Verify workload loses the overlap of its packing
// with the solves: 3.3 -> 4.2 ms.)
int ssw_context::host_pipeline(int32_t n, int64_t count, int64_t nchunks,
                               const std::function<int64_t(int64_t, int64_t)> &fill,
                               const std::function<int(int64_t)> &bad_at, int32_t *length_out, uint8_t *sync_out,
                               int64_t *expanded_out, int32_t bound, int64_t *violation_out) {
    if (violation_out) *violation_out = -1;
    uint8_t *packed = (uint8_t *)h_stage;
    // n with the three-launch chain: every chunk runs the first two launches
    // as it arrives, and the last launch (the longest searches, whose tails
    // would otherwise end every chunk) runs once over what the chunks hand
    // on; the chunks' rows and results have their place in one set of
    // whole-batch buffers
    const ssw_tuning &tn = tune[n];
    const bool defer = tn.narrow_bits >= 8 && tn.narrow_bits < n && tn.spill_arena_lds > 0 && nchunks <= 4;
    // chunk boundaries: the first chunk has half the others' size, so the
    // device starts sooner (the later chunks are packed while the earlier ones
    // run): 4.02 -> 3.87 ms for the 1M n = 16 packed rows, about 3.05 -> 2.95 ms
    // for the 2^21-automaton n = 13 Verify
    auto chunk_lo = [count, nchunks](int64_t k) -> int64_t {
        return k == 0 ? 0 : count * (2 * k - 1) / (2 * nchunks - 1);
    };
    const int workers = pool.size();
    std::vector<std::atomic<int>> filled(nchunks);
    std::vector<std::atomic<int64_t>> bad(nchunks);
    for (int64_t k = 0; k < nchunks; ++k) {
        filled[k].store(0);
        bad[k].store(-1);
    }
    // automata per device buffer set: the whole batch, or the largest chunk
    uint64_t most = (uint64_t)count;
    if (!defer) {
        most = 0;
        for (int64_t k = 0; k < nchunks; ++k) most = std::max<uint64_t>(most, (uint64_t)(chunk_lo(k + 1) - chunk_lo(k)));
    }
    const uint64_t out_row = 4u + (sync_out ? 1u : 0u) + (expanded_out ? 8u : 0u);
    const uint64_t half = align_up(most * out_row + 16, 256);
    // the shared list of the last launch: its 64 queue segments each hold
    // nchunks + 1 more entries than a 64th of the batch, so the chunks'
    // appends (at most ceil(chunk / 64) per segment and chunk) never run
    // into the next segment
    const int64_t list_items = count + 64 * (nchunks + 1);
    int rc;
    if (defer) {
        if ((rc = ensure(d_delta, (uint64_t)count * 2u * (uint64_t)n, "delta"))) return rc;
        if ((rc = ensure(d_len, 4ull * count, "length"))) return rc;
        if (sync_out && (rc = ensure(d_sync, (uint64_t)count, "sync"))) return rc;
        if (expanded_out && (rc = ensure(d_exp, 8ull * count, "expanded"))) return rc;
        if ((rc = ensure(deferred, 4ull * count, "deferred list"))) return rc;
        if ((rc = ensure(pipe_q, kQueueBytes, "work queue"))) return rc;
        if ((rc = ensure(pipe_list, 4ull * list_items, "spill list"))) return rc;
        if ((rc = ensure(pipe_patch, 16ull * (list_items + 1), "result records"))) return rc;
        if ((rc = ensure_out(half + 16u * (uint64_t)(nchunks + 1) + 16u * (kPatchFirst + 1)))) return rc;
        HIP_TRY(fill_u32((uint32_t *)pipe_patch.p, 4, 0u, stream), "clear");
        // set up before the first chunk's launches, on their stream
        HIP_TRY(init_queues((uint32_t *)pipe_q.p, list_items, stream, true), "init queues");
    } else {
        if ((rc = ensure(d_delta, most * 2u * (uint64_t)n, "delta"))) return rc;
        if ((rc = ensure(d_len, 4ull * most, "length"))) return rc;
        if (sync_out && (rc = ensure(d_sync, most, "sync"))) return rc;
        if (expanded_out && (rc = ensure(d_exp, 8ull * most, "expanded"))) return rc;
        if ((rc = ensure(d_delta2, most * 2u * (uint64_t)n, "delta"))) return rc;
        if ((rc = ensure(d_len2, 4ull * most, "length"))) return rc;
        if (sync_out && (rc = ensure(d_sync2, most, "sync"))) return rc;
        if (expanded_out && (rc = ensure(d_exp2, 8ull * most, "expanded"))) return rc;
        if ((rc = ensure_out(2 * half + 16u * (uint64_t)nchunks))) return rc;
    }
    uint8_t *heads = (uint8_t *)h_out + (defer ? 1 : 2) * half;
    // inputs and outputs on streams of their own, so that neither waits
    // behind the other's dependencies
    hipStream_t copy = class_streams[kClassStreams - 1], back = class_streams[kClassStreams - 2];
    // host part of a set's results: lengths, then flags, then counts
    const auto out_len = [&](int s) { return (int32_t *)((uint8_t *)h_out + s * half); };
    const auto out_sync = [&](int s) { return (uint8_t *)(out_len(s) + most); };
    const auto out_exp = [&](int s) {
        return (int64_t *)((uint8_t *)h_out + s * half + align_up(most * (4u + (sync_out ? 1u : 0u)), 8));
    };
    std::atomic<bool> stop{false};
    auto work = [&](int t) {
        for (int64_t k = 0; k < nchunks && !stop.load(std::memory_order_relaxed); ++k) {
            const int64_t lo = chunk_lo(k), len = chunk_lo(k + 1) - lo;
            const int64_t first = fill(lo + len * t / workers, lo + len * (t + 1) / workers);
            if (first >= 0) {
                int64_t cur = bad[k].load();
                while ((cur < 0 || first < cur) && !bad[k].compare_exchange_weak(cur, first)) {
                }
            }
            filled[k].fetch_add(1, std::memory_order_release);
        }
    };
    pool.run(work, workers);
    const auto join_all = [&] { pool.wait(); };
    const auto deliver = [&](int64_t k) -> int {  // chunk k's results to the caller
        const int s = (int)(k & 1);
        HIP_TRY(hipEventSynchronize(pipe_ev[4 + s]), "hipEventSynchronize");
        const int64_t lo = chunk_lo(k), len = chunk_lo(k + 1) - lo;
        if (length_out) std::memcpy(length_out + lo, out_len(s), 4ull * len);
        if (sync_out) std::memcpy(sync_out + lo, out_sync(s), (size_t)len);
        if (expanded_out) std::memcpy(expanded_out + lo, out_exp(s), 8ull * len);
        call_done_base.fetch_add((uint64_t)len);  // the progress feed's in-flight count
        return SSW_OK;
    };
    const auto drain = [&](int code) {
        stop.store(true);
        join_all();
        hipStreamSynchronize(copy);
        hipStreamSynchronize(stream);
        hipStreamSynchronize(back);
        return code;
    };
    // the smallest violating index of the solves whose counters are at heads
    // (part k's indices start at lo(k)); no automaton of n <= 16 is deferred
    const auto read_heads = [&](int64_t parts, const std::function<int64_t(int64_t)> &lo) -> int {
        int64_t viol = -1;
        for (int64_t k = 0; k < parts; ++k) {
            uint32_t ndef = 0;
            unsigned long long v = ~0ull;
            std::memcpy(&ndef, heads + 16 * k + kCtrDeferred, 4);
            std::memcpy(&v, heads + 16 * k + kCtrViolation, 8);
            if (ndef) return fail(SSW_ERR_INTERNAL, "an automaton of n=%d was deferred", n);
            if (v != ~0ull && (viol < 0 || lo(k) + (int64_t)v < viol)) viol = lo(k) + (int64_t)v;
        }
        if (violation_out) *violation_out = viol;
        return SSW_OK;
    };
    if (defer) {
        // chunk k's results are copied back (on the back stream) once its
        // first two launches are done and handed to the caller when the host
        // sees them there (pipe_ev[4 + k]); those of the automata the last
        // launch takes are stale in that copy and are patched from records
        // the gather kernel writes after the last launch
        int64_t delivered = 0;
        const auto deliver_done = [&](int64_t upto, bool wait) -> int {
            while (delivered < upto) {
                if (wait) {
                    HIP_TRY(hipEventSynchronize(pipe_ev[4 + delivered]), "hipEventSynchronize");
                } else {
                    const hipError_t q = hipEventQuery(pipe_ev[4 + delivered]);
                    if (q == hipErrorNotReady) return SSW_OK;
                    if (q != hipSuccess) return hip_fail(q, "hipEventQuery");
                }
                const int64_t lo = chunk_lo(delivered), len = chunk_lo(delivered + 1) - lo;
                parallel_ranges(len, 1 << 16, [&](int64_t a, int64_t b) {
                    if (length_out) std::memcpy(length_out + lo + a, out_len(0) + lo + a, 4ull * (b - a));
                    if (sync_out) std::memcpy(sync_out + lo + a, out_sync(0) + lo + a, (size_t)(b - a));
                    if (expanded_out) std::memcpy(expanded_out + lo + a, out_exp(0) + lo + a, 8ull * (b - a));
                });
                call_done_base.fetch_add((uint64_t)len);  // the progress feed's in-flight count
                ++delivered;
            }
            return SSW_OK;
        };
        ClassSpec part;
        part.spill_queue[1] = (uint32_t *)pipe_q.p;
        part.spill_list[1] = (uint32_t *)pipe_list.p;
        for (int64_t k = 0; k < nchunks; ++k) {
            while (filled[k].load(std::memory_order_acquire) < workers) {
                if ((rc = deliver_done(k, false))) return drain(rc);
                std::this_thread::yield();
            }
            if (bad[k].load() >= 0) {
                const int64_t first = bad[k].load();
                drain(0);
                return bad_at(first);
            }
            const int s = (int)(k & 1);
            const int64_t lo = chunk_lo(k), len = chunk_lo(k + 1) - lo;
            hipError_t e = hipMemcpyAsync((uint8_t *)d_delta.p + (size_t)lo * 2 * n, packed + (size_t)lo * 2 * n,
                                          (size_t)len * 2 * n, hipMemcpyHostToDevice, copy);
            if (e == hipSuccess) e = hipEventRecord(pipe_ev[s], copy);
            if (e == hipSuccess) e = hipStreamWaitEvent(stream, pipe_ev[s], 0);
            if (e != hipSuccess) return drain(hip_fail(e, "pipelined chunk input"));
            DeviceBatch db;
            db.count = len;
            db.uniform_n = n;
            db.bound = bound;
            db.delta = (const uint8_t *)d_delta.p + (size_t)lo * 2 * n;
            db.length = (int32_t *)d_len.p + lo;
            db.sync = sync_out ? (uint8_t *)d_sync.p + lo : nullptr;
            db.expanded = expanded_out ? (int64_t *)d_exp.p + lo : nullptr;
            part.list1_base = (uint32_t)lo;
            rc = solve_device(db, nullptr, false, nullptr, nullptr, heads + 16 * k, &part);
            if (rc) return drain(rc);
            e = hipEventRecord(pipe_ev[2 + s], stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(back, pipe_ev[2 + s], 0);
            if (e == hipSuccess && length_out)
                e = hipMemcpyAsync(out_len(0) + lo, db.length, 4ull * len, hipMemcpyDeviceToHost, back);
            if (e == hipSuccess && sync_out)
                e = hipMemcpyAsync(out_sync(0) + lo, db.sync, (size_t)len, hipMemcpyDeviceToHost, back);
            if (e == hipSuccess && expanded_out)
                e = hipMemcpyAsync(out_exp(0) + lo, db.expanded, 8ull * len, hipMemcpyDeviceToHost, back);
            if (e == hipSuccess) e = hipEventRecord(pipe_ev[4 + k], back);
            if (e == hipSuccess) rc = deliver_done(k, false);
            if (e != hipSuccess || rc) return drain(e != hipSuccess ? hip_fail(e, "pipelined chunk output") : rc);
        }
        join_all();
        DeviceBatch whole;
        whole.count = count;
        whole.uniform_n = n;
        whole.bound = bound;
        whole.delta = (const uint8_t *)d_delta.p;
        whole.length = (int32_t *)d_len.p;
        whole.sync = sync_out ? (uint8_t *)d_sync.p : nullptr;
        whole.expanded = expanded_out ? (int64_t *)d_exp.p : nullptr;
        rc = solve_last_launch(whole, (uint32_t *)pipe_q.p, (uint32_t *)pipe_list.p, count, heads + 16 * nchunks);
        if (rc) return drain(rc);
        // the last launch's automata: their records, the first kPatchFirst of
        // them in the same copy as the count
        unsigned long long *rec = (unsigned long long *)(heads + align_up(16u * (uint64_t)(nchunks + 1), 16));
        hipError_t e = gather_listed((const uint32_t *)pipe_q.p, (const uint32_t *)pipe_list.p, (uint32_t)list_items,
                                     (uint32_t)count, whole.length, whole.sync, whole.expanded,
                                     (unsigned long long *)pipe_patch.p, stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(rec, pipe_patch.p, 16u * (kPatchFirst + 1), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        if (e != hipSuccess) return drain(hip_fail(e, "pipelined results"));
        const uint64_t nrec = rec[0];
        std::vector<unsigned long long> more;
        if (nrec > kPatchFirst) {
            more.resize(2 * (nrec - kPatchFirst));
            e = hipMemcpy(more.data(), (uint8_t *)pipe_patch.p + 16u * (kPatchFirst + 1), 16u * (nrec - kPatchFirst),
                          hipMemcpyDeviceToHost);
            if (e != hipSuccess) return drain(hip_fail(e, "pipelined results"));
        }
        rc = deliver_done(nchunks, true);
        e = hipStreamSynchronize(back);
        join_all();
        if (rc) return rc;
        if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
        for (uint64_t q = 0; q < nrec; ++q) {
            const unsigned long long w0 = q < kPatchFirst ? rec[2 + 2 * q] : more[2 * (q - kPatchFirst)];
            const unsigned long long w1 = q < kPatchFirst ? rec[3 + 2 * q] : more[2 * (q - kPatchFirst) + 1];
            const uint32_t a = (uint32_t)w0;
            if ((int64_t)a >= count) return fail(SSW_ERR_INTERNAL, "result record of automaton %u", a);
            if (length_out) length_out[a] = (int32_t)(uint32_t)(w0 >> 32);
            if (sync_out) sync_out[a] = (uint8_t)(w1 >> 56);
            if (expanded_out) expanded_out[a] = (int64_t)(w1 & 0x00FFFFFFFFFFFFFFull);
        }
        float ms = 0;
        if (hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess) last_ms = ms;  // the last launch
        return read_heads(nchunks + 1, [&](int64_t k) { return k < nchunks ? chunk_lo(k) : (int64_t)0; });
    }
    for (int64_t k = 0; k < nchunks; ++k) {
        while (filled[k].load(std::memory_order_acquire) < workers) std::this_thread::yield();
        if (bad[k].load() >= 0) {
            const int64_t first = bad[k].load();
            drain(0);
            return bad_at(first);
        }
        const int s = (int)(k & 1);
        const int64_t lo = chunk_lo(k), len = chunk_lo(k + 1) - lo;
        DevBuf &dd = s ? d_delta2 : d_delta, &dl = s ? d_len2 : d_len, &ds = s ? d_sync2 : d_sync,
               &de = s ? d_exp2 : d_exp;
        hipError_t e = hipSuccess;
        // the set's inputs are overwritten once chunk k - 2's launches are done with them
        if (k >= 2) e = hipStreamWaitEvent(copy, pipe_ev[2 + s], 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dd.p, packed + (size_t)lo * 2 * n, (size_t)len * 2 * n, hipMemcpyHostToDevice, copy);
        if (e == hipSuccess) e = hipEventRecord(pipe_ev[s], copy);
        if (e == hipSuccess) e = hipStreamWaitEvent(stream, pipe_ev[s], 0);
        // the set's results are overwritten once chunk k - 2's have been copied back
        if (e == hipSuccess && k >= 2) e = hipStreamWaitEvent(stream, pipe_ev[4 + s], 0);
        if (e != hipSuccess) return drain(hip_fail(e, "pipelined chunk input"));
        DeviceBatch db;
        db.count = len;
        db.uniform_n = n;
        db.bound = bound;
        db.delta = (const uint8_t *)dd.p;
        db.length = (int32_t *)dl.p;
        db.sync = sync_out ? (uint8_t *)ds.p : nullptr;
        db.expanded = expanded_out ? (int64_t *)de.p : nullptr;
        rc = solve_device(db, nullptr, false, nullptr, nullptr, heads + 16 * k);
        if (rc) return drain(rc);
        e = hipEventRecord(pipe_ev[2 + s], stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(back, pipe_ev[2 + s], 0);
        if (e == hipSuccess && length_out) e = hipMemcpyAsync(out_len(s), dl.p, 4ull * len, hipMemcpyDeviceToHost, back);
        if (e == hipSuccess && sync_out) e = hipMemcpyAsync(out_sync(s), ds.p, (size_t)len, hipMemcpyDeviceToHost, back);
        if (e == hipSuccess && expanded_out)
            e = hipMemcpyAsync(out_exp(s), de.p, 8ull * len, hipMemcpyDeviceToHost, back);
        if (e == hipSuccess) e = hipEventRecord(pipe_ev[4 + s], back);
        if (e == hipSuccess && k >= 1) rc = deliver(k - 1);
        if (e != hipSuccess || rc) return drain(e != hipSuccess ? hip_fail(e, "pipelined chunk output") : rc);
    }
    rc = deliver(nchunks - 1);
    hipError_t e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipStreamSynchronize(back);
    join_all();
    if (rc) return rc;
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev0, ev1) == hipSuccess) last_ms = ms;  // the last chunk's kernels
    return read_heads(nchunks, chunk_lo);
}

int ssw_context::compute_i32(int32_t n, int64_t count, const int32_t *A, const int32_t *B,
                             int32_t *length_out, uint8_t *sync_out, uint8_t *witness_out,
                             int32_t witness_stride, int64_t *expanded_out, int32_t bound,
                             int64_t *violation_out) {
    if (n < 1 || n > SSW_MAX_N) return fail(SSW_ERR_ARG, "n=%d out of [1,30]", n);
    if (count < 0 || (count > 0 && (!A || !B))) return fail(SSW_ERR_ARG, "null argument");
    if (count > (int64_t)0x7FFFFFFF) return fail(SSW_ERR_ARG, "count above 2^31-1");
    if (violation_out) *violation_out = -1;
    // flatten like SlimGPUQueue.cs:108-126, packed to bytes in pinned staging
    // memory by host threads, checked on the way.  Large calls are cut into
    // chunks: the threads pack chunk k + 1 while chunk k is copied and solved,
    // so the host work overlaps the transfer and the kernels.
    int rc = ensure_stage((size_t)count * 2 * n);
    if (rc) return rc;
    uint8_t *packed = (uint8_t *)h_stage;
    // Witness words make the tails far longer (the longest searches also run
    // the witness passes), so those calls are solved whole.  n <= 16: the
    // chunks' transfers also overlap the kernels (host_pipeline).
    const int64_t nchunks = witness_out ? 1 : std::max<int64_t>(1, std::min<int64_t>(4, count / kMinHostChunk));
    if (nchunks > 1 && !witness_out && n <= kMaxLdsN) {
        HIP_TRY(hipSetDevice(device), "hipSetDevice");
        solve_begin((uint64_t)count);
        host_no_length = length_out == nullptr;
        rc = host_pipeline(n, count, nchunks,
                           [&](int64_t lo, int64_t hi) { return pack_range(n, A, B, packed, lo, hi); },
                           [&](int64_t first) {
                               int q = 0;
                               while (q < n && (uint32_t)A[first * n + q] < (uint32_t)n &&
                                      (uint32_t)B[first * n + q] < (uint32_t)n)
                                   ++q;
                               return fail(SSW_ERR_ARG, "automaton %lld state %d: transition out of [0,%d)",
                                           (long long)first, q, n);
                           },
                           length_out, sync_out, expanded_out, bound, violation_out);
        host_no_length = false;
        solve_end();
        const int crc = checked_verdict();
        return rc ? rc : crc;
    }
    auto chunk_lo = [&](int64_t k) { return count * k / nchunks; };
    const int workers = pool.size();
    std::vector<std::atomic<int>> packed_parts(nchunks);
    std::vector<std::atomic<int64_t>> bad(nchunks);
    for (int64_t k = 0; k < nchunks; ++k) {
        packed_parts[k].store(0);
        bad[k].store(-1);
    }
    // device buffers at the largest chunk's size before the pipeline starts
    {
        const uint64_t most = (uint64_t)((count + nchunks - 1) / nchunks);
        if ((rc = ensure(d_delta, most * 2u * (uint64_t)n, "delta"))) return rc;
        if ((rc = ensure(d_len, 4ull * most, "length"))) return rc;
        if (sync_out && (rc = ensure(d_sync, most, "sync"))) return rc;
        if (expanded_out && (rc = ensure(d_exp, 8ull * most, "expanded"))) return rc;
        if (witness_out && (rc = ensure(d_wit, (uint64_t)witness_stride * most, "witness"))) return rc;
    }
    std::atomic<bool> abort_pack{false};
    auto work = [&](int t) {
        for (int64_t k = 0; k < nchunks && !abort_pack.load(std::memory_order_relaxed); ++k) {
            const int64_t lo = chunk_lo(k), len = chunk_lo(k + 1) - lo;
            const int64_t first = pack_range(n, A, B, packed, lo + len * t / workers, lo + len * (t + 1) / workers);
            if (first >= 0) {
                int64_t cur = bad[k].load();
                while ((cur < 0 || first < cur) && !bad[k].compare_exchange_weak(cur, first)) {
                }
            }
            packed_parts[k].fetch_add(1, std::memory_order_release);
        }
    };
    pool.run(work, workers);
    auto join_all = [&] { pool.wait(); };
    static int32_t no_length;
    int64_t viol = -1;
    const auto bad_input = [&](int64_t first) {
        int q = 0;
        while (q < n && (uint32_t)A[first * n + q] < (uint32_t)n && (uint32_t)B[first * n + q] < (uint32_t)n) ++q;
        return fail(SSW_ERR_ARG, "automaton %lld state %d: transition out of [0,%d)", (long long)first, q, n);
    };
    for (int64_t k = 0; k < nchunks; ++k) {
        while (packed_parts[k].load(std::memory_order_acquire) < workers) std::this_thread::yield();
        const int64_t first = bad[k].load();
        if (first >= 0) {
            abort_pack.store(true);
            join_all();
            return bad_input(first);
        }
        const int64_t lo = chunk_lo(k), len = chunk_lo(k + 1) - lo;
        ssw_batch b;
        std::memset(&b, 0, sizeof b);
        b.count = len;
        b.uniform_n = n;
        b.delta = packed + (size_t)lo * 2 * n;
        b.length = length_out ? length_out + lo : &no_length;
        b.sync = sync_out ? sync_out + lo : nullptr;
        b.expanded = expanded_out ? expanded_out + lo : nullptr;
        b.witness = witness_out ? witness_out + (size_t)lo * witness_stride : nullptr;
        b.witness_stride = witness_stride;
        b.bound = bound;
        host_checked = true;
        host_no_length = length_out == nullptr;
        // the last chunk is packed: the pool is free for the solve's copies
        if (k == nchunks - 1) join_all();
        int64_t v = -1;
        rc = solve_entry(&b, SSW_HOST, kModeBatch, violation_out ? &v : nullptr);
        host_checked = false;
        host_no_length = false;
        if (rc) {
            abort_pack.store(true);
            join_all();
            return rc;
        }
        if (v >= 0 && (viol < 0 || lo + v < viol)) viol = lo + v;
    }
    join_all();
    if (violation_out) *violation_out = viol;
    return SSW_OK;
}

// ---------------------------------------------------------------------------
// per-launch statistics of the last batched solve
// ---------------------------------------------------------------------------

int ssw_context::launch_stats(ssw_launch_stat *out, int32_t cap, int32_t *count) {
    *count = (int32_t)records.size();
    if (records.empty() || cap == 0) return SSW_OK;
    HIP_TRY(hipSetDevice(device), "hipSetDevice");
    // the records in use, their copies summed
    const size_t used = records.size() * kCounterSlots * kStatWords;
    std::vector<unsigned long long> raw(used);
    HIP_TRY(hipMemcpyAsync(raw.data(), lstats.p, 8u * used, hipMemcpyDeviceToHost, stream), "D2H launch statistics");
    HIP_TRY(hipStreamSynchronize(stream), "hipStreamSynchronize");
    std::vector<unsigned long long> c(3 * records.size(), 0ull);
    for (uint32_t k = 0; k < (uint32_t)records.size(); ++k)
        for (uint32_t s = 0; s < kCounterSlots; ++s)
            for (uint32_t j = 0; j < 3; ++j) c[3 * k + j] += raw[((size_t)k * kCounterSlots + s) * kStatWords + j];
    for (size_t k = 0; k < records.size() && (int32_t)k < cap; ++k) {
        const LaunchRecord &r = records[k];
        ssw_launch_stat &o = out[k];
        std::memset(&o, 0, sizeof o);
        o.n = r.n;
        o.variant = r.variant;
        o.group_lanes = r.G;
        o.block_threads = r.block;
        o.grid = r.grid;
        o.lds_bytes = r.lds;
        o.ring_entries = r.ring;
        o.lane_entries = r.T;
        o.bitmap_bits = r.bitmap_bits;
        o.pass = r.pass;
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, lev[2 * k], lev[2 * k + 1]), "hipEventElapsedTime");
        o.ms = ms;
        o.automata = c[3 * k];
        o.subsets = c[3 * k + 1];
        o.expanded = c[3 * k + 2];
        const bool wave = r.G <= 64;
        if (r.variant == 0 || r.variant == 11 || r.variant == 12)
            std::snprintf(o.kernel, sizeof o.kernel, "GroupSolver<%s<%d>%s>", wave ? "WaveGroup" : "BlockGroup", r.G,
                          r.variant == 11 ? ",4" : (r.variant == 12 ? ",4,true" : ""));
        else
            std::snprintf(o.kernel, sizeof o.kernel, "SmallSolver<%s<%d>,%d,%s%s>", wave ? "WaveGroup" : "BlockGroup",
                          r.G, r.T, r.byte_tab ? "true" : "false",
                          r.variant == 10 ? ",false,true" : (r.witness ? ",true" : ""));
    }
    return SSW_OK;
}

// ---------------------------------------------------------------------------
// progress
// ---------------------------------------------------------------------------

int ssw_context::read_progress(ssw_progress *out) {
    std::lock_guard<std::mutex> lk(monitor_mu);
    // a solve running from the first read to the second: its count is the
    // one the done-in-this-solve counter belongs to
    const uint64_t g0 = solve_gen.load();
    const uint64_t cnt = call_count.load(), base = call_done_base.load();
    std::vector<unsigned long long> raw(kProgressBytes / 8);
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipMemcpyAsync(raw.data(), prog.p, kProgressBytes, hipMemcpyDeviceToHost, monitor);
    if (e == hipSuccess) e = hipStreamSynchronize(monitor);
    if (e != hipSuccess) return SSW_ERR_DEVICE;
    unsigned long long v[3] = {0, 0, 0};
    for (uint32_t s = 0; s < kCounterSlots; ++s)
        for (uint32_t j = 0; j < 3; ++j) v[j] += raw[(size_t)s * kSlotWords + j];
    const bool running = (g0 & 1u) && solve_gen.load() == g0;
    out->automata_done = v[0] + large_done.load();
    out->subsets_expanded = v[1] + large_subsets.load();
    out->launches = total_launches.load();
    out->deferred = deferred_total;
    out->in_flight = running && cnt > base + v[2] ? cnt - base - v[2] : 0;
    return SSW_OK;
}

// The thread looks every prog_ms: it reports while a solve runs (odd
// generation) and whenever the generation moved since its last look, so a
// solve that started and ended between two looks is still reported, with its
// final counts.  Stopping makes one last look.
void ssw_context::start_progress_thread() {
    stop_progress_thread();
    {
        std::lock_guard<std::mutex> lk(prog_mu);
        prog_stop = false;
    }
    prog_thread = std::thread([this] {
        uint64_t seen = solve_gen.load();
        std::unique_lock<std::mutex> lk(prog_mu);
        for (;;) {
            const bool stopping =
                prog_cv.wait_for(lk, std::chrono::milliseconds(prog_ms), [this] { return prog_stop; });
            const uint64_t g = solve_gen.load();
            if ((g & 1u) || g != seen) {
                lk.unlock();
                ssw_progress p;
                if (read_progress(&p) == SSW_OK) prog_fn(&p, prog_user);
                lk.lock();
            }
            seen = g;
            if (stopping) return;
        }
    });
}

void ssw_context::stop_progress_thread() {
    if (!prog_thread.joinable()) return;
    {
        std::lock_guard<std::mutex> lk(prog_mu);
        prog_stop = true;
    }
    prog_cv.notify_all();
    prog_thread.join();
}

// ---------------------------------------------------------------------------
// extern "C"
// ---------------------------------------------------------------------------

#define CTX_GUARD(ctx)                                           \
    if (!(ctx)) return SSW_ERR_ARG;                              \
    std::lock_guard<std::mutex> _lk((ctx)->mu)

extern "C" {

int ssw_abi_version(void) { return SSW_ABI_VERSION; }

int ssw_create(int device, uint64_t mem_budget, uint32_t flags, ssw_context **out) {
    (void)flags;
    if (!out) return SSW_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        g_create_error = "no HIP device";
        return SSW_ERR_DEVICE;
    }
    if (device < 0 || device >= ndev) {
        g_create_error = "device ordinal out of range";
        return SSW_ERR_ARG;
    }
    ssw_context *c = new (std::nothrow) ssw_context();
    if (!c) return SSW_ERR_NOMEM;
    int rc = c->open(device, mem_budget);
    if (rc) {
        g_create_error = c->err;
        c->close();
        delete c;
        return rc;
    }
    *out = c;
    return SSW_OK;
}

int ssw_destroy(ssw_context *ctx) {
    if (!ctx) return SSW_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(ctx->mu);
        ctx->close();
    }
    delete ctx;
    return SSW_OK;
}

const char *ssw_last_error(const ssw_context *ctx) {
    return ctx ? ctx->err.c_str() : g_create_error.c_str();
}

int ssw_compute(ssw_context *ctx, int32_t n, int64_t count, const int32_t *delta_a,
                const int32_t *delta_b, int32_t *length_out, uint8_t *sync_out, uint8_t *witness_out,
                int32_t witness_stride, int64_t *expanded_out, int32_t bound, int64_t *violation_out) {
    CTX_GUARD(ctx);
    if (count > 0 && !length_out) return ctx->fail(SSW_ERR_ARG, "null argument");
    return ctx->compute_i32(n, count, delta_a, delta_b, length_out, sync_out, witness_out, witness_stride,
                            expanded_out, bound, violation_out);
}

int ssw_verify(ssw_context *ctx, int32_t n, int64_t count, const int32_t *delta_a, const int32_t *delta_b,
               int32_t bound, int64_t *violation_out) {
    CTX_GUARD(ctx);
    return ctx->compute_i32(n, count, delta_a, delta_b, nullptr, nullptr, nullptr, 0, nullptr, bound, violation_out);
}

int ssw_best_parallelism(ssw_context *ctx) { return ctx ? 1 : SSW_ERR_ARG; }

int ssw_set_verify_prefilter(ssw_context *ctx, int32_t word_length) {
    CTX_GUARD(ctx);
    if (word_length != 0 && word_length != 6)
        return ctx->fail(SSW_ERR_ARG, "verify pre-filter word length %d (0 or 6)", word_length);
    ctx->verify_prefilter = word_length == 6;
    return SSW_OK;
}

int ssw_solve_batch(ssw_context *ctx, const ssw_batch *b, int32_t memory, int64_t *violation_out) {
    CTX_GUARD(ctx);
    return ctx->solve_entry(b, memory, kModeBatch, violation_out);
}

int ssw_solve_batch_noalloc(ssw_context *ctx, const ssw_batch *b, int64_t *violation_out) {
    CTX_GUARD(ctx);
    const int rc = ctx->solve_noalloc(b, violation_out);
    const int crc = ctx->checked_verdict();
    return rc ? rc : crc;
}

int ssw_solve_plain(ssw_context *ctx, const ssw_batch *b, int32_t memory, int64_t *violation_out) {
    CTX_GUARD(ctx);
    return ctx->solve_entry(b, memory, kModePlain, violation_out);
}

int ssw_solve_large(ssw_context *ctx, const ssw_batch *b, int32_t memory, int64_t *violation_out) {
    CTX_GUARD(ctx);
    return ctx->solve_entry(b, memory, kModeLarge, violation_out);
}

int ssw_get_tuning(ssw_context *ctx, int32_t n, ssw_tuning *out) {
    CTX_GUARD(ctx);
    if (n < 1 || n > SSW_MAX_N || !out) return ctx->fail(SSW_ERR_ARG, "bad n or null");
    *out = ctx->tune[n];
    return SSW_OK;
}

int ssw_set_tuning(ssw_context *ctx, int32_t n, const ssw_tuning *t) {
    CTX_GUARD(ctx);
    if (n < 1 || n > SSW_MAX_N || !t) return ctx->fail(SSW_ERR_ARG, "bad n or null");
    const int G = t->group_lanes;
    if (G != 16 && G != 32 && G != 64 && G != 256) return ctx->fail(SSW_ERR_ARG, "group_lanes %d", G);
    if (t->block_threads != 0 &&
        (t->block_threads % 64 || t->block_threads < 64 || t->block_threads > 256 ||
         (G != 256 && t->block_threads % G) || (G == 256 && t->block_threads != 256)))
        return ctx->fail(SSW_ERR_ARG, "block_threads %d", t->block_threads);
    if (t->chunk < 1 || t->arena_lds < 4) return ctx->fail(SSW_ERR_ARG, "chunk/arena_lds");
    if (t->lane_entries != 1 && t->lane_entries != 2 && t->lane_entries != 4)
        return ctx->fail(SSW_ERR_ARG, "lane_entries %d", t->lane_entries);
    if (t->spill_arena_lds < 0 || t->spill_arena_lds > 65536)
        return ctx->fail(SSW_ERR_ARG, "spill_arena_lds %d", t->spill_arena_lds);
    if (t->spill_lane_entries != 0 && t->spill_lane_entries != 1 && t->spill_lane_entries != 2 &&
        t->spill_lane_entries != 4 && t->spill_lane_entries != 8)
        return ctx->fail(SSW_ERR_ARG, "spill_lane_entries %d", t->spill_lane_entries);
    const int SG = t->spill_group_lanes;
    if (SG != 0 && SG != 16 && SG != 32 && SG != 64 && SG != 128 && SG != 256)
        return ctx->fail(SSW_ERR_ARG, "spill_group_lanes %d", SG);
    if (t->narrow_bits != 0 && (t->narrow_bits < 8 || t->narrow_bits > 16 || (n <= 16 && t->narrow_bits >= n) ||
                                (n > 16 && t->narrow_bits != 16)))
        return ctx->fail(SSW_ERR_ARG, "narrow_bits %d for n=%d (8..n-1 for n <= 16, 16 for n > 16)",
                         t->narrow_bits, n);
    ctx->tune[n] = *t;
    return SSW_OK;
}

int ssw_get_progress(ssw_context *ctx, ssw_progress *out) {
    if (!ctx || !out) return SSW_ERR_ARG;
    return ctx->read_progress(out);
}

int ssw_set_progress_callback(ssw_context *ctx, ssw_progress_fn fn, void *user, int32_t interval_ms) {
    if (!ctx) return SSW_ERR_ARG;
    ctx->stop_progress_thread();
    ctx->prog_fn = fn;
    ctx->prog_user = user;
    ctx->prog_ms = interval_ms > 0 ? interval_ms : 100;
    if (fn) ctx->start_progress_thread();
    return SSW_OK;
}

int ssw_device_alloc(ssw_context *ctx, uint64_t bytes, void **out) {
    if (!ctx || !out) return SSW_ERR_ARG;
    hipSetDevice(ctx->device);
    hipError_t e = hipMalloc(out, bytes ? bytes : 1);
    return e == hipSuccess ? SSW_OK : (e == hipErrorOutOfMemory ? SSW_ERR_NOMEM : SSW_ERR_DEVICE);
}

int ssw_device_free(ssw_context *ctx, void *p) {
    if (!ctx) return SSW_ERR_ARG;
    hipSetDevice(ctx->device);
    return hipFree(p) == hipSuccess ? SSW_OK : SSW_ERR_DEVICE;
}

int ssw_copy_to_device(ssw_context *ctx, void *dst, const void *src, uint64_t bytes) {
    CTX_GUARD(ctx);
    hipSetDevice(ctx->device);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? SSW_OK : ctx->hip_fail(e, "copy to device");
}

int ssw_copy_to_host(ssw_context *ctx, void *dst, const void *src, uint64_t bytes) {
    CTX_GUARD(ctx);
    hipSetDevice(ctx->device);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? SSW_OK : ctx->hip_fail(e, "copy to host");
}

int ssw_copy_on_device(ssw_context *ctx, void *dst, const void *src, uint64_t bytes) {
    CTX_GUARD(ctx);
    hipSetDevice(ctx->device);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? SSW_OK : ctx->hip_fail(e, "copy on device");
}

int ssw_synchronize(ssw_context *ctx) {
    CTX_GUARD(ctx);
    hipSetDevice(ctx->device);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? SSW_OK : ctx->hip_fail(e, "synchronize");
}

int ssw_last_launch_stats(ssw_context *ctx, ssw_launch_stat *out, int32_t cap, int32_t *count) {
    CTX_GUARD(ctx);
    if (!count || cap < 0 || (cap > 0 && !out)) return ctx->fail(SSW_ERR_ARG, "null argument");
    return ctx->launch_stats(out, cap, count);
}

int ssw_last_kernel_time(ssw_context *ctx, double *ms, int64_t *launches) {
    CTX_GUARD(ctx);
    if (ms) *ms = ctx->last_ms;
    if (launches) *launches = ctx->last_launches;
    return SSW_OK;
}

}  // extern "C"
