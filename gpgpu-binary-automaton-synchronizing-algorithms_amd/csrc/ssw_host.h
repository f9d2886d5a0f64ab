// ssw_host.h -- host-side state of a solver context (not part of the C ABI).
#pragma once

#include "ssw.h"
#include "ssw_internal.h"
#include "ssw_pool.h"

#include <algorithm>
#include <atomic>
#include <functional>
#include <condition_variable>
#include <cstdarg>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

namespace ssw {

constexpr int kMaxLdsN = 16;  // largest n whose visited bitmap (8 KiB) lives in LDS

enum SolveMode { kModeBatch = 0, kModePlain = 1, kModeLarge = 2 };

// layout of the small counters buffer
constexpr size_t kCtrWork = 0;         // u32 work-queue head
constexpr size_t kCtrDeferred = 4;     // u32 deferred count
constexpr size_t kCtrViolation = 8;    // u64 smallest violating index
// (progress counters: DevBuf prog, kCounterSlots copies)
constexpr size_t kCtrLarge = 64;       // device-wide engine scratch counters
constexpr size_t kCtrBytes = 256;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// A batch whose every pointer is a device pointer.
struct DeviceBatch {
    int64_t count = 0;
    int32_t uniform_n = 0;
    int32_t bound = -1;
    const uint8_t *n_per = nullptr;
    const uint64_t *offset = nullptr;
    const uint8_t *delta = nullptr;
    int32_t *length = nullptr;
    uint8_t *sync = nullptr;
    int64_t *expanded = nullptr;
    uint8_t *witness = nullptr;
    int32_t witness_stride = 0;
};

struct ClassSpec {
    int nmax = 0;
    bool lds = true;
    int64_t items = 0;
    const uint32_t *order = nullptr;
    const uint8_t *host_n = nullptr;  // plain path: per-automaton n on the host
    int plain_n = 0;
    hipStream_t stream = nullptr;     // null: the context's stream
    uint32_t *queues = nullptr;       // this class's device work queue
    uint32_t *spill_queue[3] = {nullptr, nullptr, nullptr};  // work queues and lists of the later
    uint32_t *spill_list[3] = {nullptr, nullptr, nullptr};   // launches of a spill chain (null: none)
    // allocation-free chain: the HBM rings of the last launch's groups come
    // from this caller-provided memory and every earlier launch spills
    uint8_t *static_slots = nullptr;
    uint64_t static_slot_bytes = 0;
    // one part of a larger batch: the spill chain stops before its last
    // launch, which the caller runs once for all parts (solve_last_launch)
    // over spill_queue[1] / spill_list[1]; the parts share that queue and
    // list (set up by the caller) and list their automata as the part's
    // indices plus list1_base
    bool defer_last = false;
    uint32_t list1_base = 0;
};

// launch passes of one class
enum PassKind { kPassAll = 0, kPassSpillFrom = 1, kPassSpillTo = 2 };

// What one launch of a class reads and writes besides the batch.
struct PassIO {
    int kind = kPassAll;
    const uint32_t *in_list = nullptr;  // work items: the class order, or this list ...
    uint32_t *in_queue = nullptr;       // ... with this queue, filled by the launch before
    uint32_t *out_list = nullptr;       // automata handed to the next launch
    uint32_t *out_queue = nullptr;
    uint32_t *skip_list = nullptr;      // ... or to the one after it (permutation letters)
    uint32_t *skip_queue = nullptr;
    int narrow_bits = 0;                // relabelled states, 2^narrow_bits-bit bitmap
    bool byte_tab = false;
    bool force_small = false;           // the n <= 16 solver on a class of larger n (relabelled)
    bool fam = false;                   // ... with the two-family encoding (n > 16)
    int general_entries = 1;            // frontier entries per lane of the general solver (1 or 4)
    bool general_fam = false;           // ... with the two-family encoding (n > 16, 4 entries per lane)
    bool queue_ready = false;           // the class's queue was set up with the spill queues
    uint32_t skip_base = 0;             // added to the indices written to the skip list ...
    uint32_t out_base = 0;              // ... and to the out list
};
constexpr int kClassQueues = 3;  // work queues per launch class: its own and two spill queues

// One launch of the last batched solve, for ssw_last_launch_stats: its shape,
// the HIP events recorded around it on its stream and its counters' slot.
constexpr int kMaxLaunchRecords = 32;
struct LaunchRecord {
    int n = 0, variant = 0, G = 0, block = 0, grid = 0, lds = 0, ring = 0, T = 0, bitmap_bits = 0, pass = 0;
    bool byte_tab = false, witness = false;
};

struct SlotKey {
    int G, block, lds_n, nmax, witness, arena_lds, plain;
    bool operator<(const SlotKey &o) const {
        return std::tie(G, block, lds_n, nmax, witness, arena_lds, plain) <
               std::tie(o.G, o.block, o.lds_n, o.nmax, o.witness, o.arena_lds, o.plain);
    }
};

}  // namespace ssw

struct ssw_context;

namespace ssw {

// Host mirror of one device-wide instance slot.
struct LargeSlot {
    bool used = false;
    bool harvested = false;
    bool append = false;
    int64_t index = -1;
    int n = 0;
    uint32_t *ring = nullptr;
    uint64_t ring_cap = 0;
    uint32_t *bitmap = nullptr;
    uint64_t *table = nullptr;
    uint64_t *lvl = nullptr;
    uint32_t lvl_cap = 0;
    uint64_t bytes = 0;
};

// Device-wide level loop for large instances (ssw_large.cpp).
struct LargeEngine {
    DevBuf state;                        // LargeState + LargeInstance[kLargeMaxInstances]
    std::vector<LargeSlot> slots;
    std::vector<LargeInstance> hinst;    // host copy of the instance array
    LargeState hstate;                   // staging for the output pointers
    int num = 0;                         // slots in use on the device: 0..num-1
    int grid = 1024;
    int tick_levels = 32;
    uint64_t bytes_in_use = 0;
    std::map<int, hipGraphExec_t> graphs;  // captured [step, plan] chains by level count
    // Device buffers of retired instances (bitmaps, rings, tables, level
    // tables), kept by size for the next admissions: allocating costs about
    // 0.4 ms per 128 MiB, a quarter of the C_24..C_30 set's wall time when
    // every admission allocated.  At most spare_limit bytes; freed with the
    // engine, or when an allocation fails.
    std::multimap<uint64_t, void *> spare;
    uint64_t spare_bytes = 0, spare_limit = 0;
    hipError_t take(void **p, uint64_t bytes);
    void give(void *p, uint64_t bytes);
    void drop_spare();

    LargeState *dstate() const;
    LargeInstance *dinst() const;
    void release();
    void free_slot_buffers(LargeSlot &s);
    int setup(ssw_context &ctx);
    int set_outputs(ssw_context &ctx, int32_t *length, uint8_t *sync, int64_t *expanded,
                    unsigned long long *violation, unsigned long long *progress);
    static uint64_t first_ring_cap(int n, bool append);
    static uint64_t admit_bytes(int n, bool append);
    // Admits one automaton; *slot_out = -1 when no slot or memory is free now.
    int admit(ssw_context &ctx, int64_t index, int n, const uint8_t *delta, const uint64_t *offset,
              int32_t bound, bool report, bool append, uint64_t budget, int *slot_out);
    int build_graph(ssw_context &ctx, int levels);
    int tick(ssw_context &ctx, int levels);
    int read_back(ssw_context &ctx);
    int running() const;
    int witness(ssw_context &ctx, int slot, uint8_t *out, int stride);
    void retire(int slot);
    int clear_slot(ssw_context &ctx, int slot);
    // Solves the listed automata of db (indices into db) concurrently.
    int solve_indices(ssw_context &ctx, const DeviceBatch &db, const uint8_t *host_n,
                      const std::vector<uint32_t> &idx);
    int solve_all(ssw_context &ctx, const DeviceBatch &db, const uint8_t *host_n, int64_t *violation_out);
};

}  // namespace ssw

struct ssw_context {
    int device = -1;
    int cus = 0;
    size_t lds_per_cu = 0;
    size_t lds_per_block = 65536;
    int last_grid = 0, last_block = 0, last_variant = 0;
    size_t last_lds = 0;
    hipStream_t stream = nullptr;
    hipStream_t monitor = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    static constexpr int kClassStreams = 8;        // streams for the launch classes of one call
    static constexpr int kConcurrentClasses = 4;   // of which one call uses at most this many
    hipStream_t class_streams[kClassStreams] = {};
    hipEvent_t ev_fork = nullptr;
    hipEvent_t ev_join[kClassStreams] = {};
    uint64_t budget = 0;
    std::mutex mu;           // serializes calls on the context
    std::mutex monitor_mu;   // progress reads
    std::string err;

    ssw::DevBuf ctrs, queues, deferred, order, spill;
    ssw::DevBuf sizes;  // counts of a device batch's sizes (32 u64)
    // per-launch statistics of the last batched solve: kMaxLaunchRecords
    // triples of u64 counters (zeroed with the solve's counters), the events
    // around each launch and its shape
    ssw::DevBuf lstats;  // launch statistics: kMaxLaunchRecords x kCounterSlots copies of kStatWords
    ssw::DevBuf prog;    // progress counters: kCounterSlots copies of [0] done, [1] subsets, [2] done this solve
    static constexpr uint32_t kLaunchStatWords = ssw::kMaxLaunchRecords * ssw::kCounterSlots * ssw::kStatWords;
    hipEvent_t lev[2 * ssw::kMaxLaunchRecords] = {};
    std::vector<ssw::LaunchRecord> records;
    ssw::DevBuf d_delta, d_nper, d_off, d_len, d_sync, d_exp, d_wit;
    // the second buffer set of the host pipeline (chunk k uses set k & 1)
    ssw::DevBuf d_delta2, d_len2, d_sync2, d_exp2;
    // with the last launch run once for all chunks: its work queue and list,
    // and the records of the listed automata's results (gather_listed)
    ssw::DevBuf pipe_q, pipe_list, pipe_patch;
    // its streams' events: inputs of set s copied (0, 1), set s's kernels
    // done with them (2, 3), outputs of set s copied back (4, 5); with the
    // last launch run once: chunk k's results copied back (4 + k, k < 4)
    hipEvent_t pipe_ev[8] = {};
    std::map<ssw::SlotKey, ssw::DevBuf> slot_cache;
    std::map<ssw::SlotKey, uint64_t> slot_gen;   // call that last used each entry
    uint64_t call_gen = 0;
    ssw::LargeEngine large;
    ssw_tuning tune[SSW_MAX_N + 1];

    double last_ms = 0;
    int64_t last_launches = 0;
    int64_t launches = 0;
    uint64_t deferred_total = 0;
    std::atomic<uint64_t> total_launches{0};
    std::atomic<uint64_t> large_done{0};
    std::atomic<uint64_t> large_subsets{0};
    // solve generation: bumped when a solve starts and again when it ends
    // (odd while one runs), so the progress thread sees every solve, however
    // short; call_count = automata of the running solve
    std::atomic<uint64_t> solve_gen{0};
    std::atomic<uint64_t> call_count{0};
    // automata of the running call already handed back by solves before the
    // current one (a pipelined call runs several solves)
    std::atomic<uint64_t> call_done_base{0};
    void solve_begin(uint64_t count) {
        call_count.store(count);
        call_done_base.store(0);
        solve_gen.fetch_add(1);
    }
    void solve_end() { solve_gen.fetch_add(1); }

    // progress callback
    std::thread prog_thread;
    std::mutex prog_mu;
    std::condition_variable prog_cv;
    bool prog_stop = false;  // guarded by prog_mu
    ssw_progress_fn prog_fn = nullptr;
    void *prog_user = nullptr;
    int prog_ms = 100;

    int open(int dev, uint64_t mem_budget);
    void close();
    void default_tuning();
    int fail(int code, const char *fmt, ...);
    int hip_fail(hipError_t e, const char *what);
    int ensure(ssw::DevBuf &b, size_t bytes, const char *what);
    int get_slots(const ssw::SlotKey &key, uint64_t bytes, ssw::DevBuf **out, hipStream_t st);
    int launch_class(const ssw::DeviceBatch &db, const ssw::ClassSpec &cls, bool plain);
    int launch_class_static(const ssw::DeviceBatch &db, const ssw::ClassSpec &cls);
    int launch_pass(const ssw::DeviceBatch &db, const ssw::ClassSpec &cls, const ssw_tuning &t, bool plain,
                    const ssw::PassIO &io);
    // host_n: the sizes in host memory, or null for a uniform batch or when
    // size_hist (counts per n) is given; then the sizes are copied from the
    // device only where a route needs them one by one
    // head_out (pinned, 16 bytes): the solve's counters are copied there
    // behind its launches and the call returns without waiting (a uniform
    // batch of n <= 16, whose automata are never deferred; the caller reads
    // the violation once the stream is drained)
    int solve_device(const ssw::DeviceBatch &db, const uint8_t *host_n, bool plain, int64_t *violation_out,
                     const int64_t *size_hist = nullptr, void *head_out = nullptr,
                     const ssw::ClassSpec *defer = nullptr);
    // the last launch of the parts of a uniform batch of n <= 16 whose
    // chains stopped before it (ClassSpec::defer_last), over their shared
    // list, on the whole batch; its counters copied to head_out
    int solve_last_launch(const ssw::DeviceBatch &db, uint32_t *queue1, uint32_t *list1, int64_t items,
                          void *head_out);
    int solve_host_packed(const ssw_batch *b, int mode, int64_t *violation_out);
    // a uniform batch of n <= 16 from host memory in pipelined chunks (fill
    // packs automata [lo, hi) into the pinned staging and returns the first
    // invalid one, or -1; bad_at reports it)
    int host_pipeline(int32_t n, int64_t count, int64_t nchunks, const std::function<int64_t(int64_t, int64_t)> &fill,
                      const std::function<int(int64_t)> &bad_at, int32_t *length_out, uint8_t *sync_out,
                      int64_t *expanded_out, int32_t bound, int64_t *violation_out);
    // host worker threads of the host entry points
    ssw::HostPool pool;
    // fn over [0, count) in ranges of at least per_thread on the pool's
    // threads (or on this one when the pool is busy or the count small)
    void parallel_ranges(int64_t count, int64_t per_thread, const std::function<void(int64_t, int64_t)> &fn);
    // host staging for the int32 entry points: pinned, grown on demand
    void *h_stage = nullptr;
    size_t h_stage_bytes = 0;
    int ensure_stage(size_t bytes);
    // pinned results of the host pipeline (two chunks' results and the solves' counters)
    void *h_out = nullptr;
    size_t h_out_bytes = 0;
    int ensure_out(size_t bytes);
    // set by compute_i32 for the call it makes: transitions already checked,
    // lengths not wanted by the caller (Verify)
    bool host_checked = false;
    bool host_no_length = false;
    // ssw_verify tries the 64 words of length 6 before each search
    bool verify_prefilter = false;
    int solve_entry(const ssw_batch *b, int32_t memory, int mode, int64_t *violation_out);
    // the allocation-free entry point: device buffers, uniform n <= 16, no
    // witness; every device buffer it uses is the context's (set up at
    // creation) or this GPU's static scratch
    int solve_noalloc(const ssw_batch *b, int64_t *violation_out);
    // checked build: SSW_ERR_INTERNAL if a kernel formed an index or LDS
    // address outside its buffer since the last call (SSW_OK otherwise)
    int checked_verdict();
    void *noalloc_base = nullptr;       // device address of this GPU's scratch
    hipError_t noalloc_status = hipSuccess;
    int solve_run(const ssw_batch *b, int32_t memory, int mode, int64_t *violation_out);
    int compute_i32(int32_t n, int64_t count, const int32_t *A, const int32_t *B, int32_t *length_out,
                    uint8_t *sync_out, uint8_t *witness_out, int32_t witness_stride, int64_t *expanded_out,
                    int32_t bound, int64_t *violation_out);
    int read_progress(ssw_progress *out);
    int launch_stats(ssw_launch_stat *out, int32_t cap, int32_t *count);
    void start_progress_thread();
    void stop_progress_thread();
    int measure_atomics(int32_t kind, double *ops_per_s);
    std::map<std::tuple<int, int, int, size_t>, int> occ_cache;
    hipError_t occupancy(int variant, int G, int block, size_t lds, int *out);
};
