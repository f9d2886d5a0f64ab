// ssw_internal.h -- structures shared by the HIP kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ssw {

// Nibble-sliced image tables: slice k holds, for every 4-bit pattern v of the
// states 4k..4k+3, the OR of their singleton images under a (low 32 bits) and
// b (high 32 bits).  8 slices cover n <= 32.
constexpr int kSliceBits = 4;
constexpr int kTableEntries = 8 * 16;
constexpr int kTableBytes = kTableEntries * 8;
// The general solver's two-family encoding (n > 16, visited bitmap in LDS):
// a family bit and a mask of at most kFamBits relabelled states; its tables
// hold slices 0..3 per family and one shared slice for bits 16..19.
constexpr int kFamBits = 18;
constexpr int kFamTableBytes = (2 * 4 * 16 + 16) * 8;
constexpr uint32_t kFamBitmapWords = (1u << (kFamBits + 1)) / 32;

// Device work queue: kQueues segments of kQueueWords u32 each (next, end),
// then a u64 mask of exhausted segments.
constexpr int kQueues = 64;
constexpr int kQueueWords = 32;
constexpr size_t kQueueBytes = kQueues * kQueueWords * 4 + 64;

// Scratch of the allocation-free entry point (ssw_solve_batch_noalloc): one
// file-scope device array per GPU (ssw_batch.hip) holding, for one uniform
// batch of n <= 16 with at most kNoAllocMaxCount automata, the work queues of
// its launch chain, three spill lists and the HBM rings of the last launch's
// groups (every earlier launch hands a frontier that outgrows its LDS ring
// to the next launch instead of moving it to HBM).
constexpr uint64_t kNoAllocMaxCount = 1ull << 20;
constexpr int kNoAllocMaxN = 16;
constexpr uint64_t kNoAllocQueueOff = 0;                  // 4 work queues
constexpr uint64_t kNoAllocListOff = 64ull << 10;         // 3 spill lists of kNoAllocMaxCount u32
constexpr uint64_t kNoAllocSlotOff = kNoAllocListOff + 3ull * 4ull * kNoAllocMaxCount;
constexpr uint64_t kNoAllocSlotBytes = 2ull * 65536;      // one 65536-entry u16 ring per group
constexpr uint32_t kNoAllocSlots = 256;                   // groups of the last launch at most
constexpr uint64_t kNoAllocScratchBytes = kNoAllocSlotOff + kNoAllocSlots * kNoAllocSlotBytes;
constexpr uint32_t kNoAllocLastRing = 16384;              // LDS ring entries of the last launch

// Result codes of one BFS level step.
enum StepCode : int { kContinue = 0, kFound = 1, kNoSync = 2, kOverflow = 3, kSpilled = 4, kPrefiltered = 5 };

// LDS layout of the n <= 16 path, per workgroup: the visited bitmaps of all
// groups first, each aligned to its size (a power of two, at least 256
// bytes), then one region per group -- image table (four nibble slices,
// 256 bytes, or two byte slices, 2 KiB), transitions staging (64 bytes),
// singleton images of the states (64 bytes), control words of groups of
// several wavefronts (64 bytes), and the frontier ring of ring_cap u16
// entries plus 64 dummy entries -- then the group scratch.
struct SmallLds {
    static __host__ __device__ inline uint32_t bitmap_bytes(int n) {
        const uint32_t b = n >= 3 ? (1u << (n - 3)) : 1u;
        return b < 256u ? 256u : b;
    }
    // byte tables; nibble tables (two families: slices 0..2 twice)
    static __host__ __device__ inline uint32_t tab_bytes(bool byte_tab, bool fam = false) {
        return byte_tab ? 2048u : (fam ? 512u : 256u);
    }
    static __host__ __device__ inline uint32_t region_bytes(bool byte_tab, uint32_t ring_cap, bool fam = false) {
        // tables, staging, singleton images, control words, spare, ring (+64 dummy slots)
        const uint32_t b = tab_bytes(byte_tab, fam) + 64u + 64u + 64u + 64u + 2u * (ring_cap + 64u);
        const uint32_t a = byte_tab ? 1024u : 256u;
        return (b + a - 1u) / a * a;
    }
    static __host__ __device__ inline uint32_t regions_base(uint32_t groups, uint32_t bm_bytes) {
        return (groups * bm_bytes + 1023u) & ~1023u;
    }
    static __host__ __device__ inline uint32_t block_bytes(uint32_t groups, uint32_t bm_bytes, uint32_t region,
                                                           uint32_t scratch) {
        return regions_base(groups, bm_bytes) + groups * region + scratch;
    }
};

// Counters the groups add their totals to when they flush (progress, and a
// launch's statistics): kCounterSlots copies on lines of their own, picked by
// workgroup index, which readers sum.  With one copy, the groups of a launch
// finishing together serialised on one L2 line (0.3 ms of the 1M n = 16
// solve).
constexpr uint32_t kCounterSlots = 64;
constexpr uint32_t kSlotWords = 32;     // u64 words between progress copies (256 bytes)
constexpr uint32_t kStatWords = 4;      // u64 words per launch-statistics copy
constexpr uint32_t kProgressBytes = kCounterSlots * kSlotWords * 8;

// Arguments of the batched persistent kernel (and of the plain path, which is
// the same kernel run on one automaton per launch by one 256-thread group).
struct BatchParams {
    int64_t count;
    int32_t uniform_n;
    int32_t bound;                 // < 0: (n-1)^2 per automaton
    const uint8_t *n_per;          // [count] or null
    const uint64_t *offset;        // [count] or null
    const uint8_t *delta;          // packed a[0..n) b[0..n) bytes
    const uint32_t *order;         // work item -> automaton index, or null
    int32_t *length;
    uint8_t *sync;
    int64_t *expanded;
    uint8_t *witness;
    int32_t witness_stride;
    uint32_t want_witness;
    unsigned long long *violation;     // atomicMin of violating indices
    uint32_t *queues;                  // kQueues segment counters + exhausted mask
    unsigned int *deferred_count;
    uint32_t *deferred_list;           // automata whose frontier outgrew the arena
    unsigned long long *progress;      // kCounterSlots copies (kSlotWords apart) of [0] automata done,
                                       // [1] subsets expanded, [2] done in this solve
    unsigned long long *launch_stats;  // null, or this launch's kCounterSlots copies of [0] automata finished, [1] their E
                                       // summed, [2] subsets expanded (with partial searches handed on)
    uint8_t *slot_base;                // per-group global scratch
    uint64_t slot_bytes;
    uint64_t garena_off;               // u32 frontier arena beyond the LDS part
    uint64_t hist_off;                 // n <= 16 with witness words: subsets in discovery order
    uint64_t himg_off;                 // ... and (4 entries per lane) their images, packed as the tables'
    uint64_t garena_cap;               // entries
    uint64_t gbitmap_off;              // visited bitmap for n > lds_n
    uint64_t gbitmap_words;            // its size in u32 words (0: none)
    uint64_t good_off;                 // two witness bitmaps of good_words words
    uint64_t good_words;
    uint64_t lvl_off;                  // level start offsets, lvl_cap entries
    uint32_t lvl_cap;
    int32_t lds_n;                     // the LDS bitmap serves n <= lds_n
    uint32_t lds_bitmap_words;         // per group, multiple of 4
    uint32_t lds_arena_cap;            // per group: u32 arena entries, or u16 ring entries (power of 2)
    uint32_t lds_group_bytes;          // LDS bytes of one group
    uint32_t chunk;                    // automata per work-queue fetch
    uint32_t single_item;              // plain path: solve only first_item
    int64_t first_item;
    // Ring spill (n <= 16 path): with spill = 1 an automaton whose frontier
    // outgrows the LDS ring is not moved to HBM; it is handed to the next
    // launch of the class.  spill_queue is that launch's work queue, set up
    // with empty segments at the same boundaries as this launch's queue: an
    // automaton taken from segment k is appended to segment k (atomicAdd on
    // its end word, so appends spread over kQueues addresses) and its index
    // stored in spill_list at that position.
    uint32_t spill;
    uint32_t *spill_queue;
    uint32_t *spill_list;
    // the launch after the next one (the large ring), or null: automata
    // spilled at start whose letter a or b permutes the states (a(Q) or b(Q)
    // is Q, as in every Cerny automaton) are handed there directly, since
    // their searches are the long ones and the next launch would only replay
    // their first levels
    uint32_t *skip_queue;
    uint32_t *skip_list;
    // added to the automaton indices written to the skip list and to the
    // spill list (lists shared by the parts of a larger batch)
    uint32_t skip_base;
    uint32_t spill_base;
    // n <= 16 path: > 0 relabels every automaton's states so that
    // U = a(Q) | b(Q) takes the lowest labels and uses a visited bitmap of
    // 2^narrow_bits bits; automata with |U| > narrow_bits are spilled at start
    int32_t narrow_bits;
    // Verify mode: before its search, each automaton tries the 64 words of
    // length 6 from Q, one per lane (one wavefront per automaton); one that
    // synchronizes proves L <= 6, within any bound >= 6, and the search is
    // skipped (its length, flag and expanded count are not written)
    uint32_t prefilter;
};

// Device-wide path: every instance keeps its visited bitmap (2^n bits) and a
// frontier ring in HBM; one launch expands the current level of all running
// instances.
constexpr int kLargeTile = 1024;          // frontier entries per workgroup tile
constexpr int kLargeMaxInstances = 256;   // concurrent instances per engine

enum LargeStatus : int32_t {
    kLIdle = 0,      // free slot
    kLRun = 1,
    kLPause = 2,     // ring (or level table) too small for the next level: host grows it
    kLFound = 3,
    kLNoSync = 4,
    kLError = 5,     // invalid transitions or ring overflow
};

struct LargeInstance {
    // set by the host
    uint32_t *ring;          // frontier ring of u32 subsets, capacity ring_mask + 1
    uint32_t *bitmap;        // 2^n bits
    uint64_t *table;         // 4 byte slices x 256 entries: a | b << 32
    uint64_t *lvl;           // append mode: start position of every level
    uint64_t ring_mask;
    int64_t index;           // automaton index in the caller's batch
    int32_t n;
    int32_t bound;
    uint32_t lvl_cap;
    uint32_t append;         // 1: positions never wrap (the witness needs every level)
    // device state
    uint64_t lo, hi;         // level being expanded: positions [lo, hi)
    unsigned long long wr;   // next free position
    int32_t level;
    int32_t status;
    int32_t found;           // a singleton image was met in this level
    int32_t overflow;
    int32_t length;
    int32_t report;          // 1: write the result into the batch outputs
    uint64_t expanded;
    uint64_t pad[2];
};

struct LargeState {
    int32_t num;             // slots in use (0..num-1)
    int32_t active;          // running instances with work this level
    uint64_t total_tiles;
    int32_t act[kLargeMaxInstances];
    uint64_t act_base[kLargeMaxInstances + 1];
    // batch outputs (may be null)
    int32_t *length;
    uint8_t *sync;
    int64_t *expanded;
    unsigned long long *violation;
    unsigned long long *progress;
};

}  // namespace ssw
