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

// Result codes of one BFS level step.
enum StepCode : int { kContinue = 0, kFound = 1, kNoSync = 2, kOverflow = 3 };

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
    unsigned int *work_counter;        // device work queue head
    unsigned int *deferred_count;
    uint32_t *deferred_list;           // automata whose frontier outgrew the arena
    unsigned long long *progress;      // [0] automata done, [1] subsets expanded
    uint8_t *slot_base;                // per-group global scratch
    uint64_t slot_bytes;
    uint64_t garena_off;               // u32 frontier arena beyond the LDS part
    uint64_t garena_cap;               // entries
    uint64_t gbitmap_off;              // visited bitmap for n > lds_n
    uint64_t gbitmap_words;            // its size in u32 words (0: none)
    uint64_t good_off;                 // two witness bitmaps of good_words words
    uint64_t good_words;
    uint64_t lvl_off;                  // level start offsets, lvl_cap entries
    uint32_t lvl_cap;
    int32_t lds_n;                     // the LDS bitmap serves n <= lds_n
    uint32_t lds_bitmap_words;         // per group, multiple of 4
    uint32_t lds_arena_cap;            // per group, multiple of 4
    uint32_t chunk;                    // automata per work-queue fetch
    uint32_t single_item;              // plain path: solve only first_item
    int64_t first_item;
};

// Device-wide path: one instance per automaton, all in HBM.
struct LargeInstance {
    int32_t n;
    int32_t status;          // 0 running, 1 found, 2 not synchronizing, 3 overflow
    int32_t level;           // level being expanded
    int32_t length;
    uint64_t lo, hi;         // current level [lo, hi) in ring positions
    uint64_t wr;             // write position (atomic during a step)
    uint64_t ring_mask;      // ring capacity - 1 (power of two)
    uint32_t *ring;          // frontier ring
    uint32_t *bitmap;        // 2^n bits
    uint64_t *table;         // 4 byte slices x 256 entries (a | b << 32)
    int64_t index;           // automaton index in the caller's batch
    int32_t found_flag;      // set by any thread that meets a singleton
    int32_t bound;
    uint64_t tile_base;      // first tile of this instance in the flattened grid
    uint64_t expanded;
};

struct LargeState {
    int32_t num_instances;
    int32_t active;          // instances still running
    uint32_t blocks_done;    // last-block-done counter of a step
    uint32_t step_count;
    uint64_t total_tiles;
    unsigned long long steps_executed;
};

}  // namespace ssw
