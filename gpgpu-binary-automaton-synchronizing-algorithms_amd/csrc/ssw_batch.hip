// ssw_batch.hip -- batched persistent BFS over subsets (gfx950).
//
// Replaces the level loop of SlimGPUQueue.Kernel (GPGPU/GPGPU/GPU/SlimGPUQueue.cs:210-407)
// and follows the search of SlimCPUUntweaked.ComputeMany
// (GPGPU/GPGPU/CPU/SlimCPUUntweaked.cs:89-216): BFS from Q over subsets, images
// under a then b, first-discovery wins, stop at the first singleton.
//
// Design (one automaton per thread group, many groups per CU):
//   * a group is 16, 32 or 64 lanes of one wavefront (no s_barrier, wave
//     ballots only) or a whole 256-thread workgroup (s_barrier + LDS);
//   * subset images come from nibble-sliced tables in LDS, 16 u64 entries
//     per 4 states (a in the low word, b in the high word);
//   * the visited set is a bitmap of 2^n bits, in LDS when n <= lds_n and in
//     the group's HBM slot otherwise, tested and set with atomicOr (the
//     winner is the lane that sees the bit clear);
//   * the frontier is one append-only arena per group (the FIFO of the CPU
//     search, processed level by level): its first lds_arena_cap entries in
//     LDS, the rest in the group's HBM slot; newly discovered subsets are
//     compacted into it with wavefront ballots and popcount prefix sums;
//   * groups take automata from a device work queue (one atomicAdd per chunk)
//     until the batch is exhausted, so long-tail automata (Cerny) do not hold
//     up the rest of the batch;
//   * the witness (lexicographically smallest shortest reset word) is read
//     off the retained levels: a backward pass marks the subsets lying on a
//     shortest path to a singleton, a forward pass takes letter a whenever
//     its image is such a subset of the next level.
#include "ssw_internal.h"
#include "ssw_checked.h"

#include <mutex>

namespace ssw {

// Explicit address spaces: LDS (3) and global (1).  With generic pointers the
// optimizer merges the LDS and HBM branches of the bitmap test-and-set into
// one flat atomic.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
typedef __attribute__((address_space(3))) uint64_t lds_u64;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(3))) u32x4 lds_u4;
typedef __attribute__((address_space(3))) uint16_t lds_u16;
typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef __attribute__((address_space(1))) uint32_t glb_u32;
typedef __attribute__((address_space(1))) uint16_t glb_u16;
typedef __attribute__((address_space(1))) uint64_t glb_u64;
typedef __attribute__((address_space(1))) uint8_t glb_u8;

template <class T>
__device__ __forceinline__ uint32_t or_wg(T *p, uint32_t v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ bool is_single(uint32_t x) { return (x & (x - 1u)) == 0u; }

// Marks a value as uniform across the wavefront (valid when a thread group is
// a whole wavefront or workgroup), so the compiler keeps it in an SGPR and
// branches on it with scalar branches instead of exec masks.
template <int G>
__device__ __forceinline__ uint32_t uni(uint32_t v) {
    return G >= 64 ? __builtin_amdgcn_readfirstlane(v) : v;
}
template <int G>
__device__ __forceinline__ int32_t uni(int32_t v) {
    return G >= 64 ? (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)v) : v;
}
template <int G>
__device__ __forceinline__ int64_t uni(int64_t v) {
    if (G < 64) return v;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Inclusive OR-scan within each row of 16 lanes (DPP row_shr 1, 2, 4, 8;
// lanes shifted in from outside the row read 0): lane 16k+15 ends with the
// OR of its row.
__device__ __forceinline__ uint32_t row_or_scan(uint32_t m) {
    m |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)m, 0x111, 0xF, 0xF, true);
    m |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)m, 0x112, 0xF, 0xF, true);
    m |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)m, 0x114, 0xF, 0xF, true);
    m |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)m, 0x118, 0xF, 0xF, true);
    return m;
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// G lanes of one wavefront (G = 16, 32, 64).
template <int G>
struct WaveGroup {
    static constexpr int kSize = G;
    static constexpr int kScratchBytes = 0;
    uint32_t lane, rank_, gbase;
    uint64_t gmask, below;
    __device__ explicit WaveGroup(void *) {
        lane = __lane_id();
        rank_ = lane & (G - 1);
        gbase = lane & ~(uint32_t)(G - 1);
        gmask = (G == 64) ? ~0ull : (((1ull << (G & 63)) - 1ull) << gbase);
        below = ((1ull << lane) - 1ull) & gmask;
    }
    __device__ uint32_t rank() const { return rank_; }
    __device__ void sync() const { wave_sync(); }
    __device__ bool any(bool p) const { return (__ballot(p) & gmask) != 0; }
    __device__ void exchange(bool single, bool na, bool nb, bool &any_single, uint32_t &off,
                             uint32_t &total) const {
        const uint64_t bs = __ballot(single) & gmask;
        const uint64_t ba = __ballot(na) & gmask;
        const uint64_t bb = __ballot(nb) & gmask;
        any_single = bs != 0;
        off = __popcll(ba & below) + __popcll(bb & below);
        total = __popcll(ba) + __popcll(bb);
    }
    // any(single) and the exclusive prefix / total of per-lane counts
    // c < 2^BITS, from one ballot per bit of c
    template <int BITS>
    __device__ void exchange_count(bool single, uint32_t c, bool &any_single, uint32_t &prefix,
                                   uint32_t &total) const {
        any_single = (__ballot(single) & gmask) != 0;
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int j = 0; j < BITS; ++j) {
            const uint64_t bj = __ballot((c >> j) & 1u) & gmask;
            pre += (uint32_t)__popcll(bj & below) << j;
            tot += (uint32_t)__popcll(bj) << j;
        }
        prefix = pre;
        total = tot;
    }
    // any(single), and the exclusive prefix / total of the flags f[0..K) of
    // every lane counted lane-major (all flags of lower lanes come first):
    // one ballot per flag, the prefix from mbcnt
    template <int K>
    __device__ void exchange_flags(bool single, const bool (&f)[K], bool &any_single, uint32_t &prefix,
                                   uint32_t &total) const {
        any_single = (__ballot(single) & gmask) != 0;
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint64_t m = __ballot(f[j]) & gmask;
            pre = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, pre));
            tot += (uint32_t)__popcll(m);
        }
        prefix = pre;
        total = tot;
    }
    // any(single), and for every flag f[j] of every lane its position in a
    // flag-major layout starting at base: all set f[0] of the group in lane
    // order, then all set f[1], ...; total = number of set flags
    // From wavefront ballots: sm = lanes with a singleton image, f[j] = lanes
    // whose flag j is set.  any(single), and for every flag of every lane its
    // position in a flag-major layout starting at base: all set flags 0 of
    // the group in lane order, then all set flags 1, ...; total = number of
    // set flags
    template <int K>
    __device__ void place(uint64_t sm, const uint64_t (&f)[K], uint32_t base, bool &any_single, uint32_t (&pos)[K],
                          uint32_t &total) const {
        any_single = (sm & gmask) != 0;
        uint32_t b = base;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint64_t m = f[j] & gmask;
            pos[j] = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, b));
            b += (uint32_t)__popcll(m);
        }
        total = b - base;
    }
    // bit k set when work-queue segment k still has work: the group's lanes
    // read the kQueues segments in parallel
    __device__ uint64_t live_segments(const uint32_t *queues) const;
    __device__ uint32_t bcast0(uint32_t v) const {
        if (G == 64) return __builtin_amdgcn_readfirstlane(v);
        return __shfl(v, (int)gbase, 64);
    }
    // OR of v over ranks 0..31 (0..G-1 when G < 32), for every lane of the group
    __device__ uint32_t or_low32(uint32_t v) const {
        v = row_or_scan(v);
        if (G == 64) return __builtin_amdgcn_readlane(v, 15) | __builtin_amdgcn_readlane(v, 31);
        if (G == 32) return (uint32_t)__shfl((int)v, (int)gbase + 15, 64) | (uint32_t)__shfl((int)v, (int)gbase + 31, 64);
        return (uint32_t)__shfl((int)v, (int)gbase + 15, 64);
    }
};

// A whole workgroup of NT threads (one group per block).
template <int NT>
struct BlockGroup {
    static constexpr int kSize = NT;
    static constexpr int kWaves = NT / 64;
    static constexpr int kScratchBytes = 256;
    uint32_t *scratch;
    uint32_t tid, lane, wave;
    __device__ explicit BlockGroup(void *s) : scratch((uint32_t *)s) {
        tid = threadIdx.x;
        lane = __lane_id();
        wave = threadIdx.x >> 6;
    }
    __device__ uint32_t rank() const { return tid; }
    __device__ void sync() const { __syncthreads(); }
    __device__ bool any(bool p) const {
        const uint64_t b = __ballot(p);
        __syncthreads();
        if (lane == 0) scratch[wave] = b != 0;
        __syncthreads();
        bool r = false;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) r |= scratch[w] != 0;
        return r;
    }
    __device__ void exchange(bool single, bool na, bool nb, bool &any_single, uint32_t &off,
                             uint32_t &total) const {
        const uint64_t bs = __ballot(single);
        const uint64_t ba = __ballot(na);
        const uint64_t bb = __ballot(nb);
        const uint64_t below = (1ull << lane) - 1ull;
        __syncthreads();
        if (lane == 0) {
            scratch[wave] = __popcll(ba) + __popcll(bb);
            scratch[kWaves + wave] = bs != 0;
        }
        __syncthreads();
        uint32_t pre = 0, tot = 0;
        bool anys = false;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t c = scratch[w];
            pre += (w < (int)wave) ? c : 0u;
            tot += c;
            anys |= scratch[kWaves + w] != 0;
        }
        off = pre + __popcll(ba & below) + __popcll(bb & below);
        total = tot;
        any_single = anys;
    }
    template <int BITS>
    __device__ void exchange_count(bool single, uint32_t c, bool &any_single, uint32_t &prefix,
                                   uint32_t &total) const {
        const uint64_t bs = __ballot(single);
        const uint64_t below = (1ull << lane) - 1ull;
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int j = 0; j < BITS; ++j) {
            const uint64_t bj = __ballot((c >> j) & 1u);
            pre += (uint32_t)__popcll(bj & below) << j;
            tot += (uint32_t)__popcll(bj) << j;
        }
        __syncthreads();
        if (lane == 0) {
            scratch[wave] = tot;
            scratch[kWaves + wave] = bs != 0;
        }
        __syncthreads();
        uint32_t wpre = 0, all = 0;
        bool anys = false;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t x = scratch[w];
            wpre += (w < (int)wave) ? x : 0u;
            all += x;
            anys |= scratch[kWaves + w] != 0;
        }
        prefix = wpre + pre;
        total = all;
        any_single = anys;
    }
    template <int K>
    __device__ void exchange_flags(bool single, const bool (&f)[K], bool &any_single, uint32_t &prefix,
                                   uint32_t &total) const {
        const uint64_t bs = __ballot(single);
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint64_t m = __ballot(f[j]);
            pre = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, pre));
            tot += (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (lane == 0) {
            scratch[wave] = tot;
            scratch[kWaves + wave] = bs != 0;
        }
        __syncthreads();
        uint32_t wpre = 0, all = 0;
        bool anys = false;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const uint32_t x = scratch[w];
            wpre += (w < (int)wave) ? x : 0u;
            all += x;
            anys |= scratch[kWaves + w] != 0;
        }
        prefix = wpre + pre;
        total = __builtin_amdgcn_readfirstlane(all);
        any_single = __builtin_amdgcn_readfirstlane(anys ? 1u : 0u) != 0u;
    }
    template <int K>
    __device__ void place(uint64_t sm, const uint64_t (&f)[K], uint32_t base, bool &any_single, uint32_t (&pos)[K],
                          uint32_t &total) const {
        const uint64_t bs = sm;
        uint32_t pre[K], cnt[K];
#pragma unroll
        for (int j = 0; j < K; ++j) {
            const uint64_t m = f[j];
            pre[j] = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
            cnt[j] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < K; ++j) scratch[wave * K + j] = cnt[j];
            scratch[kWaves * K + wave] = bs != 0;
        }
        __syncthreads();
        bool anys = false;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) anys |= scratch[kWaves * K + w] != 0;
        uint32_t b = base;
#pragma unroll
        for (int j = 0; j < K; ++j) {
            uint32_t before = 0, tot = 0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                const uint32_t c = scratch[w * K + j];
                before += (w < (int)wave) ? c : 0u;
                tot += c;
            }
            pos[j] = b + before + pre[j];
            b += tot;
        }
        total = __builtin_amdgcn_readfirstlane(b - base);
        any_single = __builtin_amdgcn_readfirstlane(anys ? 1u : 0u) != 0u;
    }
    // wave 0 reads the kQueues segments in parallel
    __device__ uint64_t live_segments(const uint32_t *queues) const;
    __device__ uint32_t bcast0(uint32_t v) const {
        __syncthreads();
        if (tid == 0) scratch[2 * kWaves] = v;
        __syncthreads();
        return __builtin_amdgcn_readfirstlane(scratch[2 * kWaves]);
    }
    // OR of v over threads 0..31, for every thread
    __device__ uint32_t or_low32(uint32_t v) const {
        v = row_or_scan(v);
        return bcast0(__builtin_amdgcn_readlane(v, 15) | __builtin_amdgcn_readlane(v, 31));
    }
};

// Device work queue: the batch is cut into kQueues contiguous segments, each
// with its own counter on its own 128-byte line, so the atomics of thousands
// of groups spread over many addresses.  A group starts at segment
// (slot mod kQueues); when that runs dry it reads all segments' (next, end)
// pairs at once, one per lane (one round trip, distinct addresses), and moves
// to the nearest live one.  A failed claim means that segment is dry for
// good, so the loop ends after at most kQueues + 1 claims; once every
// segment is dry a group leaves after one probe.
template <class Grp>
__device__ int64_t fetch_work(const BatchParams &P, const Grp &g, uint32_t &q, int64_t &it, int64_t &it_end) {
    if (it < it_end) return it++;
    for (int tries = 0; tries <= kQueues; ++tries) {
        uint32_t base = 0xFFFFFFFFu, end = 0;
        if (g.rank() == 0) {
            const uint32_t b = atomicAdd(P.queues + q * kQueueWords, P.chunk);
            const uint32_t e = P.queues[q * kQueueWords + 1];
            if (b < e) {
                base = b;
                end = b + P.chunk < e ? b + P.chunk : e;
            }
        }
        base = g.bcast0(base);
        end = g.bcast0(end);
        if (base != 0xFFFFFFFFu) {
            it = base;
            it_end = end;
            return it++;
        }
        const uint64_t live = g.live_segments(P.queues);
        if (live == 0) return -1;
        // nearest live segment after q, cyclically
        const uint64_t above = q + 1 < 64 ? (live >> (q + 1)) << (q + 1) : 0ull;
        q = above ? (uint32_t)__ffsll((long long)above) - 1u : (uint32_t)__ffsll((long long)live) - 1u;
    }
    return -1;
}

// (next, end) of work-queue segment k, read at device scope
__device__ __forceinline__ bool segment_live(const uint32_t *queues, uint32_t k) {
    const uint32_t nx = __hip_atomic_load(queues + k * kQueueWords, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t e = __hip_atomic_load(queues + k * kQueueWords + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return nx < e;
}

template <int G>
__device__ uint64_t WaveGroup<G>::live_segments(const uint32_t *queues) const {
    uint64_t live = 0;
#pragma unroll
    for (int j = 0; j < kQueues / G; ++j) {
        const uint64_t b = __ballot(segment_live(queues, rank_ + j * G)) & gmask;
        live |= ((b >> gbase) & (G == 64 ? ~0ull : ((1ull << (G & 63)) - 1ull))) << (j * G);
    }
    return G == 64 ? (uint64_t)uni<64>((int64_t)live) : live;
}

template <int NT>
__device__ uint64_t BlockGroup<NT>::live_segments(const uint32_t *queues) const {
    uint64_t live = 0;
    if (wave == 0) live = __ballot(segment_live(queues, lane));
    __syncthreads();
    if (tid == 0) {
        scratch[48] = (uint32_t)live;
        scratch[49] = (uint32_t)(live >> 32);
    }
    __syncthreads();
    const uint32_t lo = __builtin_amdgcn_readfirstlane(scratch[48]);
    const uint32_t hi = __builtin_amdgcn_readfirstlane(scratch[49]);
    return ((uint64_t)hi << 32) | lo;
}

// Sets up the segments of a batch of `count` work items, or (empty = 1) the
// same segments with no items yet, for a launch whose items are appended by
// the one before it.
// Sets up to three work queues over `count` items in one launch: queue j
// (non-null) gets the 64 segment boundaries, full or (bit j of empty_mask)
// empty, and its exhausted mask.
__global__ void init_queues_kernel(uint32_t *q0, uint32_t *q1, uint32_t *q2, int64_t count, uint32_t empty_mask) {
    const int k = threadIdx.x;
    uint32_t *qs[3] = {q0, q1, q2};
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        uint32_t *queues = qs[j];
        if (!queues) continue;
        if (k < kQueues) {
            const uint32_t s = (uint32_t)(count * k / kQueues), e = (uint32_t)(count * (k + 1) / kQueues);
            queues[k * kQueueWords] = s;
            queues[k * kQueueWords + 1] = ((empty_mask >> j) & 1u) ? s : e;
        }
    }
    __syncthreads();
    if (k < 3 && qs[k]) {
        uint32_t *queues = qs[k];
        unsigned long long m = 0;
        for (int j = 0; j < kQueues; ++j)
            if (queues[j * kQueueWords] >= queues[j * kQueueWords + 1]) m |= 1ull << j;
        *(unsigned long long *)(queues + kQueues * kQueueWords) = m;
    }
}

// The per-solve counters: deferred count <- 0, smallest violating index <-
// none (either may be null), automata done in this solve <- 0 in every copy
// of the progress counters (unless null), the launch statistics <- 0.
__global__ void reset_counters_kernel(uint32_t *deferred, unsigned long long *violation,
                                      unsigned long long *progress, unsigned long long *launch_stats,
                                      uint32_t stat_words) {
    if (threadIdx.x == 0) {
        if (deferred) *deferred = 0u;
        if (violation) *violation = ~0ull;
    }
    if (progress)
        for (uint32_t k = threadIdx.x; k < kCounterSlots; k += blockDim.x) progress[(size_t)k * kSlotWords + 2] = 0ull;
    for (uint32_t k = threadIdx.x; k < stat_words; k += blockDim.x) launch_stats[k] = 0ull;
}

// The results of the automata listed in a drained work queue's list (entries
// [start of segment k, end of segment k) of a queue set up empty over
// `items` slots), as 16-byte records {index | length << 32, expanded | flag
// << 56} appended after out[0..1] (out[0] counts them, zeroed by the caller):
// the host pipeline copies these back after the last launch of a batch whose
// other results it copied earlier.  An index at or past `count` (none is) is
// recorded without reading the results, for the host to report.
__global__ void gather_listed_kernel(const uint32_t *queue, const uint32_t *list, uint32_t items, uint32_t count,
                                     const int32_t *length, const uint8_t *sync, const int64_t *expanded,
                                     unsigned long long *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool have = false;
    uint32_t a = 0;
    if (i < items) {
        // segment k holds slots [items * k / kQueues, items * (k + 1) / kQueues)
        uint32_t k = (uint32_t)(((uint64_t)i * kQueues) / items);
        while (k + 1u < (uint32_t)kQueues && (uint64_t)items * (k + 1u) / kQueues <= i) ++k;
        while (k > 0u && (uint64_t)items * k / kQueues > i) --k;
        if (i < queue[k * kQueueWords + 1]) {
            have = true;
            a = list[i];
        }
    }
    const uint64_t m = __ballot(have);
    if (m == 0) return;
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    uint32_t base = 0;
    if (have && below == 0u) base = atomicAdd((unsigned int *)out, (unsigned int)__popcll(m));
    base = __builtin_amdgcn_readlane(base, (int)__builtin_ctzll(m));
    if (have) {
        const uint64_t pos = 2u + 2ull * (base + below);
        const bool in = a < count;
        const uint64_t e = in && expanded ? (uint64_t)expanded[a] : 0ull;
        const uint64_t f = in && sync ? (uint64_t)sync[a] : 0ull;
        out[pos] = (uint64_t)a | ((uint64_t)(uint32_t)(in ? length[a] : 0) << 32);
        out[pos + 1] = (e & 0x00FFFFFFFFFFFFFFull) | (f << 56);
    }
}

// The general solver (any n <= 30, HBM bitmaps past the LDS size, witness
// words for n > 16): every lane takes T frontier entries per level step, the
// test-and-sets of all of them in flight together (T > 1: the long searches
// the small path hands on).
// FAM (n > 16, no witness words): an automaton with |a(Q)|, |b(Q)| <= 18 is
// searched in the two-family encoding (every subset past Q lies inside a(Q)
// or inside b(Q); bit 18 = family B, the subsets of b(Q) not inside a(Q);
// a(Q) & b(Q) labelled alike in both families) with its visited bitmap of
// 2^19 bits in LDS, cleared through the arena after each automaton; the
// others keep the n-bit encoding and the bitmap in HBM.
template <class Grp, int T = 1, bool FAM = false>
struct GroupSolver {
    static constexpr uint32_t G = Grp::kSize;
    static constexpr int kWavesPerEu = 1;
    static constexpr uint32_t kFamBit = 1u << kFamBits;
    static constexpr uint32_t kTabBytes = FAM ? (uint32_t)kFamTableBytes : (uint32_t)kTableBytes;
    const BatchParams &P;
    Grp g;
    // LDS
    lds_u64 *tab;
    lds_u32 *bm;
    lds_u32 *arena;
    // HBM slot
    glb_u32 *garena;
    glb_u32 *gbm;
    glb_u32 *good;
    glb_u32 *lvl;
    // automaton being solved (identical in every lane of the group)
    int64_t ai;
    int32_t n;
    uint32_t Q;
    int32_t nsl;
    bool lds_bm;
    bool fam = false;         // FAM: this automaton in the two-family encoding
    bool bm_clean = false;    // FAM: the LDS bitmap is all zero
    uint32_t fam_nI = 0;      // FAM: the labels of b(Q) outside a(Q)
    uint32_t e0 = 0;          // subsets counted outside the arena (Q when encoded)
    uint32_t lo, hi, wr, rd, level;
    // work queue
    int64_t it, it_end;
    uint32_t wq;
    bool taken;
    // progress
    uint32_t done_local;
    uint64_t subsets_local;
    uint64_t expanded_local;  // subsets actually expanded (this launch's statistics)
    uint32_t found_end;       // end of the step that met a singleton

    static constexpr bool kNestedLoop = false;
    // LDS: one region of lds_group_bytes per group, then the group scratch
    __device__ GroupSolver(const BatchParams &p, uint8_t *smem, uint32_t gib, uint32_t groups, uint8_t *slot)
        : GroupSolver(p, smem + (uint64_t)groups * p.lds_group_bytes, smem + (uint64_t)gib * p.lds_group_bytes,
                      slot) {}
    __device__ GroupSolver(const BatchParams &p, void *scratch, uint8_t *gsm, uint8_t *slot)
        : P(p), g(scratch) {
        tab = (lds_u64 *)gsm;
        bm = (lds_u32 *)(gsm + kTabBytes);
        arena = bm + P.lds_bitmap_words;
        garena = (glb_u32 *)(slot + P.garena_off);
        gbm = (glb_u32 *)(slot + P.gbitmap_off);
        good = (glb_u32 *)(slot + P.good_off);
        lvl = (glb_u32 *)(slot + P.lvl_off);
        it = 0;
        it_end = 0;
        wq = 0;
        taken = false;
        done_local = 0;
        subsets_local = 0;
        expanded_local = 0;
        found_end = 0;
    }

    __device__ __forceinline__ uint32_t aget(uint32_t p) const {
        return p < P.lds_arena_cap ? arena[p] : garena[SSW_IDX(p - P.lds_arena_cap, P.garena_cap)];
    }
    __device__ __forceinline__ void aput(uint32_t p, uint32_t v) const {
        if (p < P.lds_arena_cap)
            arena[p] = v;
        else
            garena[SSW_IDX(p - P.lds_arena_cap, P.garena_cap)] = v;
    }
    __device__ __forceinline__ uint64_t image(uint32_t s) const {
        if constexpr (FAM) {
            if (fam) {
                // family B's slices 0..3 after A's, then the shared slice
                const uint32_t fb = ((s >> kFamBits) & 1u) << 6;
                uint64_t r = tab[fb | (s & 15u)] | tab[fb | 16u | ((s >> 4) & 15u)] |
                             tab[fb | 32u | ((s >> 8) & 15u)] | tab[fb | 48u | ((s >> 12) & 15u)] |
                             tab[128u | ((s >> 16) & 15u)];
                // a b-image inside a(Q) belongs to family A
                if (((uint32_t)(r >> 32) & fam_nI) == 0u) r &= ~((uint64_t)kFamBit << 32);
                return r;
            }
        }
        uint64_t r = 0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k < nsl) r |= tab[(k << 4) | ((s >> (4 * k)) & 15u)];
        return r;
    }
    // singleton test of an image (without the family bit)
    __device__ __forceinline__ bool single(uint32_t x) const {
        return is_single(FAM && fam ? (x & (kFamBit - 1u)) : x);
    }
    // test-and-set: true when this lane set the bit first
    __device__ __forceinline__ bool tas(uint32_t x) const {
        const uint32_t bit = 1u << (x & 31u);
        uint32_t old;
        if (lds_bm)
            old = or_wg(bm + SSW_IDX(x >> 5, P.lds_bitmap_words), bit);
        else
            old = or_wg(gbm + SSW_IDX(x >> 5, P.gbitmap_words), bit);
        return (old & bit) == 0u;
    }

    __device__ int64_t next_item() {
        if (P.single_item) {
            if (taken) return -1;
            taken = true;
            return P.first_item;
        }
        return fetch_work(P, g, wq, it, it_end);
    }

    int64_t nx_ai = -1;
    __device__ void prefetch() {
        const int64_t item = next_item();
        nx_ai = item < 0 ? -1 : (P.order ? (int64_t)P.order[item] : item);
    }
    __device__ bool has_next() const { return nx_ai >= 0; }
    __device__ int begin() {
        const int64_t a = nx_ai;
        prefetch();
        if (!init(a)) return -1;
        if constexpr (FAM) {
            if (fam) return init_fam();
        }
        return kContinue;
    }

    // Loads automaton `a`; false when its transitions are invalid.
    __device__ bool init(int64_t a) {
        ai = a;
        n = P.uniform_n > 0 ? P.uniform_n : (int32_t)P.n_per[ai];
        const uint32_t r = g.rank();
        fam = false;
        e0 = 0;
        bool bad = n < 1 || n > 30;
        // the group's HBM bitmap must hold 2^n bits when n > lds_n
        if (!bad && !lds_ok_for(n) && (uint64_t)(n >= 5 ? (1u << (n - 5)) : 1u) > P.gbitmap_words)
            bad = true;
        if (!bad && P.want_witness && (uint64_t)(n >= 5 ? (1u << (n - 5)) : 1u) > P.good_words)
            bad = true;
        if (g.any(bad)) {
            if (r == 0) write_invalid();
            if (P.witness) clear_witness_row();
            return false;
        }
        const uint8_t *d = P.delta + (P.offset ? P.offset[ai] : (uint64_t)ai * 2u * (uint64_t)n);
        nsl = (n + 3) >> 2;
        bool badd = false;
        for (int q = (int)r; q < 2 * n; q += (int)G) badd |= d[q] >= n;
        if (g.any(badd)) {
            if (r == 0) write_invalid();
            if (P.witness) clear_witness_row();
            return false;
        }
        Q = (1u << n) - 1u;
        lds_bm = lds_ok_for(n);
        if constexpr (FAM) {
            // |a(Q)| and |b(Q)| from the transitions (the arena's first words
            // as accumulators; the tables are built over them below)
            if (r == 0) {
                arena[0] = 0u;
                arena[1] = 0u;
            }
            g.sync();
            for (int q = (int)r; q < n; q += (int)G) {
                or_wg(arena, 1u << d[q]);
                or_wg(arena + 1, 1u << d[n + q]);
            }
            g.sync();
            const uint32_t ia = arena[0], ib = arena[1];
            g.sync();
            fam = __popc(ia) <= kFamBits && __popc(ib) <= kFamBits;
            lds_bm = fam;
            if (fam) return setup_fam(d, ia, ib);
        }
        if (lds_bm) {
            const uint32_t words = n >= 5 ? (1u << (n - 5)) : 1u;
            if (words >= 4) {
                lds_u4 *b4 = (lds_u4 *)bm;
                for (uint32_t w = r; w < (words >> 2); w += G) b4[w] = u32x4{0u, 0u, 0u, 0u};
            } else {
                for (uint32_t w = r; w < words; w += G) bm[w] = 0;
            }
        }
        for (int e = (int)r; e < nsl * 16; e += (int)G) {
            const int k = e >> 4, v = e & 15;
            uint64_t acc = 0;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int q = 4 * k + t;
                if (q < n && ((v >> t) & 1))
                    acc |= (1ull << d[q]) | (1ull << (32 + d[n + q]));
            }
            tab[e] = acc;
        }
        g.sync();
        if (r == 0) {
            const uint32_t bit = 1u << (Q & 31u);
            if (lds_bm)
                or_wg(bm + (Q >> 5), bit);
            else
                or_wg(gbm + (Q >> 5), bit);
            arena[0] = Q;
            if (P.want_witness) {
                lvl[0] = 0;
                lvl[1] = 1;
            }
        }
        g.sync();
        lo = 0;
        hi = 1;
        wr = 1;
        rd = 0;
        level = 0;
        return true;
    }

    __device__ bool lds_ok_for(int32_t nn) const { return !FAM && nn <= P.lds_n; }

    // FAM: the tables of the two-family encoding and the image of Q under
    // a and b (ia, ib, n-bit masks).  The labels: a(Q) & b(Q) first, in both
    // families, then the rest of a(Q) (family A) or of b(Q) (family B).
    uint32_t root_ia = 0, root_ib = 0, root_a = 0, root_b = 0;
    __device__ bool setup_fam(const uint8_t *d, uint32_t ia, uint32_t ib) {
        const uint32_t r = g.rank();
        const uint32_t I = ia & ib;
        const uint32_t ka = (uint32_t)__popc(ia), kb = (uint32_t)__popc(ib), ki = (uint32_t)__popc(I);
        const auto lab = [I, ki](uint32_t fs, uint32_t x) {
            const uint32_t below = (1u << x) - 1u;
            return ((I >> x) & 1u) ? (uint32_t)__popc(I & below) : ki + (uint32_t)__popc(fs & ~I & below);
        };
        // singleton images per family and label (arena words 2.., two u32
        // halves each), zero past the family's states
        lds_u64 *pimg = (lds_u64 *)(arena + 2);
        for (uint32_t k = r; k < 2u * kFamBits; k += G) pimg[k] = 0ull;
        if (!bm_clean) {
            for (uint32_t w = r; w < kFamBitmapWords / 4u; w += G) ((lds_u4 *)bm)[w] = u32x4{0u, 0u, 0u, 0u};
            bm_clean = true;
        }
        g.sync();
        for (int q = (int)r; q < n; q += (int)G) {
            const uint32_t x = (uint32_t)q;
            const uint64_t e = (uint64_t)(1u << lab(ia, d[q])) | ((uint64_t)(kFamBit | (1u << lab(ib, d[n + q]))) << 32);
            if ((ia >> x) & 1u) pimg[lab(ia, x)] = e;
            if ((ib >> x) & 1u) pimg[kFamBits + lab(ib, x)] = e;
        }
        g.sync();
        // slices 0..3 of family f at 64 f + 16 k + v; the shared slice at 128 + v
        // (v: bits 0, 1 = labels 16, 17, bit 2 = family)
        for (uint32_t e = r; e < 144u; e += G) {
            uint64_t acc = 0;
            if (e < 128u) {
                const uint32_t f = e >> 6, k = (e >> 4) & 3u, v = e & 15u;
#pragma unroll
                for (uint32_t t = 0; t < 4; ++t)
                    if ((v >> t) & 1u) acc |= pimg[f * kFamBits + 4u * k + t];
            } else {
                const uint32_t v = e - 128u, f = (v >> 2) & 1u;
                if (!(v & 8u)) {
                    if (v & 1u) acc |= pimg[f * kFamBits + 16u];
                    if (v & 2u) acc |= pimg[f * kFamBits + 17u];
                }
            }
            tab[e] = acc;
        }
        fam_nI = (kFamBit - 1u) & ~((1u << ki) - 1u);
        root_ia = ia;
        root_ib = ib;
        root_a = (1u << ka) - 1u;
        const uint32_t rb = (1u << kb) - 1u;
        root_b = (rb & fam_nI) == 0u ? rb : (rb | kFamBit);
        g.sync();
        return true;
    }
    // FAM: level 0 (Q) expanded here, Q counted once in E; level 1 = the
    // new subsets among a(Q), b(Q) in the arena
    __device__ int init_fam() {
        const bool na = root_ia != Q, nb = root_ib != Q && root_ib != root_ia;
        Q = ~0u;  // no encoded subset: the n-bit Q is never met again
        e0 = 1;
        if ((na && single(root_a)) || (nb && single(root_b))) {
            lo = hi = wr = rd = 0;
            level = 0;
            found_end = 0;
            return kFound;
        }
        if (g.rank() == 0) {
            if (na) or_wg(bm + (root_a >> 5), 1u << (root_a & 31u));
            if (nb) or_wg(bm + (root_b >> 5), 1u << (root_b & 31u));
            arena[0] = na ? root_a : root_b;
            arena[1] = root_b;
        }
        g.sync();
        lo = 0;
        rd = 0;
        hi = wr = (na ? 1u : 0u) + (nb ? 1u : 0u);
        level = 1;
        return hi == 0u ? (int)kNoSync : (int)kContinue;
    }

    __device__ void write_invalid() const {
        P.length[ai] = -3;
        if (P.sync) P.sync[ai] = 0;
        if (P.expanded) P.expanded[ai] = 0;
    }

    // an automaton without a reset word gets an all-zero witness row
    __device__ void clear_witness_row() const {
        glb_u8 *out = (glb_u8 *)(P.witness + (uint64_t)ai * (uint64_t)P.witness_stride);
        for (int32_t k = (int32_t)g.rank(); k < P.witness_stride; k += (int32_t)G) out[k] = 0;
    }

    // One level step over up to G * T frontier entries.
    __device__ int step() {
        const uint64_t cap = (uint64_t)P.lds_arena_cap + P.garena_cap;
        if ((uint64_t)wr + 2u * G * T > cap) return kOverflow;
        if (P.want_witness && level + 2 >= P.lvl_cap) return kOverflow;
        bool any_single;
        uint32_t total;
        if constexpr (T == 1) {
            const uint32_t p = rd + g.rank();
            const bool valid = p < hi;
            const uint32_t s = valid ? aget(p) : 0u;
            const uint64_t im = image(s);
            const uint32_t a = (uint32_t)im, b = (uint32_t)(im >> 32);
            // a singleton equal to Q (n = 1) is already visited: never a find,
            // as in SlimCPUUntweaked.cs:149-151
            const bool sa = valid && single(a) && a != Q;
            const bool sb = valid && single(b) && b != Q;
            const bool na = valid && !sa && tas(a);
            const bool nb = valid && !sb && tas(b);
            uint32_t off;
            g.exchange(sa || sb, na, nb, any_single, off, total);
            if (na) aput(wr + off, a);
            if (nb) aput(wr + off + (na ? 1u : 0u), b);
        } else {
            // entries rd + t G + rank; the new subsets of a lane are placed
            // together (lane-major), which leaves lengths and expanded counts
            // as they are (a level's order does not enter them)
            uint32_t v[2 * T];
            bool f[2 * T];
            bool found = false;
#pragma unroll
            for (int t = 0; t < T; ++t) {
                const uint32_t p = rd + (uint32_t)t * G + g.rank();
                const bool valid = p < hi;
                const uint32_t s = valid ? aget(p) : 0u;
                const uint64_t im = image(s);
                v[2 * t] = (uint32_t)im;
                v[2 * t + 1] = (uint32_t)(im >> 32);
                const bool sa = valid && single(v[2 * t]) && v[2 * t] != Q;
                const bool sb = valid && single(v[2 * t + 1]) && v[2 * t + 1] != Q;
                found |= sa || sb;
                f[2 * t] = valid && !sa;
                f[2 * t + 1] = valid && !sb;
            }
#pragma unroll
            for (int j = 0; j < 2 * T; ++j) f[j] = f[j] && tas(v[j]);
            uint32_t pos;
            g.template exchange_flags<2 * T>(found, f, any_single, pos, total);
            pos += wr;
#pragma unroll
            for (int j = 0; j < 2 * T; ++j) {
                if (f[j]) aput(pos, v[j]);
                pos += f[j] ? 1u : 0u;
            }
        }
        wr += total;
        if (any_single) {
            found_end = rd + G * T < hi ? rd + G * T : hi;
            return kFound;
        }
        rd += G * T;
        if (rd >= hi) {
            g.sync();
            ++level;
            lo = hi;
            hi = wr;
            rd = lo;
            if (P.want_witness && g.rank() == 0) lvl[SSW_IDX(level + 1, P.lvl_cap)] = hi;
            if (lo == hi) return kNoSync;
        }
        return kContinue;
    }

    __device__ __forceinline__ bool bit_of(const glb_u32 *m, uint32_t x) const {
        return (m[SSW_IDX(x >> 5, P.good_words)] >> (x & 31u)) & 1u;
    }

    // Lexicographically smallest shortest reset word from the retained levels.
    __device__ void witness(int32_t L) {
        glb_u32 *g0 = good;
        glb_u32 *g1 = good + P.good_words;
        const uint32_t r = g.rank();
        g.sync();
        // backward: mark the subsets of level j with a letter leading into
        // the marked part of level j+1 (level L: the singletons)
        for (int32_t j = L - 1; j >= 0; --j) {
            const uint32_t s = lvl[j], e = lvl[j + 1];
            glb_u32 *dst = (j & 1) ? g1 : g0;
            const glb_u32 *src = (j & 1) ? g0 : g1;
            for (uint32_t base = s; base < e; base += G) {
                const uint32_t p = base + r;
                if (p < e) {
                    const uint32_t x = aget(p);
                    const uint64_t im = image(x);
                    const uint32_t a = (uint32_t)im, b = (uint32_t)(im >> 32);
                    const bool on = (j == L - 1) ? (is_single(a) || is_single(b))
                                                 : (bit_of(src, a) || bit_of(src, b));
                    if (on) or_wg(dst + SSW_IDX(x >> 5, P.good_words), 1u << (x & 31u));
                }
            }
            g.sync();
        }
        // forward: greedy, letter a whenever it stays on a shortest path
        glb_u8 *out = (glb_u8 *)(P.witness + (uint64_t)ai * (uint64_t)P.witness_stride);
        uint32_t cur = Q, acc = 0;
        for (int32_t k = 0; k < L; ++k) {
            const uint64_t im = image(cur);
            const uint32_t a = (uint32_t)im, b = (uint32_t)(im >> 32);
            bool take_a;
            if (k == L - 1) {
                take_a = is_single(a);
            } else {
                const glb_u32 *gs = ((k + 1) & 1) ? g1 : g0;
                take_a = false;
                if (bit_of(gs, a)) {
                    const uint32_t s = lvl[k + 1], e = lvl[k + 2];
                    for (uint32_t base = s; base < e && !take_a; base += G) {
                        const uint32_t p = base + r;
                        take_a = g.any(p < e && aget(p) == a);
                    }
                }
            }
            if (!take_a) acc |= 1u << (k & 7);
            cur = take_a ? a : b;
            if ((k & 7) == 7 || k == L - 1) {
                if (r == 0 && (k >> 3) < P.witness_stride) out[k >> 3] = (uint8_t)acc;
                acc = 0;
            }
        }
        for (int32_t k = ((L + 7) >> 3) + (int32_t)r; k < P.witness_stride; k += (int32_t)G) out[k] = 0;
        g.sync();
        // leave both witness bitmaps clean for the next automaton
        const uint32_t end = lvl[L];
        for (uint32_t p = r; p < end; p += G) {
            const uint32_t x = aget(p);
            g0[SSW_IDX(x >> 5, P.good_words)] = 0;
            g1[SSW_IDX(x >> 5, P.good_words)] = 0;
        }
        g.sync();
    }

    __device__ void finish(int code) {
        const uint32_t r = g.rank();
        if (code == kOverflow) {
            if (r == 0) {
                const uint32_t k = atomicAdd(P.deferred_count, 1u);
                P.deferred_list[k] = (uint32_t)ai;
            }
            expanded_local += rd;  // the device-wide path searches it again
        } else {
            const bool found = code == kFound;
            const int32_t L = found ? (int32_t)level + 1 : -1;
            const int64_t E = (int64_t)hi + e0;
            if (r == 0) {
                P.length[ai] = L;
                if (P.sync) P.sync[ai] = found ? 1 : 0;
                if (P.expanded) P.expanded[ai] = E;
                const int32_t bound = P.bound >= 0 ? P.bound : (n - 1) * (n - 1);
                if (found && L > bound) atomicMin(P.violation, (unsigned long long)ai);
            }
            if (found && P.witness) witness(L);
            if (!found && P.witness) clear_witness_row();
            ++done_local;
            subsets_local += (uint64_t)E;
            expanded_local += (uint64_t)(found ? found_end : hi) + e0;
        }
        if (!lds_bm) {
            g.sync();
            for (uint32_t p = r; p < wr; p += G) gbm[SSW_IDX(aget(p) >> 5, P.gbitmap_words)] = 0;
        } else if (FAM && fam) {
            g.sync();
            for (uint32_t p = r; p < wr; p += G) bm[SSW_IDX(aget(p) >> 5, P.lds_bitmap_words)] = 0;
        }
        g.sync();
        if (done_local >= 256) flush();
    }

    __device__ void flush() {
        if (g.rank() == 0) {
            const uint32_t slot = blockIdx.x & (kCounterSlots - 1u);
            if (done_local) {
                unsigned long long *pg = P.progress + (size_t)slot * kSlotWords;
                atomicAdd(&pg[0], (unsigned long long)done_local);
                atomicAdd(&pg[1], (unsigned long long)subsets_local);
                atomicAdd(&pg[2], (unsigned long long)done_local);
            }
            if (P.launch_stats && (done_local || expanded_local)) {
                unsigned long long *ls = P.launch_stats + (size_t)slot * kStatWords;
                atomicAdd(&ls[0], (unsigned long long)done_local);
                atomicAdd(&ls[1], (unsigned long long)subsets_local);
                atomicAdd(&ls[2], (unsigned long long)expanded_local);
            }
        }
        done_local = 0;
        subsets_local = 0;
        expanded_local = 0;
    }
};

// Small automata (n <= 16) without witness: the whole live frontier sits in an
// LDS ring of 16-bit subsets (ring_cap entries; the unread window is at most
// the wider of two consecutive levels, and C_16's widest level holds 1577
// subsets), images come from packed tables (a | b << 16) of
// four nibble slices (BT = false, 256 bytes) or two byte slices (BT = true,
// 2 KiB, half the lookups), and every lane takes up to T frontier entries per
// level step.  A frontier that outgrows the ring is either listed for a
// second launch with a larger ring (spill) or moves to a 65536-entry ring in
// the group's HBM slot, which no n <= 16 frontier can outgrow.
//
// LDS (SmallLds): the bitmaps of all groups of the workgroup come first, each
// aligned to its size, then one region per group (table, transitions
// staging, singleton images, control and spare words, ring plus 64 dummy
// entries).  With aligned bases
// a bitmap word or table entry address is one OR of the base and the scaled
// index; lanes without a new entry write a dummy ring entry instead of
// branching around the write.
// WIT (one wavefront per automaton only): the search also records, in the
// group's HBM slot, every discovered subset in discovery order (hist) and the
// start of every level in it (lvlp); a found automaton then gets the
// lexicographically smallest shortest reset word from a backward pass over
// the levels (good marks of the next level in the LDS bitmap, which the
// search no longer needs, and one flag per recorded position in HBM) and a
// forward greedy pass.
// FAM (n > 16, one wavefront, T = 1, nibble tables): every subset reached
// from Q but Q itself lies inside a(Q) (the images under a) or inside b(Q)
// (those under b), so it is encoded in 16 bits as a family and a mask of at
// most 15 relabelled states: family A (bit 15 clear) holds the subsets of
// a(Q), family B (bit 15 set) those of b(Q) that are not inside a(Q).  The
// states of a(Q) & b(Q) take the lowest labels in both families, so a
// b-image inside a(Q) becomes family A by clearing bit 15.  The encoding is
// one to one on the reachable subsets, so lengths and expanded counts are
// those of the search over n-bit masks; automata with |a(Q)| or |b(Q)| above
// 15 are handed on.
template <class Grp, int T, bool BT, bool WIT = false, bool FAM = false>
struct SmallSolver {
    static_assert(!WIT || Grp::kSize == 64, "witness words: one wavefront per automaton");
    static_assert(!FAM || (Grp::kSize == 64 && T == 1 && !BT && !WIT), "two-family encoding: one wavefront, T = 1");
    static constexpr uint32_t G = Grp::kSize;
    // the witness launch of one entry per lane keeps the residency its LDS
    // allows (six waves per SIMD) with its witness pass's registers
    static constexpr int kWavesPerEu = (WIT && T == 1) ? 6 : 1;
    static constexpr uint32_t kGlobalRing = 65536;
    // one persistent loop per automaton when a group is a whole wavefront or
    // more (its control flow is uniform)
    static constexpr bool kNestedLoop = G >= 64;
    const BatchParams &P;
    Grp g;
    uint32_t tab_a, stage_a, pimg_a, ctl_a, ring_a, bm_a;  // LDS byte addresses
    // Lanes past the level take their test-and-sets to one of kDummies
    // all-ones words (once an automaton has started: the staging and
    // singleton-image words, and for one wavefront per automaton also the
    // control and spare words; an OR leaves them all ones), so their results
    // read as not new without a branch around the atomics.  A wavefront's
    // lanes get a word each: lanes sharing a word would serialise on it.
    static constexpr uint32_t kDummies = G == 64 ? 64u : (G >= 32 ? 32u : 16u);
    uint32_t dum_a;
    glb_u16 *gring;
    glb_u16 *hist;   // WIT: subsets in discovery order, 65536 entries and 64 dummies
    glb_u32 *himg;   // WIT, T >= 4: the images of the subset at each history position (once expanded)
    glb_u8 *goodp;   // WIT: 1 = the subset at this position of hist is on a shortest path
    glb_u32 *lvlp;   // WIT: start position of each level in hist
    static constexpr uint32_t kHistEntries = 65536u + 64u;
    static constexpr uint32_t kLevelEntries = 65536u + 4u;
    glb_u32 *choicep;  // WIT: bit x = letter a keeps good subset x on a shortest path
    uint32_t rmask, rcap;
    bool in_lds;
    bool narrow;  // relabelled states, visited bitmap of 2^narrow_bits bits
    uint32_t root_a, root_b;  // n > 16: a(Q), b(Q) in the relabelled states
    uint32_t fam_nI;          // FAM: the labels of b(Q) outside a(Q) (family B only)
    uint32_t fam_ia, fam_ib;  // FAM: a(Q), b(Q) in the n-bit labels
    uint32_t e0;              // subsets counted outside the ring (Q when n > 16)
    int64_t ai;
    int32_t n;
    uint32_t Q;
    uint32_t hi, wr, rd, level;
    int64_t it, it_end;
    uint32_t wq;
    bool taken;
    uint32_t done_local;
    uint64_t subsets_local;
    uint64_t expanded_local;  // subsets actually expanded (this launch's statistics)
    uint32_t found_end;       // end of the step that met a singleton (ring position)
    // the next automaton, loaded while the current one is searched
    int64_t nx_ai;
    int32_t nx_n;
    uint32_t nx_b0, nx_b1;
    uint32_t nx_seg, seg;  // work-queue segment of the next / current automaton

    __device__ SmallSolver(const BatchParams &p, uint8_t *smem, uint32_t gib, uint32_t groups, uint8_t *slot)
        : P(p), g(smem + SmallLds::block_bytes(groups, 4u * p.lds_bitmap_words, p.lds_group_bytes, 0)) {
        const uint32_t base = uni<G>((uint32_t)(size_t)(lds_u8 *)smem);
        const uint32_t bm_bytes = 4u * P.lds_bitmap_words;
        // kept in VGPRs: an address is then one v_and_or of the scaled index,
        // a mask in an SGPR and the base (one scalar operand per instruction)
        bm_a = base + gib * bm_bytes;
        tab_a = base + SmallLds::regions_base(groups, bm_bytes) + gib * P.lds_group_bytes;
        stage_a = tab_a + SmallLds::tab_bytes(BT, FAM);
        pimg_a = stage_a + 64;
        ctl_a = pimg_a + 64;
        ring_a = ctl_a + 128;
        gring = (glb_u16 *)(slot + P.garena_off);
        hist = (glb_u16 *)(slot + P.hist_off);
        himg = (glb_u32 *)(slot + P.himg_off);
        goodp = (glb_u8 *)(slot + P.good_off);
        lvlp = (glb_u32 *)(slot + P.lvl_off);
        choicep = (glb_u32 *)(slot + P.gbitmap_off);
        rcap = P.lds_arena_cap;
        rmask = rcap - 1u;
        dum_a = stage_a + 4u * (g.rank() & (kDummies - 1u));
        narrow = P.narrow_bits > 0;
        it = 0;
        it_end = 0;
        wq = 0;
        taken = false;
        done_local = 0;
        subsets_local = 0;
        expanded_local = 0;
        found_end = 0;
        nx_ai = -1;
    }
#ifdef SSW_CHECKED
    // Checked build: every LDS address the solver forms lies in its own
    // visited bitmap or in its own region (tables, staging, singleton images,
    // control words, ring); one that does not is reported with the line that
    // formed it and goes to the staging words instead.
    template <class V>
    __device__ __forceinline__ V *lds_at(uint32_t a, uint32_t line) const {
        const uint32_t w = (uint32_t)sizeof(V);
        const bool in_bm = a >= bm_a && a + w <= bm_a + 4u * P.lds_bitmap_words;
        const bool in_region = a >= tab_a && a + w <= tab_a + P.lds_group_bytes;
        if (!in_bm && !in_region) {
            checked_report(line, a);
            a = stage_a;
        }
        return (V *)(size_t)a;
    }
#define l32(a) lds_at<lds_u32>((a), __LINE__)
#define l16(a) lds_at<lds_u16>((a), __LINE__)
#define l8(a) lds_at<lds_u8>((a), __LINE__)
#define l128(a) lds_at<lds_u4>((a), __LINE__)
#else
    static __device__ __forceinline__ lds_u32 *l32(uint32_t a) { return (lds_u32 *)(size_t)a; }
    static __device__ __forceinline__ lds_u16 *l16(uint32_t a) { return (lds_u16 *)(size_t)a; }
    static __device__ __forceinline__ lds_u8 *l8(uint32_t a) { return (lds_u8 *)(size_t)a; }
    static __device__ __forceinline__ lds_u4 *l128(uint32_t a) { return (lds_u4 *)(size_t)a; }
#endif

    // LDS-only ordering inside one wavefront needs no memory fence: LDS
    // instructions of a wave complete in order
    __device__ __forceinline__ void lsync() const {
        if (G > 64 || !in_lds)
            g.sync();
        else
            __builtin_amdgcn_wave_barrier();
    }

    template <bool LDS>
    __device__ __forceinline__ uint32_t rget(uint32_t p) const {
        return LDS ? (uint32_t)*l16(ring_a + ((p & rmask) << 1)) : (uint32_t)gring[p & (kGlobalRing - 1u)];
    }
    // all slices are built (slices past n hold zeros), so the lookups are
    // unconditional and issue back to back
    __device__ __forceinline__ uint32_t image(uint32_t s) const {
        if constexpr (FAM) {
            // slices 0..2 per family (family B's 256 bytes above A's), slice
            // 3 shared: its index holds the family bit
            const uint32_t fb = tab_a + ((s >> 7) & 0x100u);
            const uint32_t im = *l32(fb | ((s << 2) & 60u)) | *l32((fb | ((s >> 2) & 60u)) + 64u) |
                                *l32((fb | ((s >> 6) & 60u)) + 128u) | *l32((tab_a | ((s >> 10) & 60u)) + 192u);
            // a b-image inside a(Q) belongs to family A
            return ((im >> 16) & fam_nI) == 0u ? (im & 0x7FFFFFFFu) : im;
        } else if constexpr (BT) {
            return *l32(tab_a | ((s << 2) & 0x3FCu)) | *l32((tab_a | ((s >> 6) & 0x3FCu)) + 1024u);
        } else {
            return *l32(tab_a | ((s << 2) & 60u)) | *l32((tab_a | ((s >> 2) & 60u)) + 64u) |
                   *l32((tab_a | ((s >> 6) & 60u)) + 128u) | *l32((tab_a | ((s >> 10) & 60u)) + 192u);
        }
    }

    __device__ int64_t next_item() {
        if (P.single_item) {
            if (taken) return -1;
            taken = true;
            return P.first_item;
        }
        return fetch_work(P, g, wq, it, it_end);
    }

    __device__ void write_invalid() const {
        P.length[ai] = -3;
        if (P.sync) P.sync[ai] = 0;
        if (P.expanded) P.expanded[ai] = 0;
    }

    // Takes the next automaton from the work queue and issues the loads of
    // its size and transitions; they are consumed by start().
    __device__ __forceinline__ void prefetch() {
        const int64_t item = next_item();
        if (item < 0) {
            nx_ai = -1;
            return;
        }
        nx_ai = uni<G>(P.order ? (int64_t)P.order[item] : item);
        nx_seg = wq;
        nx_n = uni<G>(P.uniform_n > 0 ? P.uniform_n : (int32_t)P.n_per[nx_ai]);
        nx_b0 = nx_b1 = 0;
        if (nx_n >= 1 && nx_n <= P.lds_n) {
            const uint8_t *d = P.delta + (P.offset ? P.offset[nx_ai] : (uint64_t)nx_ai * 2u * (uint64_t)nx_n);
            const uint32_t r = g.rank();
            if (r < 2u * (uint32_t)nx_n) nx_b0 = d[r];
            if (G < 32 && r + G < 2u * (uint32_t)nx_n) nx_b1 = d[r + G];
        }
    }
    __device__ bool has_next() const { return nx_ai >= 0; }

    // Starts the prefetched automaton and expands its root level: -1 when its
    // input is invalid (the result is written), kFound when a(Q) or b(Q) is
    // a singleton, kNoSync when Q is the only reachable subset (both letters
    // permutations, or n = 1: not synchronizing, one subset expanded), else
    // kContinue with level 1 in the ring.
    __device__ __forceinline__ int begin() { return start(); }
    __device__ __forceinline__ int start() {
        ai = nx_ai;
        seg = nx_seg;
        n = uni<G>(nx_n);
        const uint32_t r = g.rank();
        const uint32_t b0 = nx_b0, b1 = nx_b1;
        // n > 16 only through the relabelled path (a wavefront stages 2n <= 60 bytes)
        const bool bad_n = n < 1 || n > P.lds_n || (n > 16 && (!narrow || G < 64));
        const uint32_t two_n = bad_n ? 0u : 2u * (uint32_t)n;
        bool badd = (r < two_n && b0 >= (uint32_t)n) || (G < 32 && r + G < two_n && b1 >= (uint32_t)n);
        if (r < two_n) *l8(stage_a + r) = (uint8_t)b0;
        if (G < 32 && r + G < two_n) *l8(stage_a + r + G) = (uint8_t)b1;
        prefetch();
        in_lds = true;
        if (g.any(bad_n || badd)) {
            if (r == 0) write_invalid();
            if constexpr (WIT) zero_row(0);
            return -1;
        }
        Q = uni<G>((1u << n) - 1u);
        lsync();
        uint32_t qa = 0, qb = 0;  // the images of state r under a and b
        if (r < (uint32_t)n) {
            qa = *l8(stage_a + r);
            qb = *l8(stage_a + (uint32_t)n + r);
        }
        uint32_t words;
        if constexpr (FAM) {
            // the two families (the encoding is described above the solver);
            // lanes 0..n-1 hold the states (n <= 30, one wavefront)
            const uint32_t ia = g.or_low32(r < (uint32_t)n ? 1u << qa : 0u);
            const uint32_t ib = g.or_low32(r < (uint32_t)n ? 1u << qb : 0u);
            const uint32_t I = ia & ib;
            const uint32_t ka = (uint32_t)__popc(ia), kb = (uint32_t)__popc(ib), ki = (uint32_t)__popc(I);
            if (ka > 15u || kb > 15u) return spill_to(P.spill_queue, P.spill_list, 0u, P.spill_base);
            // label of state x in the family of `fam` (a(Q) or b(Q)): a(Q) & b(Q) first
            const auto lab = [I, ki](uint32_t fam, uint32_t x) {
                const uint32_t below = (1u << x) - 1u;
                return ((I >> x) & 1u) ? (uint32_t)__popc(I & below) : ki + (uint32_t)__popc(fam & ~I & below);
            };
            // singleton images per family and label (family B's in the ring
            // area until the tables are built), zero past the family's states
            if (r < 16u) {
                *l32(pimg_a + 4u * r) = 0u;
                *l32(ring_a + 4u * r) = 0u;
            }
            lsync();
            if (r < (uint32_t)n) {
                const uint32_t e = (1u << lab(ia, qa)) | ((0x8000u | (1u << lab(ib, qb))) << 16);
                if ((ia >> r) & 1u) *l32(pimg_a + 4u * lab(ia, r)) = e;
                if ((ib >> r) & 1u) *l32(ring_a + 4u * lab(ib, r)) = e;
            }
            fam_nI = 0x7FFFu & ~((1u << ki) - 1u);
            fam_ia = ia;
            fam_ib = ib;
            root_a = (1u << ka) - 1u;
            const uint32_t rb = (1u << kb) - 1u;
            root_b = (rb & fam_nI) == 0u ? rb : (rb | 0x8000u);
            words = 2048u;
        } else if (narrow) {
            // Every subset reached from Q but Q itself lies inside
            // U = a(Q) | b(Q) (the image of any subset is inside the image of
            // Q).  The states are relabelled so that U takes the |U| lowest
            // labels; the visited bitmap then needs 2^|U| bits, and automata
            // with |U| above the bitmap's width go to the next launch.
            // The root level is expanded here: level 1 = {a(Q), b(Q)} comes
            // from two more reductions, which costs less than the search's
            // step over Q alone.  Automata with n > 16 (lanes 0..n-1,
            // n <= 30, one wavefront) are solved this way too when |U| <= 16
            // (their root Q would not fit the 16-bit ring anyway).
            const uint32_t U = g.or_low32(r < (uint32_t)n ? (1u << qa) | (1u << qb) : 0u);
            const uint32_t k = (uint32_t)__popc(U);
            if (k > (uint32_t)P.narrow_bits) {
                if (P.skip_list && k == (uint32_t)n) {
                    const uint32_t ia = g.or_low32(r < (uint32_t)n ? 1u << qa : 0u);
                    const uint32_t ib = g.or_low32(r < (uint32_t)n ? 1u << qb : 0u);
                    if (ia == Q || ib == Q) return spill_to(P.skip_queue, P.skip_list, 0u, P.skip_base);
                }
                return spill_to(P.spill_queue, P.spill_list, 0u, P.spill_base);
            }
            const auto label = [U, k](uint32_t x) {
                const uint32_t below = (uint32_t)__popc(U & ((1u << x) - 1u));
                return ((U >> x) & 1u) ? below : k + x - below;
            };
            // singleton images of the relabelled states 0..15 (all of U), zero past n
            if (r < (uint32_t)n) {
                const uint32_t lr = label(r);
                if (lr < 16u) *l32(pimg_a + 4u * lr) = (1u << label(qa)) | (1u << (16u + label(qb)));
            } else if (r < 16u) {
                *l32(pimg_a + 4u * r) = 0u;
            }
            root_a = g.or_low32(r < (uint32_t)n ? 1u << label(qa) : 0u);
            root_b = g.or_low32(r < (uint32_t)n ? 1u << label(qb) : 0u);
            words = 1u << (P.narrow_bits - 5);
        } else {
            // singleton images of the 16 states, zero past n
            if (r < 16u) *l32(pimg_a + 4u * r) = r < (uint32_t)n ? (1u << qa) | (1u << (16u + qb)) : 0u;
            root_a = g.or_low32(r < (uint32_t)n ? 1u << qa : 0u);
            root_b = g.or_low32(r < (uint32_t)n ? 1u << qb : 0u);
            words = n >= 5 ? (1u << (n - 5)) : 1u;
        }
        if (words >= 4) {
            for (uint32_t w = r; w < (words >> 2); w += G) *l128(bm_a + 16u * w) = u32x4{0u, 0u, 0u, 0u};
        } else {
            for (uint32_t w = r; w < words; w += G) *l32(bm_a + 4u * w) = 0u;
        }
        lsync();
        if constexpr (FAM) {
            // family f's slices 0..2 at tab_a + 256 f + 64 j, entry v = OR of
            // the images of labels 4j + (bits of v); slice 3 at tab_a + 192,
            // entry v = family bit (8) and labels 12..14 (bits 0..2)
            for (uint32_t e = r; e < 112u; e += G) {
                const uint32_t f = e < 96u ? e / 48u : (e - 96u) >> 3;
                const uint32_t j = e < 96u ? (e % 48u) >> 4 : 3u;
                const uint32_t v = e < 96u ? (e & 15u) : (e & 7u);
                const u32x4 p4 = *l128((f ? ring_a : pimg_a) + 16u * j);
                const uint32_t val =
                    ((v & 1u) ? p4.x : 0u) | ((v & 2u) ? p4.y : 0u) | ((v & 4u) ? p4.z : 0u) | ((v & 8u) ? p4.w : 0u);
                *l32(e < 96u ? tab_a + 256u * f + 64u * j + 4u * v : tab_a + 192u + 4u * (e - 96u)) = val;
            }
        } else if constexpr (BT) {
            // 64 blocks of 8 entries: slice k = blk >> 5, entries v = 8 * (blk & 31) + i
            for (uint32_t blk = r; blk < 64u; blk += G) {
                const uint32_t k = blk >> 5, h5 = blk & 31u;
                const u32x4 lo4 = *l128(pimg_a + 32u * k), hi4 = *l128(pimg_a + 32u * k + 16u);
                const uint32_t h = ((h5 & 1u) ? lo4.w : 0u) | ((h5 & 2u) ? hi4.x : 0u) | ((h5 & 4u) ? hi4.y : 0u) |
                                   ((h5 & 8u) ? hi4.z : 0u) | ((h5 & 16u) ? hi4.w : 0u);
                const uint32_t p01 = lo4.x | lo4.y;
                const uint32_t e = tab_a + 1024u * k + 32u * h5;
                *l128(e) = u32x4{h, h | lo4.x, h | lo4.y, h | p01};
                *l128(e + 16u) = u32x4{h | lo4.z, h | lo4.x | lo4.z, h | lo4.y | lo4.z, h | p01 | lo4.z};
            }
        } else {
            for (uint32_t e = r; e < 64u; e += G) {
                const uint32_t v = e & 15u;
                const u32x4 p4 = *l128(pimg_a + 16u * (e >> 4));
                *l32(tab_a + 4u * e) = ((v & 1u) ? p4.x : 0u) | ((v & 2u) ? p4.y : 0u) | ((v & 4u) ? p4.z : 0u) |
                                       ((v & 8u) ? p4.w : 0u);
            }
        }
        lsync();
        if constexpr (G == 64 && !WIT && !FAM) {
            // Verify mode: lane r applies the word of length 6 whose letter k
            // is b when bit k of r is set (the 64 words one wavefront holds;
            // SuperSlimGPUBreakthrough.cs:183-256 does the same with one
            // thread per word); a singleton among the results proves L <= 6,
            // which no bound >= 6 can be below, so the search is skipped
            if (P.prefilter && (P.bound >= 0 ? P.bound : (n - 1) * (n - 1)) >= 6) {
                uint32_t cur = Q;
#pragma unroll
                for (uint32_t k = 0; k < 6u; ++k) {
                    const uint32_t im = image(cur);
                    cur = ((r >> k) & 1u) ? (im >> 16) : (im & 0xFFFFu);
                }
                if (g.any(is_single(cur))) return kPrefiltered;
            }
        }
        // The root level is expanded here: level 1 = the distinct subsets
        // among a(Q), b(Q) other than Q (a permutation letter maps Q to
        // itself) seeds the ring, and Q counts once in E.  With no new
        // subset Q is the only reachable one: not synchronizing, E = 1 (this
        // covers n = 1, where Q is already a singleton).
        // (encoded: compared in the n-bit labels, where a permutation letter
        // maps Q to itself)
        const bool na = FAM ? fam_ia != Q : root_a != Q;
        const bool nb = FAM ? (fam_ib != Q && fam_ib != fam_ia) : (root_b != Q && root_b != root_a);
        if (r == 0) {
            // relabelled, Q lies outside the narrow bitmap and is never reached again
            if (!narrow) or_wg(l32(bm_a | ((Q >> 3) & 0x1FFCu)), 1u << (Q & 31u));
            if (na) or_wg(l32(bm_a | ((root_a >> 3) & 0x1FFCu)), 1u << (root_a & 31u));
            if (nb) or_wg(l32(bm_a | ((root_b >> 3) & 0x1FFCu)), 1u << (root_b & 31u));
            *l16(ring_a) = (uint16_t)(na ? root_a : root_b);
            *l16(ring_a + 2u) = (uint16_t)root_b;
            if constexpr (G > 64) {
#pragma unroll
                for (uint32_t k = 0; k < 6u; ++k) *l32(ctl_a + 4u * k) = 0u;
            }
        }
        // the staging words are read by now: they become the all-ones dummy
        // words of the test-and-sets
        if (r < kDummies) *l32(stage_a + 4u * r) = ~0u;
        lsync();
        rd = 0;
        e0 = 1;
        if ((na && is_single(root_a)) || (nb && is_single(FAM ? (root_b & 0x7FFFu) : root_b))) {
            hi = wr = 0;
            level = 0;
            found_end = 0;
            return kFound;  // length 1, one subset expanded
        }
        hi = wr = (na ? 1u : 0u) + (nb ? 1u : 0u);
        level = 1;
        if constexpr (WIT) {
            if (r == 0 && hi > 0u) {
                if (na) hist[0] = (uint16_t)root_a;
                if (nb) hist[na ? 1u : 0u] = (uint16_t)root_b;
                lvlp[1] = 0u;
                lvlp[2] = hi;
            }
        }
        return hi == 0u ? (int)kNoSync : (int)kContinue;
    }

    __device__ void migrate() {
        for (uint32_t p = rd + g.rank(); p < wr; p += G) gring[p & (kGlobalRing - 1u)] = (uint16_t)rget<true>(p);
        g.sync();
        in_lds = false;
    }

    __device__ __forceinline__ int step() {
        if constexpr (G > 64) return step_level();
        if constexpr (T == 1 && G == 64) {
            // one entry per lane, one wavefront, LDS ring: the steps of the
            // search run in one tight loop with a single path and a single
            // exit test, its bookkeeping in selects (the launch that takes
            // every automaton spends most of its instructions here); the
            // general path below takes over when the ring is short
            if (in_lds) {
                const uint32_t lim = rcap - 2u * G;
                while (wr - rd <= lim) {
                    if (step_body<true, 1>()) {
                        found_end = min(rd + G, hi);
                        return kFound;
                    }
                    rd += G;
                    if (rd >= hi) {
                        ++level;
                        rd = hi;
                        hi = wr;
                        if (rd == hi) return kNoSync;
                        if constexpr (WIT) record_level();
                    }
                }
            }
        } else if constexpr (G == 64) {
            // the same for T > 1 (the spill launch): loop while the ring holds
            // the widest step
            if (in_lds) {
                while (wr + 2u * G * T - rd <= rcap) {
                    const uint32_t left = hi - rd;
                    int c;
                    if (T >= 8 && left > 4u * G)
                        c = step_in<true, (T >= 8 ? 8 : 1)>();
                    else if (T >= 4 && left > 2u * G)
                        c = step_in<true, (T >= 4 ? 4 : 1)>();
                    else if (T >= 2 && left > G)
                        c = step_in<true, (T >= 2 ? 2 : 1)>();
                    else
                        c = step_in<true, 1>();
                    if (c != kContinue) return c;
                }
            }
        }
        // entries per lane: the fewest steps for the width left in the level
        // (a step's instruction count grows slower than its entries, so one
        // T = 4 step over 150..256 entries costs less than T = 2 plus T = 1
        // steps, and it saves their LDS round trips)
        const uint32_t left = hi - rd;
        uint32_t tt = (T >= 8 && left > 4u * G)   ? 8u
                      : (T >= 4 && left > 2u * G) ? 4u
                                                  : ((T >= 2 && left > G) ? 2u : 1u);
        // the ring holds the unread window plus two images per entry taken:
        // fewer entries per lane while a wider step does not fit, then spill
        // (or move to the HBM ring)
        if (in_lds) {
            while (tt > 1u && wr + 2u * G * tt - rd > rcap) tt >>= 1;
            if (wr + 2u * G * tt - rd > rcap) {
                if (P.spill) return spill();
                migrate();
            }
        }
        // one code path per ring location, so the LDS path carries no wait on
        // outstanding HBM loads (the prefetched next automaton)
        if constexpr (T >= 8) {
            if (tt == 8u) return in_lds ? step_in<true, 8>() : step_in<false, 8>();
        }
        if constexpr (T >= 4) {
            if (tt == 4u) return in_lds ? step_in<true, 4>() : step_in<false, 4>();
        }
        if constexpr (T >= 2) {
            if (tt == 2u) return in_lds ? step_in<true, 2>() : step_in<false, 2>();
        }
        return in_lds ? step_in<true, 1>() : step_in<false, 1>();
    }

    // ---- groups of several wavefronts (G > 64) ----
    // The waves split each level into chunks of 64 * tt entries (wave w takes
    // chunks w, w + W, ...), append their new subsets to the shared ring after
    // one LDS atomicAdd per chunk on the level's counter, and meet at one
    // barrier per level.  The counters (ctl words 0..2) and found flags (3..5)
    // rotate over three slots: slot l % 3 serves level l, and slot
    // (l + 2) % 3 is cleared after level l's barrier, when every wave has read
    // it for level l - 1.  A level whose entries plus all the images they can
    // add do not fit the LDS ring is run on the HBM ring (copied there and
    // back at level boundaries).
    static constexpr uint32_t kW = G / 64;
    __device__ __forceinline__ lds_u32 *ctl(uint32_t k) const { return l32(ctl_a + 4u * k); }

    __device__ void relocate(bool to_lds) {
        for (uint32_t p = rd + g.rank(); p < hi; p += G) {
            if (to_lds)
                *l16(ring_a + ((p & rmask) << 1)) = gring[p & (kGlobalRing - 1u)];
            else
                gring[p & (kGlobalRing - 1u)] = (uint16_t)rget<true>(p);
        }
        g.sync();
        in_lds = to_lds;
    }

    __device__ int step_level() {
        const uint32_t lo = rd, width = hi - lo;
        const bool fits = 3u * width <= rcap;
        if (in_lds != fits) {
            if (!fits && P.spill) return spill();
            relocate(fits);
        }
        const uint32_t share = (width + kW - 1u) / kW;
        const uint32_t tt = (T >= 4 && share > 128u) ? 4u : ((T >= 2 && share > 64u) ? 2u : 1u);
        const uint32_t slot = level % 3u;
        const uint32_t wave = g.wave;
        bool single = false;
        for (uint32_t p0 = lo + wave * 64u * tt; p0 < hi; p0 += kW * 64u * tt) {
            if constexpr (T >= 4) {
                if (tt == 4u) {
                    single |= in_lds ? chunk<true, 4>(p0, slot) : chunk<false, 4>(p0, slot);
                    continue;
                }
            }
            if constexpr (T >= 2) {
                if (tt == 2u) {
                    single |= in_lds ? chunk<true, 2>(p0, slot) : chunk<false, 2>(p0, slot);
                    continue;
                }
            }
            single |= in_lds ? chunk<true, 1>(p0, slot) : chunk<false, 1>(p0, slot);
        }
        if (single && g.lane == 0) *ctl(3u + slot) = 1u;
        g.sync();
        const uint32_t added = __builtin_amdgcn_readfirstlane(*ctl(slot));
        const bool found = __builtin_amdgcn_readfirstlane(*ctl(3u + slot)) != 0u;
        if (g.rank() == 0) {
            const uint32_t c = (level + 2u) % 3u;
            *ctl(c) = 0u;
            *ctl(3u + c) = 0u;
        }
        wr = hi + added;
        if (found) {
            found_end = hi;
            return kFound;
        }
        if (added == 0u) return kNoSync;
        ++level;
        rd = hi;
        hi = wr;
        return kContinue;
    }

    // One wave's chunk of 64 * TT entries from p0 (p0 < hi); true when an
    // image is a singleton (uniform over the wave).
    template <bool LDS, int TT>
    __device__ bool chunk(uint32_t p0, uint32_t slot) {
        const uint32_t lane = g.lane;
        uint32_t s[TT];
        bool v[TT];
#pragma unroll
        for (int t = 0; t < TT; ++t) {
            const uint32_t p = p0 + t * 64u + lane;
            v[t] = p < hi;
            s[t] = rget<LDS>(min(p, hi - 1u));
        }
        uint32_t img[TT];
#pragma unroll
        for (int t = 0; t < TT; ++t) img[t] = image(s[t]);
        bool nw[2 * TT];
        uint64_t m[2 * TT], sm;
        expand<TT>(img, v, nw, m, sm);
        // wave-local flag-major positions, then one reservation on the
        // level's counter
        uint32_t pos[2 * TT], total = 0;
#pragma unroll
        for (int j = 0; j < 2 * TT; ++j) {
            pos[j] = __builtin_amdgcn_mbcnt_hi((uint32_t)(m[j] >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m[j], total));
            total += (uint32_t)__popcll(m[j]);
        }
        total = __builtin_amdgcn_readfirstlane(total);
        if (total) {
            uint32_t base = 0;
            if (lane == 0) base = __hip_atomic_fetch_add(ctl(slot), total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            base = hi + __builtin_amdgcn_readfirstlane(base);
#pragma unroll
            for (int j = 0; j < 2 * TT; ++j) {
                const uint32_t val = (j & 1) ? (img[j >> 1] >> 16) : (img[j >> 1] & 0xFFFFu);
                if constexpr (LDS) {
                    const uint32_t q = nw[j] ? ((base + pos[j]) & rmask) : (rcap + lane);
                    *l16(ring_a + (q << 1)) = (uint16_t)val;
                } else {
                    if (nw[j]) gring[(base + pos[j]) & (kGlobalRing - 1u)] = (uint16_t)val;
                }
            }
        }
        return sm != 0;
    }

    // Hands the automaton to the next launch (larger LDS ring), which
    // searches it again; `partial` = subsets expanded here so far.
    __device__ int spill() { return spill_to(P.spill_queue, P.spill_list, rd + e0, P.spill_base); }
    // list entries are the automaton's index plus `base` (a list shared by
    // several batches that are parts of one: their place in the whole)
    __device__ int spill_to(uint32_t *queue, uint32_t *list, uint32_t partial, uint32_t base) {
        if (g.rank() == 0) {
            const uint32_t k = atomicAdd(queue + seg * kQueueWords + 1, 1u);
            list[k] = (uint32_t)ai + base;
        }
        expanded_local += partial;
        return kSpilled;
    }

    // Images img[t] of up to TT entries per lane (v[t]: entry in the level):
    // their test-and-sets, nw = the image is new, and the wavefront ballots
    // m[2t], m[2t+1] (new images under a, b) and sm (a singleton image).
    // Every ballot is of a plain compare, so the lane masks are used as the
    // compares produce them.  A lane past the level repeats the level's last
    // entry hi - 1, which then lies in the same step (or chunk): its singleton
    // test is that of the lane owning that entry, and its test-and-sets go to
    // an all-ones dummy word (results not new).  With no branch around them,
    // the test-and-sets of all TT entries are issued before the first wait.
    // The halves of y = x & (x - 0x10001) are a & (a - 1) and b & (b - 1) of
    // the nonempty images a (low half) and b (high half).  Q is never a
    // singleton here (n = 1 ends at start), and the search ends at the first
    // singleton, so a singleton image is always new; it is still marked and
    // appended, which is harmless because the search stops at this level.
    template <int TT>
    __device__ __forceinline__ void expand(const uint32_t (&img)[TT], const bool (&v)[TT], bool (&nw)[2 * TT],
                                           uint64_t (&m)[2 * TT], uint64_t &sm) const {
        uint32_t oa[TT], ob[TT], ba[TT], bb[TT];
#pragma unroll
        for (int t = 0; t < TT; ++t) {
            const uint32_t x = img[t];
            ba[t] = 1u << (x & 31u);
            bb[t] = 1u << ((x >> 16) & 31u);
            // a before b, so a == b is counted once; word of a: bits 5..15 of
            // x, of b: bits 21..31
            oa[t] = or_wg(l32(v[t] ? (bm_a | ((x >> 3) & 0x1FFCu)) : dum_a), ba[t]);
            ob[t] = or_wg(l32(v[t] ? (bm_a | ((x >> 19) & 0x1FFCu)) : dum_a), bb[t]);
        }
        // every test-and-set in flight before the first result is used
        if constexpr (TT > 1) __builtin_amdgcn_sched_barrier(0);
        sm = 0;
#pragma unroll
        for (int t = 0; t < TT; ++t) {
            const uint32_t x = FAM ? (img[t] & 0x7FFF7FFFu) : img[t];  // without the family bit
            const uint32_t y = x & (x - 0x00010001u);
            // a half of y is zero: min of the halves (one v_min_u16 with a
            // word select) is zero
            const u16x2 yh = __builtin_bit_cast(u16x2, y);
            sm |= __ballot(__builtin_elementwise_min(yh, yh.yx).x == 0);
            nw[2 * t] = (oa[t] & ba[t]) == 0u;
            nw[2 * t + 1] = (ob[t] & bb[t]) == 0u;
            m[2 * t] = __ballot(nw[2 * t]);
            m[2 * t + 1] = __ballot(nw[2 * t + 1]);
        }
    }

    // One step: TT entries per lane from rd; appends the new images, advances
    // wr; true when an image is a singleton.
    template <bool LDS, int TT>
    __device__ bool step_body() {
        const uint32_t r = g.rank();
        uint32_t s[TT];
        bool v[TT];
#pragma unroll
        for (int t = 0; t < TT; ++t) {
            const uint32_t p = rd + t * G + r;
            v[t] = p < hi;
            // a lane past the level repeats entry hi - 1 (in this step): its
            // singleton test is that of the lane owning it
            s[t] = rget<LDS>(min(p, hi - 1u));
        }
        uint32_t img[TT];
#pragma unroll
        for (int t = 0; t < TT; ++t) img[t] = image(s[t]);
        bool nw[2 * TT];
        uint64_t m[2 * TT], sm;
        expand<TT>(img, v, nw, m, sm);
        if constexpr (WIT && T >= 4) {
            // the images of this step's entries at their history positions
            // (the backward pass of the witness reads them there)
#pragma unroll
            for (int t = 0; t < TT; ++t)
                if (v[t]) himg[SSW_IDX(rd + t * G + r, kHistEntries)] = img[t];
        }
        bool any_single;
        uint32_t pos[2 * TT], total;
        g.template place<2 * TT>(sm, m, wr, any_single, pos, total);
        if constexpr (WIT) {
            // lanes without a new subset write a dummy entry past the 65536
            // (one per lane) instead of branching around the store
#pragma unroll
            for (int j = 0; j < 2 * TT; ++j)
                hist[SSW_IDX(nw[j] ? (pos[j] & 0xFFFFu) : 65536u + (r & 63u), kHistEntries)] =
                    (uint16_t)((j & 1) ? (img[j >> 1] >> 16) : (img[j >> 1] & 0xFFFFu));
        }
#pragma unroll
        for (int j = 0; j < 2 * TT; ++j) {
            const uint32_t val = (j & 1) ? (img[j >> 1] >> 16) : (img[j >> 1] & 0xFFFFu);
            if constexpr (LDS) {
                const uint32_t q = nw[j] ? (pos[j] & rmask) : (rcap + (r & 63u));
                *l16(ring_a + (q << 1)) = (uint16_t)val;
            } else {
                if (nw[j]) gring[pos[j] & (kGlobalRing - 1u)] = (uint16_t)val;
            }
        }
        wr = uni<G>(wr + total);
        return any_single;
    }

    template <bool LDS, int TT>
    __device__ int step_in() {
        if (step_body<LDS, TT>()) {
            found_end = min(rd + G * TT, hi);
            return kFound;
        }
        rd = uni<G>(rd + G * TT);
        if (rd >= hi) {
            if (G > 64 || !LDS)
                g.sync();
            else
                __builtin_amdgcn_wave_barrier();
            ++level;
            rd = hi;
            hi = wr;
            if (rd == hi) return kNoSync;
            if constexpr (WIT) record_level();
        }
        return kContinue;
    }

    // ---- witness words (WIT) ----
    __device__ void record_level() {
        if (g.rank() == 0) lvlp[SSW_IDX(level + 1u, kLevelEntries)] = hi;
    }
    __device__ glb_u8 *row() const { return (glb_u8 *)(P.witness + (uint64_t)ai * (uint64_t)P.witness_stride); }
    // zero bytes from..stride-1 of the automaton's witness row
    __device__ void zero_row(int32_t from) const {
        glb_u8 *out = row();
        for (int32_t k = from + (int32_t)g.rank(); k < P.witness_stride; k += (int32_t)G) out[k] = 0;
    }
    // the LDS bitmap cleared (after the search it holds good marks)
    __device__ void clear_bitmap() const {
        const uint32_t words = P.lds_bitmap_words;
        if (words >= 4) {
            for (uint32_t w = g.rank(); w < (words >> 2); w += G) *l128(bm_a + 16u * w) = u32x4{0u, 0u, 0u, 0u};
        } else {
            for (uint32_t w = g.rank(); w < words; w += G) *l32(bm_a + 4u * w) = 0u;
        }
    }
    __device__ __forceinline__ bool marked(uint32_t y) const {
        return (*l32(bm_a | ((y >> 3) & 0x1FFCu)) >> (y & 31u)) & 1u;
    }
    // Where the passes find the recorded search.  When the whole history is
    // still in the LDS ring (no wrap, never moved to HBM) and there is room
    // after it, the flags and level starts go there too and the passes touch
    // no HBM; otherwise the history and flags are read from the HBM slot and
    // only the level starts are copied into the ring area.
    struct WitnessSpace {
        bool local;
        uint32_t flags_a, starts_a;  // LDS byte addresses (local: flags and starts; else starts only)
        bool starts_cached;
    };
    __device__ WitnessSpace witness_space(int32_t L) const {
        WitnessSpace w;
        const uint32_t room = 2u * (rcap + 64u);
        const uint32_t fo = (2u * wr + 3u) & ~3u;
        const uint32_t so = (fo + wr + 3u) & ~3u;
        w.local = in_lds && wr <= rcap && so + 4u * ((uint32_t)L + 1u) <= room;
        w.flags_a = ring_a + fo;
        w.starts_a = w.local ? ring_a + so : ring_a;
        w.starts_cached = w.local || 4u * ((uint32_t)L + 1u) <= room;
        return w;
    }
    __device__ __forceinline__ uint32_t level_start(const WitnessSpace &w, uint32_t k) const {
        return w.starts_cached ? uni<G>(*l32(w.starts_a + 4u * k)) : uni<G>(lvlp[SSW_IDX(k, kLevelEntries)]);
    }
    __device__ __forceinline__ uint32_t hist_at(const WitnessSpace &w, uint32_t p) const {
        return w.local ? (uint32_t)*l16(ring_a + 2u * p) : (uint32_t)hist[SSW_IDX(p, kHistEntries)];
    }
    __device__ __forceinline__ uint32_t flag_at(const WitnessSpace &w, uint32_t p) const {
        return w.local ? (uint32_t)*l8(w.flags_a + p) : (uint32_t)goodp[SSW_IDX(p, 65536u)];
    }
    __device__ __forceinline__ void set_flag(const WitnessSpace &w, uint32_t p, bool on) const {
        if (w.local)
            *l8(w.flags_a + p) = on ? 1u : 0u;
        else
            goodp[SSW_IDX(p, 65536u)] = on ? 1u : 0u;
    }
    // witness() for a local history of at most kRegEntries * G subsets (most
    // searches of the first launches, one entry per lane): lane r holds
    // entries r + G u, u < kRegEntries, and their images, all looked up at once; a level's good
    // flags are bits of `good`.  Per level the backward pass then costs one
    // round of bitmap reads and one of marks, and per letter the forward pass
    // one image lookup and a ballot over the registers (the LDS pass reads the
    // history, its images and the flags again at every level and letter).
    // A slot u takes part in level [s, e) only when s < (u + 1) G and
    // e > u G, a test uniform over the wavefront.
    static constexpr uint32_t kRegEntries = 2;
    __device__ void witness_regs(int32_t L, const WitnessSpace &w) {
        const uint32_t r = g.rank();
        uint32_t x[kRegEntries], im[kRegEntries];
#pragma unroll
        for (uint32_t u = 0; u < kRegEntries; ++u) {
            const uint32_t p = r + u * G;
            x[u] = p < wr ? hist_at(w, p) : 0u;
        }
#pragma unroll
        for (uint32_t u = 0; u < kRegEntries; ++u) im[u] = image(x[u]);
        uint32_t good = 0;
        for (int32_t j = L - 1; j >= 1; --j) {
            const uint32_t s = level_start(w, j), e = level_start(w, j + 1);
            const bool last = j == L - 1;
            uint32_t on = 0;
#pragma unroll
            for (uint32_t u = 0; u < kRegEntries; ++u) {
                if (s < (u + 1u) * G && e > u * G) {
                    const uint32_t p = r + u * G;
                    const uint32_t a = im[u] & 0xFFFFu, b = im[u] >> 16;
                    const bool ok = last ? (is_single(a) || is_single(b)) : (marked(a) || marked(b));
                    if (p >= s && p < e && ok) on |= 1u << u;
                }
            }
            lsync();
#pragma unroll
            for (uint32_t u = 0; u < kRegEntries; ++u)
                if ((on >> u) & 1u) or_wg(l32(bm_a | ((x[u] >> 3) & 0x1FFCu)), 1u << (x[u] & 31u));
            good |= on;
            lsync();
        }
        glb_u8 *out = row();
        uint32_t cur = Q, acc = 0;
        for (int32_t k = 0; k < L; ++k) {
            const uint32_t ic = image(cur);
            const uint32_t a = ic & 0xFFFFu, b = ic >> 16;
            bool take_a;
            if (k == L - 1) {
                take_a = is_single(a);
            } else if (k == 0) {
                take_a = marked(a);  // the bitmap holds level 1's good subsets
            } else {
                // a is taken when a(cur) is a good entry of level k + 1
                const uint32_t s = level_start(w, k + 1), e = level_start(w, k + 2);
                bool mine = false;
#pragma unroll
                for (uint32_t u = 0; u < kRegEntries; ++u) {
                    const uint32_t p = r + u * G;
                    mine |= p >= s && p < e && ((good >> u) & 1u) != 0u && x[u] == a;
                }
                take_a = g.any(mine);
            }
            if (!take_a) acc |= 1u << (k & 7);
            cur = take_a ? a : b;
            if ((k & 7) == 7 || k == L - 1) {
                if (r == 0 && (k >> 3) < P.witness_stride) out[k >> 3] = (uint8_t)acc;
                acc = 0;
            }
        }
        zero_row((L + 7) >> 3);
    }

    // The lexicographically smallest shortest reset word (a before b), the
    // word the reference's back-tracking yields (CPU.cs:262-269).  Backward,
    // level j = L-1 .. 1: a subset of level j is good when an image of it is a
    // good subset of level j + 1 (level L - 1: an image is a singleton).  The
    // LDS bitmap accumulates the good subsets of levels L-1 .. j+1 and is never
    // cleared on the way: an image of a level-j subset lies in a level <= j+1,
    // so of those marks only level j+1's can match it, and level j's flags are
    // all computed before level j's own marks are added.  Forward: letter a
    // whenever a(cur) is a good subset of the next level (for level 1 the
    // marks, where only level 1 can hold a(Q)).  A history still in the LDS ring is
    // searched there; a history in HBM (the long searches) goes through
    // witness_hbm.
    __device__ void witness(int32_t L) {
        constexpr uint32_t U = 4;  // entries per lane per round, loads issued together
        const uint32_t r = g.rank();
        const WitnessSpace w = witness_space(L);
        wave_sync();  // the recorded subsets are visible to every lane
        clear_bitmap();
        if (w.starts_cached)
            for (uint32_t k = r; k <= (uint32_t)L; k += G) *l32(w.starts_a + 4u * k) = lvlp[SSW_IDX(k, kLevelEntries)];
        lsync();
        if (!w.local) {
            witness_hbm(L, w);
            return;
        }
        if constexpr (T == 1) {
            if (wr <= kRegEntries * G) {
                witness_regs(L, w);
                return;
            }
        }
        // a level of at most G entries (most of them) takes one entry per lane
        for (int32_t j = L - 1; j >= 1; --j) {
            const uint32_t s = level_start(w, j), e = level_start(w, j + 1);
            const bool last = j == L - 1;
            if (e - s <= G) {
                const uint32_t p = s + r;
                const uint32_t x = p < e ? hist_at(w, p) : 0u;
                const uint32_t im = image(x);
                const uint32_t a = im & 0xFFFFu, b = im >> 16;
                const bool on = p < e && (last ? (is_single(a) || is_single(b)) : (marked(a) || marked(b)));
                if (p < e) set_flag(w, p, on);
                lsync();
                if (on) or_wg(l32(bm_a | ((x >> 3) & 0x1FFCu)), 1u << (x & 31u));
                lsync();
                continue;
            }
            const bool one_round = e - s <= U * G;
            uint32_t keep[U] = {};
            for (uint32_t base = s; base < e; base += U * G) {
                uint32_t x[U];
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t p = base + u * G + r;
                    x[u] = p < e ? hist_at(w, p) : 0u;
                }
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t p = base + u * G + r;
                    const uint32_t im = image(x[u]);
                    const uint32_t a = im & 0xFFFFu, b = im >> 16;
                    const bool on = p < e && (last ? (is_single(a) || is_single(b)) : (marked(a) || marked(b)));
                    if (p < e) set_flag(w, p, on);
                    keep[u] = on ? x[u] : 0u;
                }
            }
            lsync();
            if (one_round) {
#pragma unroll
                for (uint32_t u = 0; u < U; ++u)
                    if (keep[u]) or_wg(l32(bm_a | ((keep[u] >> 3) & 0x1FFCu)), 1u << (keep[u] & 31u));
            } else {
                for (uint32_t base = s; base < e; base += U * G) {
                    uint32_t x[U], f[U];
#pragma unroll
                    for (uint32_t u = 0; u < U; ++u) {
                        const uint32_t p = base + u * G + r;
                        x[u] = p < e ? hist_at(w, p) : 0u;
                        f[u] = p < e ? flag_at(w, p) : 0u;
                    }
#pragma unroll
                    for (uint32_t u = 0; u < U; ++u)
                        if (f[u]) or_wg(l32(bm_a | ((x[u] >> 3) & 0x1FFCu)), 1u << (x[u] & 31u));
                }
            }
            lsync();
        }
        glb_u8 *out = row();
        uint32_t cur = Q, acc = 0;
        for (int32_t k = 0; k < L; ++k) {
            const uint32_t im = image(cur);
            const uint32_t a = im & 0xFFFFu, b = im >> 16;
            bool take_a;
            if (k == L - 1) {
                take_a = is_single(a);
            } else if (k == 0) {
                take_a = marked(a);  // the bitmap holds level 1's good subsets
            } else {
                // a is taken when a(cur) is a flagged entry of level k + 1
                const uint32_t s = level_start(w, k + 1), e = level_start(w, k + 2);
                bool hit = false;
                if (e - s <= G) {
                    const uint32_t p = s + r;
                    hit = g.any(p < e && flag_at(w, p) != 0u && hist_at(w, p) == a);
                }
                for (uint32_t base = s; base < e && e - s > G && !hit; base += U * G) {
                    uint32_t x[U], f[U];
#pragma unroll
                    for (uint32_t u = 0; u < U; ++u) {
                        const uint32_t p = base + u * G + r;
                        x[u] = p < e ? hist_at(w, p) : 0u;
                        f[u] = p < e ? flag_at(w, p) : 0u;
                    }
                    bool mine = false;
#pragma unroll
                    for (uint32_t u = 0; u < U; ++u) mine |= f[u] != 0u && x[u] == a;
                    hit = g.any(mine);
                }
                take_a = hit;
            }
            if (!take_a) acc |= 1u << (k & 7);
            cur = take_a ? a : b;
            if ((k & 7) == 7 || k == L - 1) {
                if (r == 0 && (k >> 3) < P.witness_stride) out[k >> 3] = (uint8_t)acc;
                acc = 0;
            }
        }
        zero_row((L + 7) >> 3);
    }

    // One level of the backward pass for a history in HBM, a round of U * G
    // entries at a time: good flags into the LDS keep area (in_lds) or into
    // HBM, then this level's marks added.
    __device__ void witness_level_hbm(uint32_t s, uint32_t e, bool last, uint32_t keep_a, bool in_lds) {
        constexpr uint32_t U = 4, RW = U * G;
        const uint32_t r = g.rank();
        for (uint32_t base = s; base < e; base += RW) {
            uint32_t x[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t p = base + u * G + r;
                x[u] = p < e ? (uint32_t)hist[SSW_IDX(p, kHistEntries)] : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t p = base + u * G + r;
                const uint32_t im = image(x[u]);
                const uint32_t a = im & 0xFFFFu, b = im >> 16;
                const bool ga = last ? is_single(a) : marked(a);
                const bool on = p < e && (ga || (last ? is_single(b) : marked(b)));
                if (on && ga) choose(x[u]);
                if (p < e) {
                    if (in_lds)
                        *l16(keep_a + 2u * (p - s)) = (uint16_t)(on ? x[u] : 0u);  // no subset is empty
                    else
                        goodp[SSW_IDX(p, 65536u)] = on ? 1u : 0u;
                }
            }
        }
        if (in_lds)
            lsync();
        else
            wave_sync();  // the level's flags in HBM are visible to every lane
        for (uint32_t base = s; base < e; base += RW) {
            uint32_t x[U];
#pragma unroll
            for (uint32_t u = 0; u < U; ++u) {
                const uint32_t p = base + u * G + r;
                if (in_lds)
                    x[u] = p < e ? (uint32_t)*l16(keep_a + 2u * (p - s)) : 0u;
                else
                    x[u] = (p < e && goodp[SSW_IDX(p, 65536u)]) ? (uint32_t)hist[SSW_IDX(p, kHistEntries)] : 0u;
            }
#pragma unroll
            for (uint32_t u = 0; u < U; ++u)
                if (x[u]) or_wg(l32(bm_a | ((x[u] >> 3) & 0x1FFCu)), 1u << (x[u] & 31u));
        }
        lsync();
    }
    // bit x of the choice bitmap: letter a keeps good subset x on a shortest path
    __device__ __forceinline__ void choose(uint32_t x) const {
        __hip_atomic_fetch_or(choicep + SSW_IDX(x >> 5, P.lds_bitmap_words), 1u << (x & 31u), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_WORKGROUP);
    }

    // The passes for a history in HBM with the good subsets as lists in LDS.
    // Backward, j = L-1 .. 1: the good subsets of level j are appended to the
    // list area (the LDS ring after the cached level starts, which the search
    // no longer needs) in ballot order, followed by their count; level j's
    // marks are then set from its list.  The first round of level j - 1 is
    // loaded before level j's marks are set.  Forward: letter a when a(cur)
    // is in the next level's list (read from the top of the area down: G_1,
    // its count above it, then G_2 ...).  A shortest path's levels hold few
    // good subsets (one each for the Cerny automata), so the lists are short
    // and the passes write nothing to HBM.  Returns false when the lists
    // outgrow the area.
    __device__ bool witness_lists(int32_t L, const WitnessSpace &w) {
        constexpr uint32_t U = 4, RW = U * G;
        const uint32_t r = g.rank();
        const uint32_t starts_bytes = w.starts_cached ? 4u * ((uint32_t)L + 1u) : 0u;
        const uint32_t list_a = ring_a + ((starts_bytes + 3u) & ~3u);
        const uint32_t cap = (2u * (rcap + 64u) - (list_a - ring_a)) / 2u;  // u16 entries
        uint32_t top = 0;  // u16 entries in use (uniform)
        uint32_t x[U];
        bool have = false;  // x holds the first round of the level
        for (int32_t j = L - 1; j >= 1; --j) {
            const uint32_t s = level_start(w, j), e = level_start(w, j + 1);
            const bool last = j == L - 1;
            const uint32_t first = top;
            for (uint32_t base = s; base < e; base += RW) {
                if (!(have && base == s)) {
#pragma unroll
                    for (uint32_t u = 0; u < U; ++u) {
                        const uint32_t p = base + u * G + r;
                        x[u] = p < e ? (uint32_t)hist[SSW_IDX(p, kHistEntries)] : 0u;
                    }
                }
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t p = base + u * G + r;
                    const uint32_t im = image(x[u]);
                    const uint32_t a = im & 0xFFFFu, b = im >> 16;
                    const bool on =
                        p < e && (last ? (is_single(a) | is_single(b)) : (marked(a) | marked(b)));
                    const uint64_t m = __ballot(on);
                    const uint32_t cnt = (uint32_t)__popcll(m);
                    if (top + cnt >= cap) return false;  // no room for them and the count
                    if (on)
                        *l16(list_a + 2u * __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, top))) =
                            (uint16_t)x[u];
                    top += cnt;
                }
            }
            have = false;
            if (j > 1) {
                const uint32_t s2 = level_start(w, j - 1);
#pragma unroll
                for (uint32_t u = 0; u < U; ++u) {
                    const uint32_t p = s2 + u * G + r;
                    x[u] = p < s ? (uint32_t)hist[SSW_IDX(p, kHistEntries)] : 0u;
                }
                have = true;
            }
            if (r == 0) *l16(list_a + 2u * top) = (uint16_t)(top - first);
            ++top;
            lsync();
            for (uint32_t q = first + r; q + 1u < top; q += G) {
                const uint32_t y = *l16(list_a + 2u * q);
                or_wg(l32(bm_a | ((y >> 3) & 0x1FFCu)), 1u << (y & 31u));
            }
            lsync();
        }
        glb_u8 *out = row();
        uint32_t cur = Q, acc = 0, endp = top;
        for (int32_t k = 0; k < L; ++k) {
            const uint32_t im = image(cur);
            const uint32_t a = im & 0xFFFFu, b = im >> 16;
            bool take_a;
            if (k == L - 1) {
                take_a = is_single(a);
            } else {
                // G_{k+1}: its count at endp - 1, its entries below it
                const uint32_t c = uni<G>((uint32_t)*l16(list_a + 2u * (endp - 1u)));
                const uint32_t lo = endp - 1u - c;
                bool mine = false;
                for (uint32_t q = lo + r; q + 1u < endp; q += G) mine |= *l16(list_a + 2u * q) == a;
                take_a = g.any(mine);
                endp = lo;
            }
            if (!take_a) acc |= 1u << (k & 7);
            cur = take_a ? a : b;
            if ((k & 7) == 7 || k == L - 1) {
                if (r == 0 && (k >> 3) < P.witness_stride) out[k >> 3] = (uint8_t)acc;
                acc = 0;
            }
        }
        zero_row((L + 7) >> 3);
        return true;
    }

    // The passes of witness_lists for the long searches' launch (four entries
    // per lane), with less LDS work and fewer dependent LDS round trips per
    // level.  Backward: the history and the images the search recorded for
    // it are read in rounds of four entries per lane (from the level's start
    // rounded down to four entries), kAhead rounds ahead of the one being
    // searched and across level boundaries (the loads do not depend on the
    // marks), so below level L - 1 a subset costs two bitmap reads and no
    // table lookup, and a round's reads are issued together.  Each good
    // subset x is listed as x | ga << 16, ga = 1 when a(x) is good one level
    // further (at level L - 1: a singleton), which is the forward greedy's
    // choice at x.  A level's marks are set once its last round is done
    // (from the round's registers when it had one round, else from its
    // list) without waiting on them: the wave's LDS operations complete in
    // order, so the next level's reads see them.  Level starts come 64 at a
    // time.  Forward: at cur of level k >= 1 the letter is the ga of cur in
    // G_k, found by one read per lane of the list below its count (issued
    // with cur's image lookup), so a letter costs one round trip.  Returns
    // false when the lists outgrow the area.
    __device__ bool witness_long(int32_t L_in, const WitnessSpace &w) {
        static_assert(BT, "the long searches' launch has byte tables");
        // uniform (the compiler cannot tell): the passes' control flow is then scalar
        const int32_t L = (int32_t)__builtin_amdgcn_readfirstlane((uint32_t)L_in);
        constexpr uint32_t RW = 4u * G;  // entries per round
        constexpr int kAhead = 4;        // rounds loaded ahead
        const uint32_t r = g.rank();
        const bool cached = __builtin_amdgcn_readfirstlane(w.starts_cached ? 1u : 0u) != 0u;
        // Level starts, 64 levels per read: a window holds in lane i the
        // start of level top - i (from the LDS copy, or else from HBM), and
        // a start inside it is one lane read.  The two passes over the
        // levels (the loads' cursor and the search) go down with windows of
        // their own, so each level costs no wait on a start.
        struct StartWindow {
            uint32_t v;
            int32_t top;
        };
        const auto window = [&](int32_t top) -> StartWindow {
            const int32_t k = top - (int32_t)r;
            uint32_t v = 0;
            if (cached) {
                v = k >= 0 ? (uint32_t)*l32(w.starts_a + 4u * (uint32_t)max(k, 0)) : 0u;
            } else {
                v = k >= 0 ? (uint32_t)lvlp[SSW_IDX((uint32_t)max(k, 0), kLevelEntries)] : 0u;
            }
            return StartWindow{v, top};
        };
        const auto start_in = [&](StartWindow &sw, int32_t k) -> uint32_t {
            if (k > sw.top || k <= sw.top - 64) sw = window(k);
            return __builtin_amdgcn_readlane(sw.v, sw.top - k);
        };
        StartWindow fwin = window(L), cwin = fwin;
        const auto start = [&](int32_t k) -> uint32_t { return start_in(cwin, k); };
        const uint32_t starts_bytes = cached ? 4u * ((uint32_t)L + 1u) : 0u;
        const uint32_t list_off = (starts_bytes + 3u) & ~3u;
        const uint32_t list_a = ring_a + list_off;
        // from uniform values only (ring_a lives in a VGPR), so that the room
        // test below is a scalar branch
        const uint32_t cap = (2u * (rcap + 64u) - list_off) / 4u;  // u32 entries
        uint32_t top = 0;  // u32 entries in use (uniform)
        // the loads' cursor: level fj (done below 1), its start fs and end
        // fe, the next round's base fb (the history's and the images' slots
        // are 256-byte aligned, so every round's loads are aligned: 8 bytes
        // of subsets and 16 of their images per lane)
        const glb_u64 *h64 = (const glb_u64 *)hist;
        const __attribute__((address_space(1))) u32x4 *i128 = (const __attribute__((address_space(1))) u32x4 *)himg;
        struct Round {
            uint64_t x;  // four subsets
            u32x4 im;    // their images (levels below L - 1: every subset there was expanded)
        };
        int32_t fj = L - 1;
        uint32_t fs = start_in(fwin, fj), fe = start_in(fwin, fj + 1), fb = fs & ~3u;
        const auto fetch = [&]() -> Round {
            // the loads are issued on every call, past the last level too
            // (at the first words): the same number of loads per round, so
            // the waits for a slot count exactly the rounds issued after it
            const uint32_t q = fj >= 1 ? min(fb + 4u * r, kHistEntries - 4u) >> 2 : 0u;
            Round v;
            v.x = h64[SSW_IDX(q, kHistEntries / 4u)];
            v.im = i128[SSW_IDX(q, kHistEntries / 4u)];
            if (fj >= 1) {
                fb += RW;
                if (fb >= fe) {
                    fe = fs;
                    if (--fj >= 1) {
                        fs = start_in(fwin, fj);
                        fb = fs & ~3u;
                    }
                }
            }
            return v;
        };
        // slot d holds the round kAhead rounds after the one it last held
        Round ahead[kAhead];
#pragma unroll
        for (int d = 0; d < kAhead; ++d) ahead[d] = fetch();
        // the searching cursor: level j, its start s and end e, round base,
        // rounds taken in the level, the first list entry of the level
        int32_t j = L - 1;
        uint32_t e = start(L), s = start(L - 1);
        uint32_t base = s & ~3u, rounds = 0, first = 0;
        uint32_t x[4] = {}, ent[4] = {};
        bool on[4] = {};
        while (j >= 1) {
#pragma unroll
            for (int d = 0; d < kAhead; ++d) {
                if (j < 1) break;
                // slot d is refilled once its round's values are taken out
                // of it, so the refill can land in the same registers (no
                // copy of a load in flight at the loop's back edge)
                const Round v = ahead[d];
                // both halves of the slot consumed on every path (level L - 1
                // reads no images): a load left pending on one path would
                // make the slot's refill below wait for every load in flight
                asm volatile("" ::"v"(v.x), "v"(v.im));
                const bool last = j == L - 1;
                uint32_t im[4] = {v.im.x, v.im.y, v.im.z, v.im.w};
                bool ga[4], gb[4];
#pragma unroll
                for (uint32_t u = 0; u < 4u; ++u) x[u] = (uint32_t)(v.x >> (16u * u)) & 0xFFFFu;
                if (last) {
                    // level L - 1 (the search stopped inside it): images from
                    // the tables, good = an image is a singleton
                    uint32_t lo[4], hi[4];
#pragma unroll
                    for (uint32_t u = 0; u < 4u; ++u) {
                        lo[u] = *l32(tab_a | ((x[u] << 2) & 0x3FCu));
                        hi[u] = *l32((tab_a | ((x[u] >> 6) & 0x3FCu)) + 1024u);
                    }
#pragma unroll
                    for (uint32_t u = 0; u < 4u; ++u) {
                        im[u] = lo[u] | hi[u];
                        ga[u] = is_single(im[u] & 0xFFFFu);
                        gb[u] = is_single(im[u] >> 16);
                    }
                } else {
                    // the recorded images' eight bitmap reads issued together
                    // (nothing moved across the barrier: the kernel's register
                    // pressure would otherwise pull each read next to its use)
                    uint32_t wa[4], wb[4];
#pragma unroll
                    for (uint32_t u = 0; u < 4u; ++u) {
                        wa[u] = *l32(bm_a | ((im[u] >> 3) & 0x1FFCu));
                        wb[u] = *l32(bm_a | ((im[u] >> 19) & 0x1FFCu));
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (uint32_t u = 0; u < 4u; ++u) {
                        ga[u] = ((wa[u] >> (im[u] & 31u)) & 1u) != 0u;
                        gb[u] = ((wb[u] >> ((im[u] >> 16) & 31u)) & 1u) != 0u;
                    }
                }
                ahead[d] = fetch();
                uint64_t m[4];
                uint32_t total = 0;
#pragma unroll
                for (uint32_t u = 0; u < 4u; ++u) {
                    const uint32_t p = base + 4u * r + u;
                    on[u] = (p >= s) & (p < e) & (ga[u] | gb[u]);
                    ent[u] = x[u] | (ga[u] ? 0x10000u : 0u);
                    m[u] = __ballot(on[u]);
                    total += (uint32_t)__popcll(m[u]);
                }
                if (top + total >= cap) return false;  // no room for them and the count
#pragma unroll
                for (uint32_t u = 0; u < 4u; ++u) {
                    if (on[u])
                        *l32(list_a + 4u * __builtin_amdgcn_mbcnt_hi((uint32_t)(m[u] >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m[u], top))) =
                            ent[u];
                    top += (uint32_t)__popcll(m[u]);
                }
                ++rounds;
                base += RW;
                if (base >= e) {
                    // level j done: its count, then its marks
                    if (r == 0) *l32(list_a + 4u * top) = top - first;
                    ++top;
                    if (rounds == 1) {
                        // every good subset of the level is in this round's registers
#pragma unroll
                        for (uint32_t u = 0; u < 4u; ++u)
                            if (on[u]) or_wg(l32(bm_a | ((x[u] >> 3) & 0x1FFCu)), 1u << (x[u] & 31u));
                    } else {
                        lsync();
                        for (uint32_t q = first + r; q + 1u < top; q += G) {
                            const uint32_t y = *l32(list_a + 4u * q) & 0xFFFFu;
                            or_wg(l32(bm_a | ((y >> 3) & 0x1FFCu)), 1u << (y & 31u));
                        }
                    }
                    lsync();
                    --j;
                    e = s;
                    if (j >= 1) s = start(j);
                    base = s & ~3u;
                    rounds = 0;
                    first = top;
                }
            }
        }
        glb_u8 *out = row();
        uint32_t cur = Q, acc = 0, endp = top;
        for (int32_t k = 0; k < L; ++k) {
            bool take_a;
            const uint32_t im = image(cur);
            if (k == 0) {
                take_a = L == 1 ? is_single(im & 0xFFFFu) : marked(im & 0xFFFFu);  // level 1's marks
            } else {
                // G_k: its count at endp - 1, its entries below it; lane 0
                // reads the count, lane i the i-th entry down (one read
                // covers lists of up to 63 entries)
                const uint32_t v0 = r < endp ? *l32(list_a + 4u * (endp - 1u - r)) : 0u;
                const uint32_t c = __builtin_amdgcn_readfirstlane(v0);
                uint64_t hit = __ballot(r >= 1u && r <= c && (v0 & 0xFFFFu) == cur);
                uint32_t val = hit ? __builtin_amdgcn_readlane(v0, (int)__ffsll((long long)hit) - 1) : 0u;
                for (uint32_t q0 = 64u; !hit && q0 <= c; q0 += G) {
                    const uint32_t q = q0 + r;
                    const uint32_t vq = q <= c ? *l32(list_a + 4u * (endp - 1u - q)) : 0u;
                    hit = __ballot(q <= c && (vq & 0xFFFFu) == cur);
                    if (hit) val = __builtin_amdgcn_readlane(vq, (int)__ffsll((long long)hit) - 1);
                }
                take_a = (val >> 16) != 0u;
                endp -= c + 1u;
            }
            if (!take_a) acc |= 1u << (k & 7);
            cur = take_a ? (im & 0xFFFFu) : (im >> 16);
            if ((k & 7) == 7 || k == L - 1) {
                if (r == 0 && (k >> 3) < P.witness_stride) out[k >> 3] = (uint8_t)acc;
                acc = 0;
            }
        }
        zero_row((L + 7) >> 3);
        return true;
    }

    // The passes for a history in HBM (a long search: its level chain is the
    // cost, so round trips per level are what is cut).  First with the good
    // subsets as lists (witness_lists); when those outgrow their room, again
    // from a clean bitmap with choice bits.  Backward: the good subsets of a
    // level are kept in the LDS ring area, which the search no longer needs
    // (after the cached level starts), so its marks are set from there (no
    // flags stored in HBM, waited for and read back); a level wider than that
    // room stores its flags in HBM and reads them back.  Each good subset x
    // also gets its letter: bit x of the choice bitmap (the group's HBM slot)
    // is set when a(x) is good one level further (level L - 1: a(x) is a
    // singleton), which is exactly the forward greedy's test at x, since
    // every subset lies in one level.  Forward: the first letter from the
    // marks of level 1, then the choice bits, copied into the LDS bitmap, one
    // LDS read per letter (no search of a level's positions).
    __device__ void witness_hbm(int32_t L, const WitnessSpace &w) {
        if constexpr (T >= 4) {
            if (witness_long(L, w)) return;
        } else {
            if (witness_lists(L, w)) return;
        }
        const uint32_t r = g.rank();
        lsync();
        clear_bitmap();
        lsync();
        const uint32_t starts_bytes = w.starts_cached ? 4u * ((uint32_t)L + 1u) : 0u;
        const uint32_t keep_a = ring_a + ((starts_bytes + 3u) & ~3u);
        const uint32_t keep_cap = (2u * (rcap + 64u) - (keep_a - ring_a)) / 2u;  // u16 entries
        for (int32_t j = L - 1; j >= 1; --j) {
            const uint32_t s = level_start(w, j), e = level_start(w, j + 1);
            witness_level_hbm(s, e, j == L - 1, keep_a, e - s <= keep_cap);
        }
        // the first letter from the marks of level 1, then the choice bits
        // (visible to every lane once the wave's atomics have completed)
        const uint32_t im0 = image(Q);
        const bool first_a = L == 1 ? is_single(im0 & 0xFFFFu) : marked(im0 & 0xFFFFu);
        wave_sync();
        const uint32_t words = P.lds_bitmap_words;
        for (uint32_t k = r; k < words; k += G) {
            // read where the atomics were performed (past the vector L1)
            *l32(bm_a + 4u * k) = __hip_atomic_load(choicep + SSW_IDX(k, P.lds_bitmap_words), __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
            choicep[SSW_IDX(k, P.lds_bitmap_words)] = 0u;  // clean for the next automaton
        }
        lsync();
        glb_u8 *out = row();
        uint32_t cur = Q, acc = 0;
        for (int32_t k = 0; k < L; ++k) {
            const uint32_t im = k == 0 ? im0 : image(cur);
            const bool take_a = k == 0 ? first_a : marked(cur);
            if (!take_a) acc |= 1u << (k & 7);
            cur = take_a ? (im & 0xFFFFu) : (im >> 16);
            if ((k & 7) == 7 || k == L - 1) {
                if (r == 0 && (k >> 3) < P.witness_stride) out[k >> 3] = (uint8_t)acc;
                acc = 0;
            }
        }
        zero_row((L + 7) >> 3);
    }

    __device__ __forceinline__ void finish(int code) {
        if (code == kSpilled) {
            lsync();
            return;
        }
        if (code == kPrefiltered) {
            ++done_local;
            lsync();
            if (done_local >= 256) flush();
            return;
        }
        const bool found = code == kFound;
        const int32_t L = found ? (int32_t)level + 1 : -1;
        const int64_t E = (int64_t)hi + e0;
        if constexpr (WIT) {
            if (found)
                witness(L);
            else
                zero_row(0);
        }
        if (g.rank() == 0) {
            P.length[ai] = L;
            if (P.sync) P.sync[ai] = found ? 1 : 0;
            if (P.expanded) P.expanded[ai] = E;
            const int32_t bound = P.bound >= 0 ? P.bound : (n - 1) * (n - 1);
            if (found && L > bound) atomicMin(P.violation, (unsigned long long)ai);
        }
        ++done_local;
        subsets_local += (uint64_t)E;
        expanded_local += (uint64_t)(found ? found_end : hi) + e0;
        lsync();
        if (done_local >= 256) flush();
    }

    __device__ __forceinline__ void flush() {
        if (g.rank() == 0) {
            const uint32_t slot = blockIdx.x & (kCounterSlots - 1u);
            if (done_local) {
                unsigned long long *pg = P.progress + (size_t)slot * kSlotWords;
                atomicAdd(&pg[0], (unsigned long long)done_local);
                atomicAdd(&pg[1], (unsigned long long)subsets_local);
                atomicAdd(&pg[2], (unsigned long long)done_local);
            }
            if (P.launch_stats && (done_local || expanded_local)) {
                unsigned long long *ls = P.launch_stats + (size_t)slot * kStatWords;
                atomicAdd(&ls[0], (unsigned long long)done_local);
                atomicAdd(&ls[1], (unsigned long long)subsets_local);
                atomicAdd(&ls[2], (unsigned long long)expanded_local);
            }
        }
        done_local = 0;
        subsets_local = 0;
        expanded_local = 0;
    }
};
#ifdef SSW_CHECKED
#undef l32
#undef l16
#undef l8
#undef l128
#endif

template <class Solver, class Grp>
__device__ __forceinline__ void batch_body(const BatchParams &P) {
    extern __shared__ __align__(16) uint8_t smem[];
    constexpr uint32_t G = Grp::kSize;
    const uint32_t groups_per_block = blockDim.x / G;
    const uint32_t gib = threadIdx.x / G;
    const uint64_t slot = (uint64_t)blockIdx.x * groups_per_block + gib;
    Solver S(P, smem, gib, groups_per_block, P.slot_base + slot * P.slot_bytes);
    S.wq = (uint32_t)(slot & (kQueues - 1));
    S.prefetch();
    if constexpr (Solver::kNestedLoop) {
        // uniform control flow: one automaton at a time, one tight step loop
        while (S.has_next()) {
            int code = S.begin();
            if (code < 0) continue;
            while (code == kContinue) code = S.step();
            S.finish(code);
        }
    } else {
        // several groups per wavefront: each group starts its next automaton
        // while the others keep stepping
        bool active = true, need_new = true;
        for (;;) {
            if (active && need_new) {
                if (!S.has_next()) {
                    active = false;
                } else {
                    const int c = S.begin();
                    if (c == kContinue)
                        need_new = false;
                    else if (c > 0)
                        S.finish(c);
                }
            }
            if (G < 64) {
                if (__ballot(active) == 0) break;
            } else if (!active) {
                break;
            }
            if (active && !need_new) {
                const int code = S.step();
                if (code != kContinue) {
                    S.finish(code);
                    need_new = true;
                }
            }
        }
    }
    S.flush();
}

template <class Solver, class Grp>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Solver::kWavesPerEu))) void batch_kernel(
    BatchParams P) {
    batch_body<Solver, Grp>(P);
}

// Scratch of the allocation-free entry point (layout in ssw_internal.h): one
// copy per GPU, held while the process runs.  In host code the name
// `scratch` is a host-side stand-in, not a device address: the device
// address of each GPU's copy comes from hipGetSymbolAddress
// (noalloc_scratch below), and only that address is handed to kernels.
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
static_assert(4 * kQueueBytes <= kNoAllocListOff, "no-alloc work queues overlap the spill lists");

// Clears `words` u32 words (kernel instead of a memset so the call can sit
// inside a captured graph).
__global__ void fill_u32_kernel(uint32_t *p, uint64_t words, uint32_t value) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

}  // namespace ssw

// ---- launch wrappers (host side of this translation unit) -------------------

namespace ssw {

template <class Solver, class Grp>
static hipError_t launch_t(const BatchParams &P, int grid, int block, size_t lds, hipStream_t st) {
    hipLaunchKernelGGL((batch_kernel<Solver, Grp>), dim3(grid), dim3(block), lds, st, P);
    return hipGetLastError();
}

template <class Solver, class Grp>
static hipError_t occ_t(int block, size_t lds, int *blocks) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, batch_kernel<Solver, Grp>, block, lds);
}

// variant: 0 = general (witness, HBM bitmaps); 1..3 = small path with T = 1, 2, 4
// and nibble tables; 4..6 = small path with T = 1, 2, 4 and byte tables; 7 =
// small path with T = 8 and byte tables, one wavefront per automaton; 8, 9 =
// small path with witness words (one wavefront per automaton), T = 1 with
// nibble tables and T = 4 with byte tables; 10 = small path with the
// two-family encoding of n > 16 (one wavefront, T = 1, nibble tables); 11 =
// general with four entries per lane (groups of 64 or 256 lanes); 12 = the
// same with the two-family encoding and the LDS bitmap of 2^19 bits
#define SSW_SMALL_CASES(FN, V, TT, BT, ...)                                                                   \
    case V * 1000 + 16: return FN<SmallSolver<WaveGroup<16>, TT, BT>, WaveGroup<16>>(__VA_ARGS__);            \
    case V * 1000 + 32: return FN<SmallSolver<WaveGroup<32>, TT, BT>, WaveGroup<32>>(__VA_ARGS__);            \
    case V * 1000 + 64: return FN<SmallSolver<WaveGroup<64>, TT, BT>, WaveGroup<64>>(__VA_ARGS__);            \
    case V * 1000 + 128: return FN<SmallSolver<BlockGroup<128>, TT, BT>, BlockGroup<128>>(__VA_ARGS__);       \
    case V * 1000 + 256: return FN<SmallSolver<BlockGroup<256>, TT, BT>, BlockGroup<256>>(__VA_ARGS__);
#define SSW_DISPATCH(FN, ...)                                                              \
    switch (variant * 1000 + group_lanes) {                                                \
        case 16: return FN<GroupSolver<WaveGroup<16>>, WaveGroup<16>>(__VA_ARGS__);         \
        case 32: return FN<GroupSolver<WaveGroup<32>>, WaveGroup<32>>(__VA_ARGS__);         \
        case 64: return FN<GroupSolver<WaveGroup<64>>, WaveGroup<64>>(__VA_ARGS__);         \
        case 256: return FN<GroupSolver<BlockGroup<256>>, BlockGroup<256>>(__VA_ARGS__);    \
        case 11064: return FN<GroupSolver<WaveGroup<64>, 4>, WaveGroup<64>>(__VA_ARGS__);  \
        case 11256: return FN<GroupSolver<BlockGroup<256>, 4>, BlockGroup<256>>(__VA_ARGS__); \
        case 12064: return FN<GroupSolver<WaveGroup<64>, 4, true>, WaveGroup<64>>(__VA_ARGS__); \
        case 12256: return FN<GroupSolver<BlockGroup<256>, 4, true>, BlockGroup<256>>(__VA_ARGS__); \
        SSW_SMALL_CASES(FN, 1, 1, false, __VA_ARGS__)                                      \
        SSW_SMALL_CASES(FN, 2, 2, false, __VA_ARGS__)                                      \
        SSW_SMALL_CASES(FN, 3, 4, false, __VA_ARGS__)                                      \
        SSW_SMALL_CASES(FN, 4, 1, true, __VA_ARGS__)                                       \
        SSW_SMALL_CASES(FN, 5, 2, true, __VA_ARGS__)                                       \
        SSW_SMALL_CASES(FN, 6, 4, true, __VA_ARGS__)                                       \
        case 7064: return FN<SmallSolver<WaveGroup<64>, 8, true>, WaveGroup<64>>(__VA_ARGS__); \
        case 8064: return FN<SmallSolver<WaveGroup<64>, 1, false, true>, WaveGroup<64>>(__VA_ARGS__); \
        case 9064: return FN<SmallSolver<WaveGroup<64>, 4, true, true>, WaveGroup<64>>(__VA_ARGS__); \
        case 10064: return FN<SmallSolver<WaveGroup<64>, 1, false, false, true>, WaveGroup<64>>(__VA_ARGS__); \
        default: return hipErrorInvalidValue;                                              \
    }

hipError_t launch_batch(int variant, int group_lanes, const BatchParams &P, int grid, int block,
                        size_t lds, hipStream_t st) {
    SSW_DISPATCH(launch_t, P, grid, block, lds, st)
}

hipError_t batch_occupancy(int variant, int group_lanes, int block, size_t lds, int *blocks) {
    SSW_DISPATCH(occ_t, block, lds, blocks)
}

size_t batch_scratch_bytes(int group_lanes) {
    return group_lanes > 64 ? (size_t)BlockGroup<256>::kScratchBytes : 0;
}

// LDS bytes of one workgroup of `groups` groups (with the group scratch), and
// the per-group figure the kernel reads from BatchParams.lds_group_bytes.
// variants 4..7 and 9 use byte tables, 1..3 and 8 nibble tables
static bool variant_byte_tab(int variant) { return (variant >= 4 && variant <= 7) || variant == 9; }

size_t batch_group_lds(int variant, uint32_t bitmap_words, uint32_t arena_cap) {
    if (variant == 12) return (size_t)(kFamTableBytes + 4ull * bitmap_words + 4ull * arena_cap);
    return variant == 0 || variant == 11 ? (size_t)(kTableBytes + 4ull * bitmap_words + 4ull * arena_cap)
                        : (size_t)SmallLds::region_bytes(variant_byte_tab(variant), arena_cap, variant == 10);
}

size_t batch_block_lds(int variant, int group_lanes, int groups, uint32_t bitmap_words, uint32_t arena_cap) {
    const size_t scratch = batch_scratch_bytes(group_lanes);
    if (variant == 0 || variant == 11 || variant == 12)
        return batch_group_lds(variant, bitmap_words, arena_cap) * groups + scratch;
    return SmallLds::block_bytes((uint32_t)groups, 4u * bitmap_words,
                                 SmallLds::region_bytes(variant_byte_tab(variant), arena_cap, variant == 10),
                                 (uint32_t)scratch);
}

hipError_t init_queues3(uint32_t *q0, uint32_t *q1, uint32_t *q2, int64_t count, uint32_t empty_mask,
                        hipStream_t st) {
    hipLaunchKernelGGL(init_queues_kernel, dim3(1), dim3(kQueues), 0, st, q0, q1, q2, count, empty_mask);
    return hipGetLastError();
}

hipError_t init_queues(uint32_t *queues, int64_t count, hipStream_t st, bool empty) {
    return init_queues3(queues, nullptr, nullptr, count, empty ? 1u : 0u, st);
}

hipError_t reset_counters(uint32_t *deferred, unsigned long long *violation, unsigned long long *progress,
                          unsigned long long *launch_stats, uint32_t stat_words, hipStream_t st) {
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(256), 0, st, deferred, violation, progress,
                       launch_stats, stat_words);
    return hipGetLastError();
}

hipError_t gather_listed(const uint32_t *queue, const uint32_t *list, uint32_t items, uint32_t count,
                         const int32_t *length, const uint8_t *sync, const int64_t *expanded, unsigned long long *out,
                         hipStream_t st) {
    if (items == 0) return hipSuccess;
    hipLaunchKernelGGL(gather_listed_kernel, dim3((items + 255u) / 256u), dim3(256), 0, st, queue, list, items, count,
                       length, sync, expanded, out);
    return hipGetLastError();
}

// Device address of the allocation-free entry point's scratch on `device`,
// which must be current: looked up once per device (at context set-up) and
// kept, since every GPU holds its own copy of the array at its own address.
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
std::mutex g_noalloc_call_mu[kMaxDevices];
}  // namespace

hipError_t noalloc_scratch(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}

// Calls of the allocation-free entry point on one GPU share its scratch: the
// library serialises them per GPU with this lock (within one process).
std::mutex &noalloc_call_lock(int device) { return g_noalloc_call_mu[device < 0 || device >= kMaxDevices ? 0 : device]; }

// Counts of the sizes n_per[0..count) into hist[0..31] (31: n >= 31), which
// the caller has cleared.
__global__ void count_sizes_kernel(const uint8_t *n_per, int64_t count, unsigned long long *hist) {
    __shared__ uint32_t h[32];
    if (threadIdx.x < 32) h[threadIdx.x] = 0u;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t n = n_per[i];
        atomicAdd(&h[n < 31u ? n : 31u], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 32 && h[threadIdx.x]) atomicAdd(hist + threadIdx.x, (unsigned long long)h[threadIdx.x]);
}

hipError_t count_sizes(const uint8_t *n_per, int64_t count, unsigned long long *hist, hipStream_t st) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3(1), dim3(64), 0, st, (uint32_t *)hist, (uint64_t)64, 0u);
    const int64_t want = (count + 1023) / 1024;
    const int64_t blocks = want < 1 ? 1 : (want > 1024 ? 1024 : want);
    hipLaunchKernelGGL(count_sizes_kernel, dim3((unsigned)blocks), dim3(256), 0, st, n_per, count, hist);
    return hipGetLastError();
}

hipError_t fill_u32(uint32_t *p, uint64_t words, uint32_t value, hipStream_t st) {
    if (words == 0) return hipSuccess;
    uint64_t blocks = (words + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)blocks), dim3(256), 0, st, p, words, value);
    return hipGetLastError();
}

#ifdef SSW_CHECKED
SSW_CHECKED_TAKE(checked_take_batch)
#endif

}  // namespace ssw
