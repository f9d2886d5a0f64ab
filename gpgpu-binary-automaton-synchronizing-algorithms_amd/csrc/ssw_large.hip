// ssw_large.hip -- device-wide level loop for large automata (n up to 30).
//
// The same search as the batched kernel (SlimCPUUntweaked.cs:89-216: BFS from
// Q over subsets, a before b, first discovery wins, stop at the first
// singleton), laid out for instances whose levels hold up to ~2.4 % of 2^n
// subsets (C_30: ~16 M):
//   * every instance owns a bitmap of 2^n bits in HBM (128 MiB at n = 30) and
//     a frontier ring of u32 subsets that the host grows between graph
//     replays when the next level might not fit;
//   * large_step expands the current level of all running instances in one
//     launch: the workgroups walk a flattened list of 1024-entry tiles, load
//     the tile's byte-sliced image table (4 x 256 u64) into LDS, test-and-set
//     the bitmap with device-scope atomicOr, compact the new subsets in LDS
//     (wave ballots + an LDS counter) and reserve ring space with one
//     atomicAdd per tile;
//   * large_plan (one workgroup) closes the level of every instance, writes
//     finished results and lays out the tiles of the next level, so a
//     captured graph of [large_step, large_plan] pairs runs many levels with
//     no host round trip;
//   * the witness pass reuses the retained levels: large_back marks, level by
//     level from the last, the subsets with a letter into the marked part of
//     the next level; large_fwd then reads off the lexicographically smallest
//     shortest word.
#include "ssw_internal.h"
#include "ssw_checked.h"

namespace ssw {

__device__ __forceinline__ bool lg_single(uint32_t x) { return (x & (x - 1u)) == 0u; }

// u32 words of a bitmap over the 2^n subsets of an n-state automaton
__device__ __forceinline__ uint64_t lg_words(int n) { return n > 5 ? 1ull << (n - 5) : 1ull; }

__device__ __forceinline__ uint32_t or_agent(uint32_t *p, uint32_t v) {
    return __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint64_t lg_image(const uint64_t *tab, uint32_t s) {
    return tab[s & 255u] | tab[256u + ((s >> 8) & 255u)] | tab[512u + ((s >> 16) & 255u)] |
           tab[768u + (s >> 24)];
}

// Builds the tables, seeds the frontier with Q.  One workgroup of 256.
__global__ __launch_bounds__(256) void large_init(LargeInstance *inst, int slot, const uint8_t *delta,
                                                  const uint64_t *offset, int64_t index) {
    LargeInstance &I = inst[slot];
    const int n = I.n;
    const uint8_t *d = delta + (offset ? offset[index] : (uint64_t)index * 2u * (uint64_t)n);
    __shared__ int bad;
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int q = threadIdx.x; q < 2 * n; q += blockDim.x)
        if (d[q] >= n) bad = 1;
    __syncthreads();
    if (bad) {
        if (threadIdx.x == 0) {
            I.status = kLError;
            I.length = -3;
        }
        return;
    }
    for (int e = threadIdx.x; e < 1024; e += blockDim.x) {
        const int k = e >> 8, v = e & 255;
        uint64_t acc = 0;
        for (int t = 0; t < 8; ++t) {
            const int q = 8 * k + t;
            if (q < n && ((v >> t) & 1)) acc |= (1ull << d[q]) | (1ull << (32 + d[n + q]));
        }
        I.table[e] = acc;
    }
    if (threadIdx.x == 0) {
        const uint32_t Q = (uint32_t)((1ull << n) - 1ull);
        I.ring[0] = Q;
        I.bitmap[Q >> 5] |= 1u << (Q & 31u);
        if (I.append) {
            I.lvl[0] = 0;
            I.lvl[1] = 1;
        }
        I.lo = 0;
        I.hi = 1;
        I.wr = 1;
        I.level = 0;
        I.found = 0;
        I.overflow = 0;
        I.length = -1;
        I.expanded = 0;
        I.status = kLRun;
    }
}

__device__ void lg_report(LargeState *st, LargeInstance &I) {
    if (!I.report) return;
    if (st->length) st->length[I.index] = I.length;
    if (st->sync) st->sync[I.index] = I.status == kLFound ? 1 : 0;
    if (st->expanded) st->expanded[I.index] = (int64_t)I.expanded;
    if (I.status == kLFound && st->violation) {
        const int32_t bound = I.bound >= 0 ? I.bound : (I.n - 1) * (I.n - 1);
        if (I.length > bound) atomicMin(st->violation, (unsigned long long)I.index);
    }
    if (st->progress && (I.status == kLFound || I.status == kLNoSync)) {
        atomicAdd(&st->progress[0], 1ull);
        atomicAdd(&st->progress[1], (unsigned long long)I.expanded);
        atomicAdd(&st->progress[2], 1ull);
    }
}

// Closes the level just expanded (advance = 1) and lays out the tiles of the
// next one.  One workgroup.
__global__ __launch_bounds__(256) void large_plan(LargeState *st, LargeInstance *inst, int advance) {
    __shared__ uint64_t tiles[kLargeMaxInstances];
    const int num = st->num;
    for (int i = threadIdx.x; i < kLargeMaxInstances; i += blockDim.x) {
        uint64_t t = 0;
        if (i < num) {
            LargeInstance &I = inst[i];
            if (advance && I.status == kLRun) {
                if (I.overflow) {
                    I.status = kLError;
                    I.length = -4;
                } else if (I.found) {
                    I.status = kLFound;
                    I.length = I.level + 1;
                    I.expanded = I.hi;
                    lg_report(st, I);
                } else {
                    I.lo = I.hi;
                    I.hi = I.wr;
                    I.level += 1;
                    if (I.lo == I.hi) {
                        I.status = kLNoSync;
                        I.length = -1;
                        I.expanded = I.hi;
                        lg_report(st, I);
                    } else if (I.append) {
                        if ((uint32_t)I.level + 1 < I.lvl_cap) I.lvl[I.level + 1] = I.hi;
                    }
                }
            }
            if (I.status == kLRun) {
                const uint64_t width = I.hi - I.lo;
                // the next level holds at most 2 * width subsets, and the
                // ring never holds more than the 2^n distinct subsets
                const uint64_t bound = I.append ? I.hi + 2 * width : 3 * width;
                const uint64_t all = 1ull << I.n;
                const uint64_t need = bound < all ? bound : all;
                if (need > I.ring_mask + 1 || (I.append && (uint32_t)I.level + 2 >= I.lvl_cap))
                    I.status = kLPause;
                else
                    t = (width + kLargeTile - 1) / kLargeTile;
            }
        }
        tiles[i] = t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0;
        uint64_t base = 0;
        for (int i = 0; i < num; ++i) {
            if (tiles[i]) {
                st->act[a] = i;
                st->act_base[a] = base;
                base += tiles[i];
                ++a;
            }
        }
        st->act_base[a] = base;
        st->active = a;
        st->total_tiles = base;
    }
}

// Expands the current level of every running instance.
__global__ __launch_bounds__(256) void large_step(LargeState *st, LargeInstance *inst) {
    __shared__ uint64_t tab[1024];
    __shared__ uint32_t outbuf[2 * kLargeTile];
    __shared__ uint32_t s_count;
    __shared__ unsigned long long s_base;
    const uint64_t total = st->total_tiles;
    const int na = st->active;
    const uint32_t tid = threadIdx.x, lane = __lane_id();
    const uint64_t below = (1ull << lane) - 1ull;
    int cur = -1;
    for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {
        // instance owning tile t
        int lo = 0, hi = na;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (st->act_base[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
        const int i = st->act[lo];
        const uint64_t tile = t - st->act_base[lo];
        LargeInstance &I = inst[i];
        if (i != cur) {
            __syncthreads();
            for (int e = tid; e < 1024; e += blockDim.x) tab[e] = I.table[e];
            cur = i;
        }
        if (tid == 0) s_count = 0;
        __syncthreads();
        const uint64_t lev_lo = I.lo, lev_hi = I.hi, mask = I.ring_mask;
        const uint32_t *ring = I.ring;
        uint32_t *bitmap = I.bitmap;
        const uint32_t Q = (uint32_t)((1ull << I.n) - 1ull);
        const uint64_t e0 = lev_lo + tile * kLargeTile;
        bool single = false;
#pragma unroll
        for (int j = 0; j < kLargeTile / 256; ++j) {
            const uint64_t e = e0 + (uint64_t)j * 256u + tid;
            const bool valid = e < lev_hi;
            const uint32_t s = valid ? ring[e & mask] : 0u;
            const uint64_t im = lg_image(tab, s);
            const uint32_t a = (uint32_t)im, b = (uint32_t)(im >> 32);
            const bool sa = valid && lg_single(a) && a != Q;
            const bool sb = valid && lg_single(b) && b != Q;
            single |= sa | sb;
            bool na = false, nb = false;
            if (valid && !sa) {
                const uint32_t bit = 1u << (a & 31u);
                na = (or_agent(bitmap + SSW_IDX(a >> 5, lg_words(I.n)), bit) & bit) == 0u;
            }
            if (valid && !sb) {
                const uint32_t bit = 1u << (b & 31u);
                nb = (or_agent(bitmap + SSW_IDX(b >> 5, lg_words(I.n)), bit) & bit) == 0u;
            }
            const uint64_t ba = __ballot(na), bb = __ballot(nb);
            const uint32_t wtot = __popcll(ba) + __popcll(bb);
            uint32_t wbase = 0;
            if (lane == 0 && wtot) wbase = atomicAdd(&s_count, wtot);
            wbase = __shfl(wbase, 0, 64);
            uint32_t k = wbase + __popcll(ba & below) + __popcll(bb & below);
            if (na) outbuf[SSW_IDX(k++, 2u * kLargeTile)] = a;
            if (nb) outbuf[SSW_IDX(k, 2u * kLargeTile)] = b;
        }
        if (__ballot(single) != 0 && lane == 0) atomicOr(&I.found, 1);
        __syncthreads();
        const uint32_t cnt = s_count;
        if (tid == 0) {
            unsigned long long base = ~0ull;
            if (cnt) {
                base = atomicAdd(&I.wr, (unsigned long long)cnt);
                const uint64_t cap = mask + 1;
                const bool over = I.append ? (base + cnt > cap) : (base + cnt - lev_lo > cap);
                if (over) {
                    atomicOr(&I.overflow, 1);
                    base = ~0ull;
                }
            }
            s_base = base;
        }
        __syncthreads();
        const unsigned long long base = s_base;
        if (base != ~0ull)
            for (uint32_t k = tid; k < cnt; k += blockDim.x) I.ring[(base + k) & mask] = outbuf[k];
        __syncthreads();
    }
}

// Copies the live part of a ring into a larger one (same logical positions).
__global__ void large_regrow(const uint32_t *src, uint64_t src_mask, uint32_t *dst, uint64_t dst_mask,
                             uint64_t from, uint64_t to) {
    for (uint64_t p = from + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < to;
         p += (uint64_t)gridDim.x * blockDim.x)
        dst[p & dst_mask] = src[p & src_mask];
}

// Witness, backward pass over level j: marks in dst the subsets x of level j
// with a letter into the marked part of level j+1 (j = L-1: into a singleton).
__global__ __launch_bounds__(256) void large_back(const LargeInstance *inst, int slot, int j, int L,
                                                  const uint32_t *src, uint32_t *dst) {
    __shared__ uint64_t tab[1024];
    const LargeInstance &I = inst[slot];
    for (int e = threadIdx.x; e < 1024; e += blockDim.x) tab[e] = I.table[e];
    __syncthreads();
    const uint64_t s = I.lvl[SSW_IDX(j, I.lvl_cap)], e = I.lvl[SSW_IDX(j + 1, I.lvl_cap)];
    for (uint64_t p = s + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < e;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t x = I.ring[p & I.ring_mask];
        const uint64_t im = lg_image(tab, x);
        const uint32_t a = (uint32_t)im, b = (uint32_t)(im >> 32);
        const uint64_t words = lg_words(I.n);
        const bool on = (j == L - 1) ? (lg_single(a) || lg_single(b))
                                     : (((src[SSW_IDX(a >> 5, words)] >> (a & 31u)) & 1u) ||
                                        ((src[SSW_IDX(b >> 5, words)] >> (b & 31u)) & 1u));
        if (on) or_agent(dst + SSW_IDX(x >> 5, words), 1u << (x & 31u));
    }
}

// Witness, forward pass: from Q take letter a whenever its image is a marked
// subset of the next level.  One workgroup.
__global__ __launch_bounds__(256) void large_fwd(const LargeInstance *inst, int slot, int L, const uint32_t *g0,
                                                 const uint32_t *g1, uint8_t *out, int stride) {
    __shared__ uint64_t tab[1024];
    __shared__ int hit;
    const LargeInstance &I = inst[slot];
    for (int e = threadIdx.x; e < 1024; e += blockDim.x) tab[e] = I.table[e];
    __syncthreads();
    uint32_t cur = (uint32_t)((1ull << I.n) - 1ull), acc = 0;
    for (int k = 0; k < L; ++k) {
        const uint64_t im = lg_image(tab, cur);
        const uint32_t a = (uint32_t)im, b = (uint32_t)(im >> 32);
        bool take_a;
        if (k == L - 1) {
            take_a = lg_single(a);
        } else {
            const uint32_t *gs = ((k + 1) & 1) ? g1 : g0;
            take_a = false;
            if ((gs[SSW_IDX(a >> 5, lg_words(I.n))] >> (a & 31u)) & 1u) {
                if (threadIdx.x == 0) hit = 0;
                __syncthreads();
                const uint64_t s = I.lvl[SSW_IDX(k + 1, I.lvl_cap)], e = I.lvl[SSW_IDX(k + 2, I.lvl_cap)];
                for (uint64_t base = s; base < e; base += blockDim.x) {
                    const uint64_t p = base + threadIdx.x;
                    if (p < e && I.ring[p & I.ring_mask] == a) hit = 1;
                    __syncthreads();
                    const int h = hit;
                    __syncthreads();
                    if (h) break;
                }
                take_a = hit != 0;
                __syncthreads();  // every thread has read `hit` before it is reset
            }
        }
        if (!take_a) acc |= 1u << (k & 7);
        cur = take_a ? a : b;
        if ((k & 7) == 7 || k == L - 1) {
            if (threadIdx.x == 0 && (k >> 3) < stride) out[k >> 3] = (uint8_t)acc;
            acc = 0;
        }
    }
    for (int k = ((L + 7) >> 3) + threadIdx.x; k < stride; k += blockDim.x) out[k] = 0;
}

// ---- launch wrappers -------------------------------------------------------

hipError_t large_launch_init(LargeInstance *inst, int slot, const uint8_t *delta, const uint64_t *offset,
                             int64_t index, hipStream_t st) {
    hipLaunchKernelGGL(large_init, dim3(1), dim3(256), 0, st, inst, slot, delta, offset, index);
    return hipGetLastError();
}

hipError_t large_launch_plan(LargeState *s, LargeInstance *inst, int advance, hipStream_t st) {
    hipLaunchKernelGGL(large_plan, dim3(1), dim3(256), 0, st, s, inst, advance);
    return hipGetLastError();
}

hipError_t large_launch_step(LargeState *s, LargeInstance *inst, int grid, hipStream_t st) {
    hipLaunchKernelGGL(large_step, dim3(grid), dim3(256), 0, st, s, inst);
    return hipGetLastError();
}

hipError_t large_launch_regrow(const uint32_t *src, uint64_t src_mask, uint32_t *dst, uint64_t dst_mask,
                               uint64_t from, uint64_t to, hipStream_t st) {
    if (to <= from) return hipSuccess;
    uint64_t blocks = (to - from + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(large_regrow, dim3((unsigned)blocks), dim3(256), 0, st, src, src_mask, dst, dst_mask,
                       from, to);
    return hipGetLastError();
}

hipError_t large_launch_back(const LargeInstance *inst, int slot, int j, int L, const uint32_t *src,
                             uint32_t *dst, int grid, hipStream_t st) {
    hipLaunchKernelGGL(large_back, dim3(grid), dim3(256), 0, st, inst, slot, j, L, src, dst);
    return hipGetLastError();
}

hipError_t large_launch_fwd(const LargeInstance *inst, int slot, int L, const uint32_t *g0, const uint32_t *g1,
                            uint8_t *out, int stride, hipStream_t st) {
    hipLaunchKernelGGL(large_fwd, dim3(1), dim3(256), 0, st, inst, slot, L, g0, g1, out, stride);
    return hipGetLastError();
}

#ifdef SSW_CHECKED
SSW_CHECKED_TAKE(checked_take_large)
#endif

}  // namespace ssw
