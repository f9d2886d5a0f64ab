// ssw_probe.hip -- test-and-set throughput ceilings of the device, for the
// measurement report (no counterpart in the reference): the searches' visited
// sets are bitmaps tested and set with atomicOr-with-return, in LDS for n <= 16
// (an 8 KiB bitmap per automaton) and in HBM for the device-wide path (up to
// 128 MiB per automaton).  These kernels issue nothing but such test-and-sets,
// at uniformly random words and bits, with every wavefront of the device busy,
// so their rate bounds what the solvers can reach.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "ssw.h"
#include "ssw_host.h"

namespace {

constexpr int kProbeThreads = 256;
constexpr uint32_t kLdsWords = 2048;  // 8 KiB per wavefront

__device__ __forceinline__ uint32_t xorshift(uint32_t x) {
    x ^= x << 13;
    x ^= x >> 17;
    x ^= x << 5;
    return x;
}

// Each wavefront owns 8 KiB of LDS and tests-and-sets random bits in it.  A
// lane draws 8 random (word, bit) pairs once; iteration k XORs the word
// indices with one wave-uniform value, which maps a random pattern to another
// random one with the same bank conflicts, so an atomic costs two VALU
// instructions (its address and the use of its result) and the rate is the
// LDS's own.
__global__ __launch_bounds__(kProbeThreads) void lds_probe_kernel(uint32_t iters, uint32_t *sink) {
    __shared__ uint32_t bm[kProbeThreads / 64][kLdsWords];
    uint32_t *flat = &bm[0][0];
    for (uint32_t i = threadIdx.x; i < (kProbeThreads / 64) * kLdsWords; i += kProbeThreads) flat[i] = 0u;
    __syncthreads();
    uint32_t *mine = bm[threadIdx.x >> 6];
    uint32_t x = (blockIdx.x * kProbeThreads + threadIdx.x) * 2654435761u + 0x9E3779B9u;
    uint32_t word[8], bit[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        x = xorshift(x);
        word[j] = x & (kLdsWords - 1u);
        bit[j] = 1u << (x >> 27);
    }
    uint32_t acc = 0;
    for (uint32_t k = 0; k < iters; k += 8) {
        const uint32_t u = ((k >> 3) * 2654435761u) >> 21;  // 11 random bits, uniform
#pragma unroll
        for (int j = 0; j < 8; ++j)
            acc ^= __hip_atomic_fetch_or(mine + (word[j] ^ u), bit[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    if (acc == 0x5EED5EEDu) sink[0] = acc;  // keeps the returned values live
}

// device-scope test-and-sets at random words of an HBM bitmap
__global__ __launch_bounds__(kProbeThreads) void hbm_probe_kernel(uint32_t *bitmap, uint32_t word_mask,
                                                                  uint32_t iters, uint32_t *sink) {
    uint32_t x = (blockIdx.x * kProbeThreads + threadIdx.x) * 2654435761u + 0x7F4A7C15u;
    uint32_t acc = 0;
    for (uint32_t k = 0; k < iters; ++k) {
        x = xorshift(x);
        acc += __hip_atomic_fetch_or(bitmap + (x & word_mask), 1u << (x >> 27), __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
    }
    if (acc == 0x5EED5EEDu) sink[0] = acc;
}

}  // namespace

// Times `iters` test-and-sets per thread over `blocks` workgroups; returns
// operations per second.
int ssw_context::measure_atomics(int32_t kind, double *ops_per_s) {
    if (!ops_per_s || (kind != SSW_ATOMICS_LDS && kind != SSW_ATOMICS_HBM)) return fail(SSW_ERR_ARG, "kind %d", kind);
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    const int blocks = cus * 8;
    const uint32_t iters = kind == SSW_ATOMICS_LDS ? 4096u : 256u;
    constexpr uint64_t kHbmBytes = 128ull << 20;  // the n = 30 visited bitmap
    ssw::DevBuf buf;
    int rc = ensure(buf, kind == SSW_ATOMICS_HBM ? kHbmBytes : 256, "probe buffer");
    if (rc) return rc;
    uint32_t *p = (uint32_t *)buf.p;
    e = hipMemsetAsync(p, 0, buf.bytes, stream);
    float ms = 0.f;
    for (int pass = 0; pass < 2 && e == hipSuccess; ++pass) {  // pass 0 warms up
        e = hipEventRecord(ev0, stream);
        if (e != hipSuccess) break;
        if (kind == SSW_ATOMICS_LDS)
            hipLaunchKernelGGL(lds_probe_kernel, dim3(blocks), dim3(kProbeThreads), 0, stream, iters, p);
        else
            hipLaunchKernelGGL(hbm_probe_kernel, dim3(blocks), dim3(kProbeThreads), 0, stream, p,
                               (uint32_t)(kHbmBytes / 4 - 1), iters, p);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(ev1, stream);
        if (e == hipSuccess) e = hipEventSynchronize(ev1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev0, ev1);
    }
    (void)hipFree(buf.p);
    if (e != hipSuccess) return hip_fail(e, "atomic probe");
    *ops_per_s = (double)blocks * kProbeThreads * iters / (ms * 1e-3);
    return SSW_OK;
}

extern "C" int ssw_measure_atomics(ssw_context *ctx, int32_t kind, double *ops_per_s) {
    if (!ctx) return SSW_ERR_ARG;
    std::lock_guard<std::mutex> lk(ctx->mu);
    return ctx->measure_atomics(kind, ops_per_s);
}
