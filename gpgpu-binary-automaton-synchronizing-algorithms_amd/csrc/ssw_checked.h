// ssw_checked.h -- bounds checks of the checked build.
//
// Compiled with -DSSW_CHECKED (build.py --checked, libssw_checked.so), the
// kernels check the LDS addresses the n <= 16 solver forms and the indexes of
// the HBM buffers (arena, visited bitmaps, recorded history, level starts,
// witness flags) against their sizes.  An access out of bounds is not made:
// it goes to a spare word (LDS) or to element 0 (HBM), and the first such
// access of the code object -- its source line and the offending address or
// index -- is kept in checked_violation, which the library reads after every
// solve and reports as SSW_ERR_INTERNAL.  Without SSW_CHECKED the macros are
// the plain expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#ifdef SSW_CHECKED
namespace ssw {
// bit 63: set; bits 40..62: source line; bits 0..39: address or index
static __device__ unsigned long long checked_violation;

__device__ __forceinline__ void checked_report(uint32_t site, uint64_t value) {
    atomicCAS(&checked_violation, 0ull,
              (1ull << 63) | ((unsigned long long)(site & 0x7FFFFFu) << 40) | (value & ((1ull << 40) - 1)));
}

template <class I>
__device__ __forceinline__ I checked_index(I i, uint64_t limit, uint32_t site) {
    if ((uint64_t)i >= limit) {
        checked_report(site, (uint64_t)i);
        return (I)0;
    }
    return i;
}
}  // namespace ssw
#define SSW_IDX(i, limit) ::ssw::checked_index((i), (uint64_t)(limit), __LINE__)
// The host side of one translation unit: reads and clears its violation word
// on the current device, ordered on stream s.  Solves that run at once on
// one device share the word; the first violation is seen by one of them.
#define SSW_CHECKED_TAKE(fn)                                                                                 \
    hipError_t fn(hipStream_t s, unsigned long long *out) {                                                  \
        static const unsigned long long zero = 0;                                                            \
        hipError_t e = hipMemcpyFromSymbolAsync(out, HIP_SYMBOL(checked_violation), sizeof *out, 0,          \
                                                hipMemcpyDeviceToHost, s);                                   \
        if (e == hipSuccess)                                                                                 \
            e = hipMemcpyToSymbolAsync(HIP_SYMBOL(checked_violation), &zero, sizeof zero, 0,                 \
                                       hipMemcpyHostToDevice, s);                                            \
        if (e == hipSuccess) e = hipStreamSynchronize(s);                                                    \
        return e;                                                                                            \
    }
#else
#define SSW_IDX(i, limit) (i)
#endif
