// ssw_gen.cpp -- problem generators (the SharedDefinitions Problem.* part of
// the contract: GPGPU/SharedDefinitions/Problem.cs:13-58).
#include "ssw.h"

#include <cstdint>
#include <cstdlib>

namespace {

inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// value in [0, n) for (seed, automaton i, state q, letter x): the upper 32
// bits of a splitmix64 hash scaled by n (multiply-shift)
inline int32_t draw(uint64_t seed_key, int64_t i, int q, int x, int n) {
    const uint64_t h = splitmix64(seed_key ^ (((uint64_t)i << 6) | ((uint64_t)q << 1) | (uint64_t)x));
    return (int32_t)(((h >> 32) * (uint64_t)n) >> 32);
}

// .NET Framework System.Random (Knuth's subtractive generator as published
// in Numerical Recipes in C, 2nd ed., with the .NET seeding and sampling).
struct DotnetRandom {
    static constexpr int32_t MBIG = 0x7FFFFFFF;
    static constexpr int32_t MSEED = 161803398;
    int32_t seeds[56];
    int inext = 0, inextp = 21;

    explicit DotnetRandom(int32_t seed) {
        const int32_t subtraction = seed == INT32_MIN ? MBIG : (seed < 0 ? -seed : seed);
        int32_t mj = (int32_t)((uint32_t)MSEED - (uint32_t)subtraction);
        seeds[0] = 0;
        seeds[55] = mj;
        int32_t mk = 1;
        for (int i = 1; i < 55; i++) {
            const int ii = (21 * i) % 55;
            seeds[ii] = mk;
            mk = (int32_t)((uint32_t)mj - (uint32_t)mk);
            if (mk < 0) mk += MBIG;
            mj = seeds[ii];
        }
        for (int k = 1; k < 5; k++)
            for (int i = 1; i < 56; i++) {
                seeds[i] = (int32_t)((uint32_t)seeds[i] - (uint32_t)seeds[1 + (i + 30) % 55]);
                if (seeds[i] < 0) seeds[i] += MBIG;
            }
    }
    int32_t internal_sample() {
        if (++inext >= 56) inext = 1;
        if (++inextp >= 56) inextp = 1;
        int32_t r = (int32_t)((uint32_t)seeds[inext] - (uint32_t)seeds[inextp]);
        if (r == MBIG) r--;
        if (r < 0) r += MBIG;
        seeds[inext] = r;
        return r;
    }
    int32_t next() { return internal_sample(); }
    double sample() { return internal_sample() * (1.0 / MBIG); }
    int32_t next(int32_t lo, int32_t hi) {
        const int64_t range = (int64_t)hi - lo;
        return (int32_t)(sample() * (double)range) + lo;  // range <= INT32_MAX here
    }
};

}  // namespace

extern "C" {

int ssw_generate_uniform(uint64_t seed, int32_t n, int64_t first, int64_t count, int32_t *a, int32_t *b) {
    if (n < 1 || n > SSW_MAX_N || count < 0 || first < 0 || (count && (!a || !b))) return SSW_ERR_ARG;
    const uint64_t key = splitmix64(seed);
    for (int64_t k = 0; k < count; ++k) {
        const int64_t i = first + k;
        for (int q = 0; q < n; ++q) {
            a[k * n + q] = draw(key, i, q, 0, n);
            b[k * n + q] = draw(key, i, q, 1, n);
        }
    }
    return SSW_OK;
}

int ssw_generate_cerny(int32_t n, int32_t convention, int32_t *a, int32_t *b) {
    if (n < 2 || n > SSW_MAX_N || !a || !b || (convention != 0 && convention != 1)) return SSW_ERR_ARG;
    int32_t *merge = convention == 0 ? a : b;  // 0 -> 1, identity elsewhere (Problem.cs:21-23)
    int32_t *cycle = convention == 0 ? b : a;  // q -> q+1 mod n (Problem.cs:24-26)
    for (int q = 0; q < n; ++q) {
        merge[q] = q == 0 ? 1 : q;
        cycle[q] = (q + 1) % n;
    }
    return SSW_OK;
}

int ssw_generate_dotnet(int32_t seed, int32_t n, int64_t count, int32_t *a, int32_t *b) {
    if (n < 1 || n > SSW_MAX_N || count < 0 || (count && (!a || !b))) return SSW_ERR_ARG;
    DotnetRandom outer(seed);                        // Problem.cs:34
    for (int64_t i = 0; i < count; ++i) {
        DotnetRandom r(outer.next());                // Problem.cs:37, :43
        for (int q = 0; q < n; ++q) {                // Problem.cs:51-55
            a[i * n + q] = r.next(0, n);
            b[i * n + q] = r.next(0, n);
        }
    }
    return SSW_OK;
}

}  // extern "C"
