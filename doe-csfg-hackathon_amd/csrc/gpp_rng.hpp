// gpp_rng.hpp -- counter-based generator for the seeded GPP inputs.
//
// Every element is a pure function of (seed, array id, element index,
// component), so the host and device generators produce identical bits and
// any sub-range can be regenerated on its own.  The hash is SplitMix64
// (Steele, Lea & Flood, "Fast splittable pseudorandom number generators",
// OOPSLA 2014); a uniform double is its top 53 bits times 2^-53.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpp {

enum SeededArray : uint32_t {
    kArrAqsm = 1,
    kArrAqsn = 2,
    kArrIeps = 3,
    kArrWtilde = 4,
    kArrVcoul = 5,
    kArrIndex = 6,
};

__host__ __device__ __forceinline__ uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Uniform in [0, 1) for (seed, array, counter).
__host__ __device__ __forceinline__ double seeded_u01(uint64_t seed, uint32_t array_id, uint64_t counter)
{
    uint64_t h = splitmix64(seed ^ splitmix64(((uint64_t)array_id << 56) ^ counter));
    return (double)(h >> 11) * 0x1p-53;
}

// lo + (hi - lo) * u with no FMA contraction, so host and device round alike.
__host__ __device__ __forceinline__ double seeded_uniform(uint64_t seed, uint32_t array_id,
                                                          uint64_t counter, double lo, double hi)
{
#pragma clang fp contract(off)
    double u = seeded_u01(seed, array_id, counter);
    double span = hi - lo;
    double t = span * u;
    return lo + t;
}

}  // namespace gpp
