// Device PCG32 generator, one stream per history.
// Generator: PCG XSH-RR 64/32 as in the reference RandomState::pcg32
// (dxmcrandom.hpp:156-165): s' = s * 6364136223846793005 + (inc | 1), output
// rotr32(((s >> 18) ^ s) >> 27, s >> 59) of the old state.
// Stream of history h under a run seed: inc = 2h+1, s0 = splitmix64(seed ^ splitmix64(h)).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dxmc_dev {

__host__ __device__ inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

struct Rng {
    uint64_t s;
    uint64_t inc; // already or'ed with 1

    __device__ static Rng forHistory(uint64_t seed, uint64_t h)
    {
        Rng r;
        r.s = splitmix64(seed ^ splitmix64(h));
        r.inc = (2 * h + 1) | 1ULL;
        return r;
    }
    __device__ inline uint32_t next()
    {
        const uint64_t old = s;
        s = old * 6364136223846793005ULL + inc;
        const uint32_t xs = static_cast<uint32_t>(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = static_cast<uint32_t>(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    // u32 * 2^-32 (the float conversion rounds, so 1.0f can occur)
    template <typename T>
    __device__ inline T uniform()
    {
        return static_cast<T>(next()) * static_cast<T>(2.32830643653869628906e-010);
    }
    template <typename T>
    __device__ inline T uniform(T max) { return uniform<T>() * max; }
    template <typename T>
    __device__ inline T uniform(T min, T max)
    {
        const T r = uniform<T>();
        return min + r * (max - min);
    }
    // integer in [0, n) by rejection against the 32-bit truncation of
    // 2^64 mod n (threshold precomputed by the host)
    __device__ inline uint32_t index(uint32_t n, uint32_t threshold)
    {
        for (;;) {
            const uint32_t r = next();
            if (r >= threshold)
                return r % n;
        }
    }
};
}
