// rt_pcg.h -- PCG32 (O'Neill, "PCG: A Family of Simple Fast Space-Efficient
// Statistically Good Algorithms for Random Number Generation", 2014; the
// pcg32 XSH-RR 64/32 generator with pcg32_srandom seeding), usable from host
// C++ and HIP device code.  The reference draws from a shared, unseeded
// std::mt19937 (sampling/Sampler.cu:139-145, 210-214); this project needs
// reproducible per-pixel streams instead.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define RT_HD __host__ __device__ __forceinline__
#else
#define RT_HD inline
#endif

namespace rt {

struct Pcg32 {
    uint64_t state;
    uint64_t inc;
};

RT_HD uint32_t pcg32_next(Pcg32 &g)
{
    uint64_t old = g.state;
    g.state = old * 6364136223846793005ULL + g.inc;
    uint32_t xorshifted = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

RT_HD Pcg32 pcg32_seed(uint64_t initstate, uint64_t initseq)
{
    Pcg32 g;
    g.state = 0u;
    g.inc = (initseq << 1u) | 1u;
    pcg32_next(g);
    g.state += initstate;
    pcg32_next(g);
    return g;
}

// Uniform float in [0, 1): the top 24 bits scaled by 2^-24 (exact).
RT_HD float pcg32_float(Pcg32 &g) { return (float)(pcg32_next(g) >> 8) * (1.0f / 16777216.0f); }

// Uniform double in [0, 1) from two draws (53 bits).
RT_HD double pcg32_double(Pcg32 &g)
{
    uint64_t hi = pcg32_next(g) >> 5, lo = pcg32_next(g) >> 6;
    return (double)(hi * 67108864ULL + lo) * (1.0 / 9007199254740992.0);
}

}  // namespace rt
