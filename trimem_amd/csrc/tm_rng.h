// tm_rng.h -- counter-based random numbers shared by the flip sweeps and the
// device HMC step.  Every draw is a pure function of (seed, stream, index),
// so results depend on neither launch shape nor thread scheduling.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tmh {

// SplitMix64 finaliser (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014)
__device__ __forceinline__ uint64_t splitmix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// streams: 0 flip proposals, 1 flip uniforms, 2 and 3 HMC momenta (the two
// uniforms of a Box-Muller pair), 4 HMC acceptance uniform
__device__ __forceinline__ uint64_t rng_draw(uint64_t seed, uint64_t stream, uint64_t i)
{
    return splitmix64(splitmix64(seed ^ (stream * 0xd1b54a32d192ed03ull)) + i);
}

// uniform on [0, 1) with 53 random bits
__device__ __forceinline__ double rng_unit(uint64_t bits) { return (double)(bits >> 11) * (1.0 / 9007199254740992.0); }

}    // namespace tmh
