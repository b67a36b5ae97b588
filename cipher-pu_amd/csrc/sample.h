// Counter-based draws shared by the samplers (k_sample.hip) and the fused
// encryption (k_ntt.hip): element k of a call with seed s draws
// splitmix64(s + k * 0x9E3779B97F4A7C15) (Steele, Lea and Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "launch.h"

namespace smp {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ uint64_t draw(uint64_t seed, uint64_t k) { return mix64(seed + k * 0x9E3779B97F4A7C15ull); }

// Uniform residue in [0, q) of element k: floor(u q / 2^64).
__device__ __forceinline__ uint64_t uniform(uint64_t seed, uint64_t k, uint64_t q) {
    return __umul64hi(draw(seed, k), q);
}

// Rounded-Gaussian integer of element k, as (|e|, sign): |e| = the first j
// with u < cdf[j] over the table, the sign from a second draw (none for 0).
__device__ __forceinline__ int gauss(uint64_t seed, uint64_t k, const GaussTable& tab, bool& neg) {
    const uint64_t u = draw(seed, 2 * k), v = draw(seed, 2 * k + 1);
    int j = 0;
    while (j < GAUSS_TABLE - 1 && u >= tab.cdf[j]) ++j;
    neg = (v & 1) && j;
    return j;
}

}  // namespace smp
