// Device-side sampling for RLWE encryption: uniform residues (the mask a)
// and rounded-Gaussian errors (sigma = 3.2, the reference's dgauss of
// scripts/exp_e33_real_ciphertext.py:75-85, a rounded Box-Muller normal).
//
// The generator is counter-based: element k of a call with seed s draws
// splitmix64(s + k * 0x9E3779B97F4A7C15) (Steele, Lea and Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014), so a call is
// reproducible from its seed and independent of the launch shape.  It is a
// simulation generator, not a cryptographic one -- as is the reference's
// numpy / random.Random sampling.
//
// Uniform: floor(u q / 2^64) for a 64-bit u; the bias is below q / 2^64.
// Gaussian: |e| from a cumulative table of the rounded normal over
// |k| <= 47 (integer comparisons only), the sign from a second draw; the
// same integer e is written into every limb of the row.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "launch.h"
#include "modarith.h"
#include "sample.h"

namespace {


// out[row][i][n] uniform in [0, q_i)
__global__ void k_sample_uniform(uint64_t* __restrict__ out, const PrimeConst* __restrict__ pcs, LimbMap map, int l,
                                 int logN, size_t total, uint64_t seed) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int i = (int)((g >> logN) % l);
    out[g] = smp::uniform(seed, g, pcs[map.idx[i]].q);
}

// out[row][i][n] = e[row][n] mod q_i, e rounded-Gaussian
__global__ void k_sample_gauss(uint64_t* __restrict__ out, const PrimeConst* __restrict__ pcs, LimbMap map, int l,
                               int logN, size_t total, uint64_t seed, GaussTable tab) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;   // (row, n)
    if (g >= total) return;
    const size_t N = (size_t)1 << logN;
    const size_t row = g >> logN, n = g & (N - 1);
    bool neg;
    const int k = smp::gauss(seed, g, tab, neg);
    uint64_t* o = out + row * l * N + n;
    for (int i = 0; i < l; ++i) {
        const uint64_t q = pcs[map.idx[i]].q;
        o[(size_t)i * N] = neg ? q - (uint64_t)k : (uint64_t)k;
    }
}

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

hipError_t launch_sample_uniform(const DevTables& t, uint64_t* out, int rows, int l, const LimbMap& map,
                                 uint64_t seed, hipStream_t s) {
    const size_t total = ((size_t)rows * l) << t.logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_sample_uniform, dim3(blocks(total)), dim3(256), 0, s, out, t.pc, map, l, t.logN, total, seed);
    return hipGetLastError();
}

hipError_t launch_sample_gauss(const DevTables& t, uint64_t* out, int rows, int l, const LimbMap& map,
                               uint64_t seed, const GaussTable& tab, hipStream_t s) {
    const size_t total = (size_t)rows << t.logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_sample_gauss, dim3(blocks(total)), dim3(256), 0, s, out, t.pc, map, l, t.logN, total, seed,
                       tab);
    return hipGetLastError();
}
