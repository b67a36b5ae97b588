// Measurement probe for the matrix cores: one 64-point stage of the 45-bit
// transforms (the inner step of a four-step N = 64 x 128 NTT) on the integer
// VALU and on the int8 matrix cores.  Both kernels compute y = G^R x mod q
// for a batch of V length-64 vectors stored element-major ([64][V] words),
// where G is the 64-point negacyclic forward transform as the radix-2 stages
// compute it (natural order in, bit-reversed order out).
//
//  * k_dft64_radix2: one vector per thread, the six radix-2 stages in
//    registers with the library's lazy Shoup butterflies and scalar twiddle
//    loads.  No data exchange at all: the VALU path's best case.
//  * k_dft64_mfma: the exact integer product G x through
//    v_mfma_i32_16x16x64_i8 (a 16 x 16 x 64 int8 product per instruction,
//    int32 accumulators).  A lane keeps each value as a word z below 2^48
//    with z = x + BIAS (mod q), BIAS = 128 (1 + 2^8 + ... + 2^40); its six
//    bytes u_s, fed as signed digits u_s - 128 (an XOR of 0x80), then sum
//    to exactly z - BIAS.  G's entries are split into balanced signed
//    digits.  The 36 digit-plane products accumulate per weight 2^(8k),
//    k = s + t, into 11 int32 sums A_k, recombined exactly (pairs merged in
//    32 bits, two 64-bit words L and M with G x = L + 2^40 M) and reduced by
//    one lazy Shoup product by 2^40; BIAS for the next product rides in the
//    same product's constant.  The result is below 2^48 and feeds the next
//    byte split with no reduction.
// Operand layout: lane l of a wave holds, for column (vector) l & 15, the 16
// values of data index pi(l >> 4, j), j = 0..15.  The hardware pairs byte j
// of the A operand of K group g with byte j of the B operand of the same
// group, so any bijection pi gives the right contraction as long as G's
// digits are laid out with the same pi; pi is chosen equal to the rows the
// accumulator layout leaves in the lane (tile rt = j / 4, register e = j % 4),
// so a product's output is the next product's input with no exchange.  CM
// selects between the two candidate accumulator layouts (row 4 g + e, or
// g + 4 e, within the 16-row tile); only the first gives the radix-2 path's
// words.  PIPE issues row tile rt + 1's products between the instructions of
// tile rt's recombination, so that one wave keeps the matrix core and the
// VALU busy together.  No counterpart in the reference: it settles whether
// a matrix-core formulation of a radix stage beats the VALU path.
#include "launch.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));

template <int CM>
__device__ __forceinline__ int pi_index(int g, int j) {
    return CM == 0 ? 16 * (j >> 2) + 4 * g + (j & 3) : 16 * (j >> 2) + g + 4 * (j & 3);
}

// tw[m + i] = (psi^bitrev6(m + i), floor(. * 2^63 / q)), m + i in 1..63.
// rep_stride is 0: it only keeps the compiler from hoisting all 63 twiddle
// pairs out of the repetition loop (252 scalar registers, spilled), so each
// repetition loads its twiddles as one transform does.
__global__ void __launch_bounds__(256) k_dft64_radix2(const uint64_t* __restrict__ X, uint64_t* __restrict__ Y,
                                                      const ulonglong2* __restrict__ tw, uint64_t q, size_t V,
                                                      int reps, int rep_stride) {
    const size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    const uint64_t nq = 0 - q, four_q = 4 * q;
    uint64_t x[64];
#pragma unroll
    for (int k = 0; k < 64; ++k) x[k] = X[(size_t)k * V + v];
    // lazy bounds: from x < q each rep adds 24 q (six stages of 4 q); reps
    // are capped by the launcher so every value stays below 2^62
    for (int r = 0; r < reps; ++r) {
        const ulonglong2* twr = tw + r * rep_stride;
#pragma unroll
        for (int m = 1; m < 64; m <<= 1) {
            const int t = 32 / m;
#pragma unroll
            for (int i = 0; i < m; ++i) {
                const ulonglong2 w = twr[m + i];
#pragma unroll
                for (int j = 2 * i * t; j < 2 * i * t + t; ++j) ct_bfly_lazy(x[j], x[j + t], w.x, w.y, nq, four_q);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 64; ++i) Y[(size_t)i * V + v] = x[i] % q;
}

// z = sum_k a_k 2^(8k) + BIAS (mod q), returned in [0, 4q + 2^40).  Bounds
// (the host checks them for the digits it builds): |a_k| < 2^23, the merged
// pairs below 2^31, |L| < 2^55, |M| < 2^59; kq = K q + (BIAS / 2^40 mod q)
// keeps M + floor(L / 2^40) + kq in (0, 2^62).
// Stage 1: the two exact words; stage 2: the lift and the Shoup product.
__device__ __forceinline__ void recombine1(const v4i (&acc)[11], int e, uint64_t& L, uint64_t& M) {
    auto u = [&](int k) { return (uint32_t)acc[k][e]; };
    auto sx = [](uint32_t w) { return (uint64_t)(int64_t)(int32_t)w; };
    const uint32_t p0 = u(0) + (u(1) << 8), p1 = u(2) + (u(3) << 8);
    L = sx(p0) + (sx(p1) << 16) + ((uint64_t)u(4) << 32);
    const uint32_t q0 = u(5) + (u(6) << 8), q1 = u(7) + (u(8) << 8), q2 = u(9) + (u(10) << 8);
    M = sx(q0) + (sx(q1) << 16) + ((uint64_t)q2 << 32);
}

__device__ __forceinline__ uint64_t recombine2(uint64_t L, uint64_t M, uint64_t kq, uint64_t ws40, uint64_t nq) {
    const uint64_t lo = L & ((1ull << 40) - 1);
    const uint64_t mp = M + (uint64_t)((int64_t)L >> 40) + kq;
    return mul_shoup4_nq(mp, 1ull << 40, ws40, nq) + lo;   // [0, 4q) + [0, 2^40)
}

__device__ __forceinline__ uint64_t recombine(const v4i (&acc)[11], int e, uint64_t kq, uint64_t ws40,
                                              uint64_t nq) {
    uint64_t L, M;
    recombine1(acc, e, L, M);
    return recombine2(L, M, kq, ws40, nq);
}

// Byte planes of four words below 2^48 (a 4 x 6 byte transpose in 12
// permutes): plane s holds byte s of w0..w3 in bytes 0..3, each XORed with
// 0x80 (the signed digit u - 128).
__device__ __forceinline__ void byte_planes(uint64_t w0, uint64_t w1, uint64_t w2, uint64_t w3, int (&P)[6]) {
    const uint32_t l0 = (uint32_t)w0, l1 = (uint32_t)w1, l2 = (uint32_t)w2, l3 = (uint32_t)w3;
    const uint32_t h0 = (uint32_t)(w0 >> 32), h1 = (uint32_t)(w1 >> 32), h2 = (uint32_t)(w2 >> 32),
                   h3 = (uint32_t)(w3 >> 32);
    // perm(a, b, sel): byte k of the result is byte sel_k of (a << 32 | b)
    const uint32_t a01 = __builtin_amdgcn_perm(l1, l0, 0x05010400u), b01 = __builtin_amdgcn_perm(l1, l0, 0x07030602u);
    const uint32_t a23 = __builtin_amdgcn_perm(l3, l2, 0x05010400u), b23 = __builtin_amdgcn_perm(l3, l2, 0x07030602u);
    const uint32_t c01 = __builtin_amdgcn_perm(h1, h0, 0x05010400u), c23 = __builtin_amdgcn_perm(h3, h2, 0x05010400u);
    P[0] = (int)(__builtin_amdgcn_perm(a23, a01, 0x05040100u) ^ 0x80808080u);
    P[1] = (int)(__builtin_amdgcn_perm(a23, a01, 0x07060302u) ^ 0x80808080u);
    P[2] = (int)(__builtin_amdgcn_perm(b23, b01, 0x05040100u) ^ 0x80808080u);
    P[3] = (int)(__builtin_amdgcn_perm(b23, b01, 0x07060302u) ^ 0x80808080u);
    P[4] = (int)(__builtin_amdgcn_perm(c23, c01, 0x05040100u) ^ 0x80808080u);
    P[5] = (int)(__builtin_amdgcn_perm(c23, c01, 0x07060302u) ^ 0x80808080u);
}

// The 36 digit-plane products of row tile rt: acc[k] = sum over s + t = k.
__device__ __forceinline__ void tile_product(const v4i* sA, int rt, int lane, const v4i (&B)[6], v4i (&acc)[11]) {
    v4i A[6];
#pragma unroll
    for (int t = 0; t < 6; ++t) A[t] = sA[(rt * 6 + t) * 64 + lane];
#pragma unroll
    for (int k = 0; k < 11; ++k) acc[k] = (v4i){0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < 6; ++t)
#pragma unroll
        for (int s = 0; s < 6; ++s) acc[s + t] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[t], B[s], acc[s + t], 0, 0, 0);
}

// One wave per 16 vectors, 4 waves a workgroup; V is a multiple of 64.
// Fpk[rt][t][lane] = digit t of G[16 rt + (lane & 15)][pi(lane >> 4, j)] at
// byte j, staged in LDS (24 KiB) and read per row tile; rep_stride is 0 and
// only keeps the compiler from hoisting those reads out of the repetition
// loop (96 registers), which would leave one wave per SIMD and the
// accumulators in AGPRs.  cst = {kq, (-BIAS) mod q, BIAS}.
template <int CM, bool PIPE>
__global__ void __launch_bounds__(256) k_dft64_mfma(const uint64_t* __restrict__ X, uint64_t* __restrict__ Y,
                                                    const v4i* __restrict__ Fpk, const uint64_t* __restrict__ cst,
                                                    uint64_t ws40, uint64_t q, size_t V, int reps, int rep_stride) {
    __shared__ v4i sA[4 * 6 * 64];
    for (int i = threadIdx.x; i < 4 * 6 * 64; i += 256) sA[i] = Fpk[i];
    const int lane = threadIdx.x & 63, g = lane >> 4, col = lane & 15;
    const size_t v0 = ((size_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16 + col;
    const uint64_t nq = 0 - q, kq = cst[0], unbias = cst[1], bias = cst[2];
    uint64_t xs[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) xs[j] = X[(size_t)pi_index<CM>(g, j) * V + v0] + bias;
    __syncthreads();
    for (int r = 0; r < reps; ++r) {
        v4i B[6];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            int P[6];
            byte_planes(xs[4 * d], xs[4 * d + 1], xs[4 * d + 2], xs[4 * d + 3], P);
#pragma unroll
            for (int s = 0; s < 6; ++s) B[s][d] = P[s];
        }
        const v4i* sAr = sA + r * rep_stride;
        if constexpr (!PIPE) {
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) {
                v4i acc[11];
                tile_product(sAr, rt, lane, B, acc);
#pragma unroll
                for (int e = 0; e < 4; ++e) xs[4 * rt + e] = recombine(acc, e, kq, ws40, nq);
            }
        } else {
            // tile rt + 1's 36 products are issued in eight fenced groups of
            // four or five, each beside one stage of tile rt's recombination
            // (four values x two stages), so one wave keeps the matrix core
            // and the VALU busy together
            v4i acc[2][11];
            tile_product(sAr, 0, lane, B, acc[0]);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) {
                v4i (&cur)[11] = acc[rt & 1];
                v4i (&nxt)[11] = acc[(rt + 1) & 1];
                v4i A[6];
                if (rt < 3) {
#pragma unroll
                    for (int t = 0; t < 6; ++t) A[t] = sAr[((rt + 1) * 6 + t) * 64 + lane];
#pragma unroll
                    for (int k = 0; k < 11; ++k) nxt[k] = (v4i){0, 0, 0, 0};
                }
                uint64_t L[4], M[4];
#pragma unroll
                for (int c = 0; c < 8; ++c) {
                    if (rt < 3) {
#pragma unroll
                        for (int m = (c * 36) / 8; m < ((c + 1) * 36) / 8; ++m)
                            nxt[m / 6 + m % 6] =
                                __builtin_amdgcn_mfma_i32_16x16x64_i8(A[m / 6], B[m % 6], nxt[m / 6 + m % 6], 0, 0, 0);
                    }
                    const int e = c >> 1;
                    if (c & 1) xs[4 * rt + e] = recombine2(L[e], M[e], kq, ws40, nq);
                    else recombine1(cur, e, L[e], M[e]);
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        }
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) Y[(size_t)pi_index<CM>(g, j) * V + v0] = (xs[j] + unbias) % q;
}

}  // namespace

hipError_t launch_dft64_probe(const uint64_t* X, uint64_t* Y, const void* tables, uint64_t q, uint64_t ws40,
                              size_t V, int method, int reps, hipStream_t s) {
    const ulonglong2* tw = (const ulonglong2*)tables;                       // 64 twiddle pairs
    const uint64_t* cst = (const uint64_t*)(tw + 64);                       // kq, unbias, BIAS (128 words)
    const v4i* Fpk = (const v4i*)(cst + 128);                               // [2][4][6][64]
    const dim3 grid((unsigned)(V / 64));
    switch (method) {
        case 0:
            hipLaunchKernelGGL(k_dft64_radix2, dim3((unsigned)((V + 255) / 256)), dim3(256), 0, s, X, Y, tw, q, V,
                               reps, 0);
            break;
        case 1:
            hipLaunchKernelGGL((k_dft64_mfma<0, false>), grid, dim3(256), 0, s, X, Y, Fpk, cst, ws40, q, V, reps, 0);
            break;
        case 2:
            hipLaunchKernelGGL((k_dft64_mfma<1, false>), grid, dim3(256), 0, s, X, Y, Fpk + 4 * 6 * 64, cst, ws40, q, V,
                               reps, 0);
            break;
        case 3:
            hipLaunchKernelGGL((k_dft64_mfma<0, true>), grid, dim3(256), 0, s, X, Y, Fpk, cst, ws40, q, V, reps, 0);
            break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}
