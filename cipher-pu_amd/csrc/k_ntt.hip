// Batched NTT / INTT, fused negacyclic polynomial product, bit-reversal
// permutation and element-wise RNS arithmetic.
//
// Tensors are uint64 [poly][limb][N], limb i reduced modulo the prime
// map.idx[i].  One workgroup transforms one (poly, limb) row.
#include "launch.h"
#include "ntt_core.h"
#include "ntt_f64.h"
#include "sample.h"

namespace {

using ntt::Cfg;

__device__ __forceinline__ uint64_t reduce4q(uint64_t x, const PrimeConst& pc) {
    return csub(csub(x, pc.two_q), pc.q);
}

template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_ntt_fwd(uint64_t* dst, const uint64_t* src, const ulonglong2* __restrict__ tw_all,
                                               const PrimeConst* __restrict__ pcs, LimbMap map, int n_limbs,
                                               int reduce_in) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    const int tid = threadIdx.x;
    const int limb = wave_uniform(blockIdx.x % n_limbs);
    const int p = map.idx[limb];
    const PrimeConst pc = pcs[p];
    const size_t off = (size_t)blockIdx.x * C::N;
    const ulonglong2* tw = tw_all + (size_t)p * 2 * C::N;
    uint64_t x[C::E];
    ntt::load_coalesced<C>(x, src + off, tid);
    if (reduce_in) {
#pragma unroll
        for (int e = 0; e < C::E; ++e) x[e] = reduce64(x[e], pc);
    }
    ntt::fwd_ntt<C>(x, lds, tw, pc.q, pc.two_q, tid);
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = reduce64(x[e], pc);
    __syncthreads();
    ntt::store_eval<C>(x, lds, dst + off, tid);
}

template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_ntt_inv(uint64_t* dst, const uint64_t* src, const ulonglong2* __restrict__ tw_all,
                                               const PrimeConst* __restrict__ pcs, LimbMap map, int n_limbs,
                                               int reduce_in) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    const int tid = threadIdx.x;
    const int limb = wave_uniform(blockIdx.x % n_limbs);
    const int p = map.idx[limb];
    const PrimeConst pc = pcs[p];
    const size_t off = (size_t)blockIdx.x * C::N;
    const ulonglong2* itw = tw_all + ((size_t)p * 2 + 1) * C::N;
    uint64_t x[C::E];
    ntt::load_eval<C>(x, lds, src + off, tid);
    if (reduce_in) {
#pragma unroll
        for (int e = 0; e < C::E; ++e) x[e] = reduce64(x[e], pc);
    }
    ntt::inv_ntt<C>(x, lds, itw, pc, tid);
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = csub(x[e], pc.q);
    ntt::store_coalesced<C>(x, dst + off, tid);
}

// out = a * b mod (X^N + 1, q): forward transforms of both operands, the
// point-wise product and the inverse transform in one workgroup
// (rtl/m2/poly_mult.sv:1-43 runs the same three phases on one NTT unit).
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_polymul(const uint64_t* __restrict__ A, const uint64_t* __restrict__ B,
                                               uint64_t* __restrict__ out, const ulonglong2* __restrict__ tw_all,
                                               const PrimeConst* __restrict__ pcs, LimbMap map, int n_limbs) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    const int tid = threadIdx.x;
    const int limb = wave_uniform(blockIdx.x % n_limbs);
    const int p = map.idx[limb];
    const PrimeConst pc = pcs[p];
    const size_t off = (size_t)blockIdx.x * C::N;
    const ulonglong2* tw = tw_all + (size_t)p * 2 * C::N;
    uint64_t x[C::E], y[C::E];
    ntt::load_coalesced<C>(x, A + off, tid);
    ntt::load_coalesced<C>(y, B + off, tid);
    if constexpr (ntt::pair_pays<C>()) {
        ntt::fwd_ntt2<C>(x, y, lds, tw, pc.q, pc.two_q, tid);
    } else {
        ntt::fwd_ntt<C>(x, lds, tw, pc.q, pc.two_q, tid);
        ntt::fwd_ntt<C>(y, lds, tw, pc.q, pc.two_q, tid);
    }
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = mul_barrett(reduce64(x[e], pc), reduce64(y[e], pc), pc);
    // the pair's two exchange buffers serve the inverse as alternating ones
    if constexpr (ntt::pair_pays<C>()) ntt::inv_ntt_db<C>(x, lds, tw + C::N, pc, tid);
    else ntt::inv_ntt<C>(x, lds, tw + C::N, pc, tid);
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = csub(x[e], pc.q);
    ntt::store_coalesced<C>(x, out + off, tid);
}

// k_polymul on the FP64 pipeline (modf64.h): the same three phases, the
// same words out.  Operands below q; the point-wise products take both
// transforms reduced to (q + 1) / 2, so every product is exact.
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_polymul_f(const uint64_t* __restrict__ A, const uint64_t* __restrict__ B,
                                                 uint64_t* __restrict__ out, const double* __restrict__ twf_all,
                                                 const PrimeF* __restrict__ pfs, LimbMap map, int n_limbs) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    double* ldsf = reinterpret_cast<double*>(lds);
    const int tid = threadIdx.x;
    const int limb = wave_uniform(blockIdx.x % n_limbs);
    const int p = map.idx[limb];
    const PrimeF pf = pfs[p];
    const size_t off = (size_t)blockIdx.x * C::N;
    const double* tw = twf_all + (size_t)p * 2 * C::N;
    double x[C::E], y[C::E];
    nttf::load_words<C>(x, A + off, tid);
    nttf::load_words<C>(y, B + off, tid);
    if constexpr (ntt::pair_pays<C>()) {
        nttf::fwd_ntt2<C>(x, y, ldsf, tw, pf, tid);
    } else {
        nttf::fwd_ntt<C>(x, ldsf, tw, pf, tid);
        nttf::fwd_ntt<C>(y, ldsf, tw, pf, tid);
    }
#pragma unroll
    for (int e = 0; e < C::E; ++e)
        x[e] = muld(rmod(x[e], pf.qinv, pf.nq), rmod(y[e], pf.qinv, pf.nq), pf.qinv, pf.nq);
    if constexpr (ntt::pair_pays<C>()) nttf::inv_ntt_db<C>(x, ldsf, tw + C::N, pf, tid);
    else nttf::inv_ntt<C>(x, ldsf, tw + C::N, pf, tid);
    nttf::store_words<C>(x, out + off, pf, tid);
}

// Symmetric RLWE encryption, one (ciphertext b, limb i) row a workgroup:
// c1 = -a, c0 = a s + e + m (exp_e33_real_ciphertext.py:171-175).  The mask
// a and the error e are drawn as k_sample_uniform and k_sample_gauss draw
// them (element (b l + i) N + n of seed_a; element b N + n of seed_e, one
// integer for every limb), so the words are those of the two samplers, the
// product a s and the point-wise sums.  s_eval: s in evaluation order, [l][N].
// The mask is drawn again for c1 rather than kept across the transforms.
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_encrypt_f(const uint64_t* __restrict__ pt, const uint64_t* __restrict__ s_eval,
                                                 uint64_t* __restrict__ out, const double* __restrict__ twf_all,
                                                 const PrimeF* __restrict__ pfs, const PrimeConst* __restrict__ pcs,
                                                 LimbMap map, int l, uint64_t seed_a, uint64_t seed_e,
                                                 GaussTable tab) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    double* ldsf = reinterpret_cast<double*>(lds);
    const int tid = threadIdx.x;
    const int row = (int)blockIdx.x;   // b l + i
    const int i = wave_uniform(row % l), b = wave_uniform(row / l);
    const int p = map.idx[i];
    const PrimeF pf = pfs[p];
    const uint64_t q = pcs[p].q;
    const double* tw = twf_all + (size_t)p * 2 * C::N;
    const size_t base = (size_t)row * C::N;
    double x[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = u2d(smp::uniform(seed_a, base + e * C::T + tid, q));
    nttf::fwd_ntt<C>(x, ldsf, tw, pf, tid);
    const uint64_t* sr = s_eval + (size_t)i * C::N;
#pragma unroll
    for (int e = 0; e < C::E; ++e)
        x[e] = muld(rmod(x[e], pf.qinv, pf.nq), u2d(sr[ntt::eval_index<C>(e, tid)]), pf.qinv, pf.nq);
    nttf::inv_ntt<C>(x, ldsf, tw + C::N, pf, tid);
    const uint64_t* m = pt + base;
    uint64_t* c0 = out + ((size_t)b * 2 * l + i) * C::N;
    uint64_t* c1 = c0 + (size_t)l * C::N;
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int n = e * C::T + tid;
        bool neg;
        const uint64_t k = (uint64_t)smp::gauss(seed_e, (size_t)b * C::N + n, tab, neg);
        uint64_t v = d2u(canon(x[e], pf.qinv, pf.nq)) + (neg ? q - k : k);
        v = v >= q ? v - q : v;
        v += m[n];
        c0[n] = v >= q ? v - q : v;
        const uint64_t a = smp::uniform(seed_a, base + n, q);
        c1[n] = a ? q - a : 0;
    }
}

// Decryption, one (ciphertext b, limb i) row a workgroup: m = c0 + c1 s
// (exp_e33_real_ciphertext.py:178-180), the words of ckks_poly_mul of c1 and
// s followed by the point-wise sum.  ct [batch][2][l][N], s_eval [l][N]
// evaluation order, out [batch][l][N].
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_decrypt_f(const uint64_t* __restrict__ ct, const uint64_t* __restrict__ s_eval,
                                                 uint64_t* __restrict__ out, const double* __restrict__ twf_all,
                                                 const PrimeF* __restrict__ pfs, const PrimeConst* __restrict__ pcs,
                                                 LimbMap map, int l) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    double* ldsf = reinterpret_cast<double*>(lds);
    const int tid = threadIdx.x;
    const int row = (int)blockIdx.x;   // b l + i
    const int i = wave_uniform(row % l), b = wave_uniform(row / l);
    const int p = map.idx[i];
    const PrimeF pf = pfs[p];
    const uint64_t q = pcs[p].q;
    const double* tw = twf_all + (size_t)p * 2 * C::N;
    const uint64_t* c0 = ct + ((size_t)b * 2 * l + i) * C::N;
    double x[C::E];
    nttf::load_words<C>(x, c0 + (size_t)l * C::N, tid);
    nttf::fwd_ntt<C>(x, ldsf, tw, pf, tid);
    const uint64_t* sr = s_eval + (size_t)i * C::N;
#pragma unroll
    for (int e = 0; e < C::E; ++e)
        x[e] = muld(rmod(x[e], pf.qinv, pf.nq), u2d(sr[ntt::eval_index<C>(e, tid)]), pf.qinv, pf.nq);
    nttf::inv_ntt<C>(x, ldsf, tw + C::N, pf, tid);
    uint64_t* o = out + (size_t)row * C::N;
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int n = e * C::T + tid;
        const uint64_t v = d2u(canon(x[e], pf.qinv, pf.nq)) + c0[n];
        o[n] = v >= q ? v - q : v;
    }
}

// Generic transform for small rings (N <= 2048): one stage per barrier.
__device__ void small_fwd(uint64_t* s, const ulonglong2* tw, const PrimeConst& pc, int logN) {
    const int H = 1 << (logN - 1);
    for (int b = logN - 1; b >= 0; --b) {
        for (int t = threadIdx.x; t < H; t += blockDim.x) {
            const int j = ((t >> b) << (b + 1)) | (t & ((1 << b) - 1));
            const ulonglong2 w = tw[(1 << (logN - 1 - b)) + (j >> (b + 1))];
            ct_bfly(s[j], s[j + (1 << b)], w.x, w.y, pc.q, pc.two_q);
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < (1 << logN); k += blockDim.x) s[k] = reduce4q(s[k], pc);
    __syncthreads();
}

__device__ void small_inv(uint64_t* s, const ulonglong2* itw, const PrimeConst& pc, int logN) {
    const int H = 1 << (logN - 1);
    for (int b = 0; b < logN; ++b) {
        for (int t = threadIdx.x; t < H; t += blockDim.x) {
            const int j = ((t >> b) << (b + 1)) | (t & ((1 << b) - 1));
            uint64_t& X = s[j];
            uint64_t& Y = s[j + (1 << b)];
            if (b == logN - 1) {
                const uint64_t u = X, v = Y;
                X = mul_shoup(u + v, pc.ninv, pc.ninv_s, pc.q);
                Y = mul_shoup(u - v + pc.two_q, pc.ninv_w1, pc.ninv_w1_s, pc.q);
            } else {
                const ulonglong2 w = itw[(1 << (logN - 1 - b)) + (j >> (b + 1))];
                gs_bfly(X, Y, w.x, w.y, pc.q, pc.two_q);
            }
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < (1 << logN); k += blockDim.x) s[k] = csub(s[k], pc.q);
    __syncthreads();
}

__global__ void k_ntt_small(uint64_t* dst, const uint64_t* src, const ulonglong2* __restrict__ tw_all,
                            const PrimeConst* __restrict__ pcs, LimbMap map, int n_limbs, int logN, int inverse,
                            int reduce_in) {
    extern __shared__ uint64_t s[];
    const int N = 1 << logN;
    const int p = map.idx[wave_uniform(blockIdx.x % n_limbs)];
    const PrimeConst pc = pcs[p];
    const size_t off = (size_t)blockIdx.x * N;
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        uint64_t v = src[off + k];
        s[k] = reduce_in ? reduce64(v, pc) : v;
    }
    __syncthreads();
    if (inverse)
        small_inv(s, tw_all + ((size_t)p * 2 + 1) * N, pc, logN);
    else
        small_fwd(s, tw_all + (size_t)p * 2 * N, pc, logN);
    for (int k = threadIdx.x; k < N; k += blockDim.x) dst[off + k] = s[k];
}

__global__ void k_polymul_small(const uint64_t* __restrict__ A, const uint64_t* __restrict__ B,
                                uint64_t* __restrict__ out, const ulonglong2* __restrict__ tw_all,
                                const PrimeConst* __restrict__ pcs, LimbMap map, int n_limbs, int logN) {
    extern __shared__ uint64_t s[];
    const int N = 1 << logN;
    const int p = map.idx[wave_uniform(blockIdx.x % n_limbs)];
    const PrimeConst pc = pcs[p];
    const size_t off = (size_t)blockIdx.x * N;
    const ulonglong2* tw = tw_all + (size_t)p * 2 * N;
    for (int k = threadIdx.x; k < N; k += blockDim.x) {
        s[k] = A[off + k];
        s[N + k] = B[off + k];
    }
    __syncthreads();
    small_fwd(s, tw, pc, logN);
    small_fwd(s + N, tw, pc, logN);
    for (int k = threadIdx.x; k < N; k += blockDim.x) s[k] = mul_barrett(s[k], s[N + k], pc);
    __syncthreads();
    small_inv(s, tw + N, pc, logN);
    for (int k = threadIdx.x; k < N; k += blockDim.x) out[off + k] = s[k];
}

__global__ void k_bitrev(uint64_t* __restrict__ data, int logN, size_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const uint32_t N = 1u << logN;
    const uint32_t i = (uint32_t)(g & (N - 1));
    const uint32_t j = logN ? (__brev(i) >> (32 - logN)) : 0;
    if (i < j) {
        uint64_t* row = data + (g - i);
        const uint64_t t = row[i];
        row[i] = row[j];
        row[j] = t;
    }
}

// op 0: a*b, 1: a+b, 2: a-b, 3: -a  (all mod the limb's prime)
__global__ void k_pointwise(const uint64_t* __restrict__ a, const uint64_t* __restrict__ b,
                            uint64_t* __restrict__ out, const PrimeConst* __restrict__ pcs, LimbMap map,
                            int n_limbs, int logN, size_t total, int op) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const int limb = (int)((g >> logN) % n_limbs);
    const PrimeConst& pc = pcs[map.idx[limb]];
    const uint64_t x = a[g];
    uint64_t r;
    switch (op) {
        case 0: r = mul_barrett(x, b[g], pc); break;
        case 1: r = add_mod(x, b[g], pc.q); break;
        case 2: r = sub_mod(x, b[g], pc.q); break;
        case 4: r = mul_special(x, b[g], pc.q); break;
        default: r = x ? pc.q - x : 0; break;
    }
    out[g] = r;
}

// (w, floor(w 2^63 / q)) pairs for key material; w < q < 2^45, the quotient
// by restoring division.
__global__ void k_shoup_pairs(const uint64_t* __restrict__ in, ulonglong2* __restrict__ out,
                              const PrimeConst* __restrict__ pcs, LimbMap map, int n_limbs, int logN, size_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const uint64_t q = pcs[map.idx[(g >> logN) % n_limbs]].q;
    const uint64_t w = in[g];
    uint64_t r = w, quo = 0;
    for (int k = 0; k < 63; ++k) {
        r <<= 1;
        quo <<= 1;
        if (r >= q) {
            r -= q;
            quo |= 1;
        }
    }
    out[g] = make_ulonglong2(w, quo);
}

// (w, RN(w / q)) FP64 pairs of reduced words w < q: the prepared form of
// key and plaintext rows.  The double division is correctly rounded (IEEE
// 754, as the gfx950 division sequence computes it).
__global__ void k_fp64_pairs(const uint64_t* __restrict__ in, double2* __restrict__ out,
                             const PrimeF* __restrict__ pfs, LimbMap map, int n_limbs, int logN, size_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const double q = pfs[map.idx[(g >> logN) % n_limbs]].q;
    const double w = u2d(in[g]);
    out[g] = make_double2(w, w / q);
}

// Key rows on the FP64 pipeline: the doubles of reduced words w < q (the
// key products form RN(w / q) themselves, modf64.h:mulk).
__global__ void k_fp64_words(const uint64_t* __restrict__ in, double* __restrict__ out, size_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    out[g] = u2d(in[g]);
}

// Batched transforms on the FP64 pipeline, the same words as k_ntt_fwd /
// k_ntt_inv.  reduce_in: inputs are any 64-bit words, reduced exactly by
// the integer reduce64 first; otherwise below 2q, the integer kernels'
// contract (the inverse reduces its inputs on load either way).
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_ntt_fwd_f(uint64_t* dst, const uint64_t* src, const double* __restrict__ twf_all,
                                                 const PrimeF* __restrict__ pfs, const PrimeConst* __restrict__ pcs,
                                                 LimbMap map, int n_limbs, int reduce_in) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    double* ldsf = reinterpret_cast<double*>(lds);
    const int tid = threadIdx.x;
    const int p = map.idx[wave_uniform(blockIdx.x % n_limbs)];
    const PrimeF pf = pfs[p];
    const size_t off = (size_t)blockIdx.x * C::N;
    uint64_t u[C::E];
    ntt::load_coalesced<C>(u, src + off, tid);
    if (reduce_in) {
        const PrimeConst pc = pcs[p];
#pragma unroll
        for (int e = 0; e < C::E; ++e) u[e] = reduce64(u[e], pc);
    }
    double x[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = u2d(u[e]);
    nttf::fwd_ntt<C>(x, ldsf, twf_all + (size_t)p * 2 * C::N, pf, tid);
#pragma unroll
    for (int e = 0; e < C::E; ++e) u[e] = d2u(canon(x[e], pf.qinv, pf.nq));
    __syncthreads();
    ntt::store_eval<C>(u, lds, dst + off, tid);
}

template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_ntt_inv_f(uint64_t* dst, const uint64_t* src, const double* __restrict__ twf_all,
                                                 const PrimeF* __restrict__ pfs, const PrimeConst* __restrict__ pcs,
                                                 LimbMap map, int n_limbs, int reduce_in) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    double* ldsf = reinterpret_cast<double*>(lds);
    const int tid = threadIdx.x;
    const int p = map.idx[wave_uniform(blockIdx.x % n_limbs)];
    const PrimeF pf = pfs[p];
    const size_t off = (size_t)blockIdx.x * C::N;
    uint64_t u[C::E];
    ntt::load_eval<C>(u, lds, src + off, tid);
    if (reduce_in) {
        const PrimeConst pc = pcs[p];
#pragma unroll
        for (int e = 0; e < C::E; ++e) u[e] = reduce64(u[e], pc);
    }
    double x[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = rmod(u2d(u[e]), pf.qinv, pf.nq);
    nttf::inv_ntt<C>(x, ldsf, twf_all + ((size_t)p * 2 + 1) * C::N, pf, tid);
    nttf::store_words<C>(x, dst + off, pf, tid);
}

template <typename K>
hipError_t set_lds(K kernel, int bytes) {
    return hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

template <int LOGN, int T>
hipError_t run_ntt(const DevTables& t, uint64_t* dst, const uint64_t* src, int rows, int n_limbs, const LimbMap& map, bool inv,
                   bool red, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    if (t.fp64) {
        auto k = inv ? k_ntt_inv_f<LOGN, T> : k_ntt_fwd_f<LOGN, T>;
        hipError_t e = set_lds(k, C::LDS_BYTES);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3(rows), dim3(T), C::LDS_BYTES, s, dst, src, t.twf, t.pf, t.pc, map, n_limbs,
                           red ? 1 : 0);
        return hipGetLastError();
    }
    auto k = inv ? k_ntt_inv<LOGN, T> : k_ntt_fwd<LOGN, T>;
    hipError_t e = set_lds(k, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3(rows), dim3(T), C::LDS_BYTES, s, dst, src, t.tw, t.pc, map, n_limbs, red ? 1 : 0);
    return hipGetLastError();
}

template <int LOGN, int T>
hipError_t run_polymul(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* out, int rows,
                       int n_limbs, const LimbMap& map, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    constexpr int lds = ntt::pair_pays<C>() ? 2 * C::LDS_BYTES : C::LDS_BYTES;
    hipError_t e = set_lds(k_polymul<LOGN, T>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_polymul<LOGN, T>), dim3(rows), dim3(T), lds, s, a, b, out, t.tw, t.pc, map,
                       n_limbs);
    return hipGetLastError();
}

template <int LOGN, int T>
hipError_t run_polymul_f(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* out, int rows,
                         int n_limbs, const LimbMap& map, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    constexpr int lds = ntt::pair_pays<C>() ? 2 * C::LDS_BYTES : C::LDS_BYTES;
    hipError_t e = set_lds(k_polymul_f<LOGN, T>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_polymul_f<LOGN, T>), dim3(rows), dim3(T), lds, s, a, b, out, t.twf, t.pf, map,
                       n_limbs);
    return hipGetLastError();
}

template <int LOGN, int T>
hipError_t run_encrypt_f(const DevTables& t, const uint64_t* pt, const uint64_t* s_eval, uint64_t* out, int batch,
                         int l, const LimbMap& map, uint64_t seed_a, uint64_t seed_e, const GaussTable& tab,
                         hipStream_t s) {
    using C = Cfg<LOGN, T>;
    hipError_t e = set_lds(k_encrypt_f<LOGN, T>, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_encrypt_f<LOGN, T>), dim3(batch * l), dim3(T), C::LDS_BYTES, s, pt, s_eval, out, t.twf,
                       t.pf, t.pc, map, l, seed_a, seed_e, tab);
    return hipGetLastError();
}

template <int LOGN, int T>
hipError_t run_decrypt_f(const DevTables& t, const uint64_t* ct, const uint64_t* s_eval, uint64_t* out, int batch,
                         int l, const LimbMap& map, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    hipError_t e = set_lds(k_decrypt_f<LOGN, T>, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_decrypt_f<LOGN, T>), dim3(batch * l), dim3(T), C::LDS_BYTES, s, ct, s_eval, out, t.twf,
                       t.pf, t.pc, map, l);
    return hipGetLastError();
}

// Untuned launch shape of the batched transforms (fastest standalone
// transform on MI355X at N = 8192: 512 threads).
int default_threads(int logN) { return logN >= 13 ? 512 : 256; }

}  // namespace

hipError_t launch_ntt(const DevTables& t, uint64_t* dst, const uint64_t* src, int n_polys, int n_limbs, const LimbMap& map,
                      bool inverse, bool reduce_input, int threads, hipStream_t s) {
    const int rows = n_polys * n_limbs;
    if (rows == 0) return hipSuccess;
    if (!threads) threads = default_threads(t.logN);
    switch (t.logN * 10000 + threads) {
        case 120128: return run_ntt<12, 128>(t, dst, src, rows, n_limbs, map, inverse, reduce_input, s);
        case 120256: return run_ntt<12, 256>(t, dst, src, rows, n_limbs, map, inverse, reduce_input, s);
        case 120512: return run_ntt<12, 512>(t, dst, src, rows, n_limbs, map, inverse, reduce_input, s);
        case 130256: return run_ntt<13, 256>(t, dst, src, rows, n_limbs, map, inverse, reduce_input, s);
        case 130512: return run_ntt<13, 512>(t, dst, src, rows, n_limbs, map, inverse, reduce_input, s);
        case 131024: return run_ntt<13, 1024>(t, dst, src, rows, n_limbs, map, inverse, reduce_input, s);
        case 140512: return run_ntt<14, 512>(t, dst, src, rows, n_limbs, map, inverse, reduce_input, s);
        case 141024: return run_ntt<14, 1024>(t, dst, src, rows, n_limbs, map, inverse, reduce_input, s);
        default: break;
    }
    if (t.logN > 11) return hipErrorInvalidValue;
    const int N = 1 << t.logN;
    const int thr = N / 2 < 64 ? 64 : (N / 2 > 1024 ? 1024 : N / 2);
    hipLaunchKernelGGL(k_ntt_small, dim3(rows), dim3(thr), N * 8, s, dst, src, t.tw, t.pc, map, n_limbs, t.logN,
                       inverse ? 1 : 0, reduce_input ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_polymul(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* out, int n_polys,
                          int n_limbs, const LimbMap& map, int threads, hipStream_t s) {
    const int rows = n_polys * n_limbs;
    if (rows == 0) return hipSuccess;
    if (!threads) threads = default_threads(t.logN);
    if (t.fp64) {
        switch (t.logN * 10000 + threads) {
            case 120128: return run_polymul_f<12, 128>(t, a, b, out, rows, n_limbs, map, s);
            case 120256: return run_polymul_f<12, 256>(t, a, b, out, rows, n_limbs, map, s);
            case 120512: return run_polymul_f<12, 512>(t, a, b, out, rows, n_limbs, map, s);
            case 130256: return run_polymul_f<13, 256>(t, a, b, out, rows, n_limbs, map, s);
            case 130512: return run_polymul_f<13, 512>(t, a, b, out, rows, n_limbs, map, s);
            case 131024: return run_polymul_f<13, 1024>(t, a, b, out, rows, n_limbs, map, s);
            case 140512: return run_polymul_f<14, 512>(t, a, b, out, rows, n_limbs, map, s);
            case 141024: return run_polymul_f<14, 1024>(t, a, b, out, rows, n_limbs, map, s);
            default: break;
        }
    }
    switch (t.logN * 10000 + threads) {
        case 120128: return run_polymul<12, 128>(t, a, b, out, rows, n_limbs, map, s);
        case 120256: return run_polymul<12, 256>(t, a, b, out, rows, n_limbs, map, s);
        case 120512: return run_polymul<12, 512>(t, a, b, out, rows, n_limbs, map, s);
        case 130256: return run_polymul<13, 256>(t, a, b, out, rows, n_limbs, map, s);
        case 130512: return run_polymul<13, 512>(t, a, b, out, rows, n_limbs, map, s);
        case 131024: return run_polymul<13, 1024>(t, a, b, out, rows, n_limbs, map, s);
        case 140512: return run_polymul<14, 512>(t, a, b, out, rows, n_limbs, map, s);
        case 141024: return run_polymul<14, 1024>(t, a, b, out, rows, n_limbs, map, s);
        default: break;
    }
    if (t.logN > 11) return hipErrorInvalidValue;
    const int N = 1 << t.logN;
    const int thr = N / 2 < 64 ? 64 : (N / 2 > 1024 ? 1024 : N / 2);
    hipLaunchKernelGGL(k_polymul_small, dim3(rows), dim3(thr), 2 * N * 8, s, a, b, out, t.tw, t.pc, map, n_limbs,
                       t.logN);
    return hipGetLastError();
}

hipError_t launch_encrypt_sk(const DevTables& t, const uint64_t* pt, const uint64_t* s_eval, uint64_t* out, int batch,
                             int l, const LimbMap& map, uint64_t seed_a, uint64_t seed_e, const GaussTable& tab,
                             hipStream_t s) {
    if (batch * l == 0) return hipSuccess;
    if (!t.fp64) return hipErrorInvalidValue;
    switch (t.logN) {
        case 12: return run_encrypt_f<12, 256>(t, pt, s_eval, out, batch, l, map, seed_a, seed_e, tab, s);
        case 13: return run_encrypt_f<13, 512>(t, pt, s_eval, out, batch, l, map, seed_a, seed_e, tab, s);
        case 14: return run_encrypt_f<14, 512>(t, pt, s_eval, out, batch, l, map, seed_a, seed_e, tab, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_decrypt_sk(const DevTables& t, const uint64_t* ct, const uint64_t* s_eval, uint64_t* out, int batch,
                             int l, const LimbMap& map, hipStream_t s) {
    if (batch * l == 0) return hipSuccess;
    if (!t.fp64) return hipErrorInvalidValue;
    switch (t.logN) {
        case 12: return run_decrypt_f<12, 256>(t, ct, s_eval, out, batch, l, map, s);
        case 13: return run_decrypt_f<13, 512>(t, ct, s_eval, out, batch, l, map, s);
        case 14: return run_decrypt_f<14, 512>(t, ct, s_eval, out, batch, l, map, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_bitrev(uint64_t* data, int n_rows, int logN, hipStream_t s) {
    const size_t total = (size_t)n_rows << logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_bitrev, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, data, logN, total);
    return hipGetLastError();
}

hipError_t launch_pointwise(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* out, int n_polys,
                            int n_limbs, const LimbMap& map, int op, hipStream_t s) {
    const size_t total = ((size_t)n_polys * n_limbs) << t.logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_pointwise, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a, b, out, t.pc, map,
                       n_limbs, t.logN, total, op);
    return hipGetLastError();
}

hipError_t launch_prepare_pairs(const DevTables& t, const uint64_t* in, void* out, int n_polys, int n_limbs,
                                const LimbMap& map, hipStream_t s) {
    const size_t total = ((size_t)n_polys * n_limbs) << t.logN;
    if (!total) return hipSuccess;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (t.fp64)
        hipLaunchKernelGGL(k_fp64_pairs, grid, dim3(256), 0, s, in, (double2*)out, t.pf, map, n_limbs, t.logN, total);
    else
        hipLaunchKernelGGL(k_shoup_pairs, grid, dim3(256), 0, s, in, (ulonglong2*)out, t.pc, map, n_limbs, t.logN,
                           total);
    return hipGetLastError();
}

hipError_t launch_prepare_key(const DevTables& t, const uint64_t* in, void* out, int n_polys, int n_limbs,
                              const LimbMap& map, hipStream_t s) {
    if (!t.fp64) return launch_prepare_pairs(t, in, out, n_polys, n_limbs, map, s);
    const size_t total = ((size_t)n_polys * n_limbs) << t.logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_fp64_words, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, (double*)out, total);
    return hipGetLastError();
}
