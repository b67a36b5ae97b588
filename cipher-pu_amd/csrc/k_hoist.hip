// Hoisted rotations: several slot rotations of the same ciphertexts share
// one digit decomposition ("ModUp") -- the method of Halevi and Shoup,
// "Faster homomorphic linear transformations in HElib" (CRYPTO 2018).
// The reference rotates one step at a time (exp_e24_rotation_bootstrap.py:
// 143-176, automorphism then key switch); there is no hoisted counterpart.
//
// For ct = (c0, c1) at l limbs, digits D_j = lift of [c1]_(q_j), and the
// evaluation-domain automorphism pi_g, NTT_t(sigma_g(D_j)) = pi_g(NTT_t(D_j))
// where sigma_g acts over Z (a coefficient sent past X^N changes sign).  So
//   acc_t = sum_j pi_g(NTT_t(D_j)) * K_(j,t) = pi_g( sum_j NTT_t(D_j) * pi_g^-1(K_(j,t)) ).
// The key of each rotation is stored pre-permuted by pi_g^-1 (ckks_key_permute),
// the sum reads every operand in place, and pi_g is applied once per
// accumulator through LDS before the inverse transform.  The digits here
// are sigma_g of the canonical lifts -- a different (equally small)
// representative than the lift of sigma_g(c1) that the one-at-a-time rotation
// decomposes, so the two agree after decryption, not bit for bit.
//
// Evaluation order: position i of a forward transform holds
// a(psi^(2 brv(i) + 1)), so pi_g(v)[i] = v[j(i)] with
// 2 brv(j(i)) + 1 = g (2 brv(i) + 1) mod 2N.
#include "launch.h"
#include "ntt_core.h"
#include "ntt_f64.h"

using ntt::Cfg;

namespace {

__device__ __forceinline__ int eval_perm(int i, uint32_t g, int logN) {
    const uint32_t shift = 32 - logN;
    const uint32_t e = 2 * (__brev((uint32_t)i) >> shift) + 1;
    const uint32_t e2 = (g * e) & ((2u << logN) - 1);
    return (int)(__brev((e2 - 1) >> 1) >> shift);
}

// D[b][j][t] = NTT_t(lift [c1_b]_(q_j)), evaluation order, reduced; one
// workgroup per (ciphertext, digit j, target t), t = l is the special prime.
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_modup(const uint64_t* __restrict__ ct, uint64_t* __restrict__ D,
                                             const ulonglong2* __restrict__ tw_all,
                                             const PrimeConst* __restrict__ pcs, LimbMap map, int l, int pidx_p,
                                             int reduce_digits) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    const int tid = threadIdx.x;
    const int per_b = l * (l + 1);
    const int b = wave_uniform(blockIdx.x / per_b), rem = wave_uniform(blockIdx.x % per_b);
    const int j = rem / (l + 1), t = rem % (l + 1);
    const int pt = t < l ? map.idx[t] : pidx_p;
    const PrimeConst pc = pcs[pt];
    uint64_t x[C::E];
    ntt::load_coalesced<C>(x, ct + ((size_t)b * 2 + 1) * l * C::N + (size_t)j * C::N, tid);
    if (reduce_digits) {
#pragma unroll
        for (int k = 0; k < C::E; ++k) x[k] = reduce64(x[k], pc);
    }
    ntt::fwd_ntt<C>(x, lds, tw_all + (size_t)pt * 2 * C::N, pc.q, pc.two_q, tid);
#pragma unroll
    for (int k = 0; k < C::E; ++k) x[k] = reduce64(x[k], pc);
    __syncthreads();
    ntt::store_eval<C>(x, lds, D + (((size_t)b * l + j) * (l + 1) + t) * C::N, tid);
}

// k_modup on the FP64 pipeline: the same canonical words.  The digit's
// residues (below 2 q_t on a narrow chain, reduced first otherwise) enter the
// transform as doubles; the transform's outputs (below 16 q) go through LDS
// to the coalesced evaluation order and are stored canonical.
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_modup_f(const uint64_t* __restrict__ ct, uint64_t* __restrict__ D,
                                               const double* __restrict__ twf_all, const PrimeF* __restrict__ pfs,
                                               LimbMap map, int l, int pidx_p, int reduce_digits) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds_u[];
    double* lds = reinterpret_cast<double*>(lds_u);
    const int tid = threadIdx.x;
    const int per_b = l * (l + 1);
    const int b = wave_uniform(blockIdx.x / per_b), rem = wave_uniform(blockIdx.x % per_b);
    const int j = rem / (l + 1), t = rem % (l + 1);
    const int pt = t < l ? map.idx[t] : pidx_p;
    const PrimeF pf = pfs[pt];
    double x[C::E];
    nttf::load_words<C>(x, ct + ((size_t)b * 2 + 1) * l * C::N + (size_t)j * C::N, tid);
    if (reduce_digits) {
#pragma unroll
        for (int k = 0; k < C::E; ++k) x[k] = rmod(x[k], pf.qinv, pf.nq);
    }
    nttf::fwd_ntt<C>(x, lds, twf_all + (size_t)pt * 2 * C::N, pf, tid);
    __syncthreads();
    ntt::to_lds<C, C::NP - 1>(x, lds, tid);
    __syncthreads();
    uint64_t* dst = D + (((size_t)b * l + j) * (l + 1) + t) * C::N;
#pragma unroll
    for (int e = 0; e < C::E; ++e)
        dst[e * C::T + tid] = d2u(canon(lds[ntt::pad<C>(e * C::T + tid)], pf.qinv, pf.nq));
}

// Reduce an accumulator whose register k holds evaluation slot k T + tid
// (the coalesced order the key products are summed in; FP64, below 14 q),
// apply pi_g through LDS, inverse-transform and store it in coefficient form.
template <class C>
__device__ __forceinline__ void permute_inverse(double (&a)[C::E], uint32_t g, double* lds,
                                                const double* __restrict__ itw, const PrimeF& pf,
                                                uint64_t* __restrict__ out, int tid) {
#pragma unroll
    for (int k = 0; k < C::E; ++k) lds[ntt::pad<C>(k * C::T + tid)] = rmod(a[k], pf.qinv, pf.nq);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) a[k] = lds[ntt::pad<C>(eval_perm(ntt::eval_index<C>(k, tid), g, C::LOGN))];
    __syncthreads();
    nttf::inv_ntt<C>(a, lds, itw, pf, tid);
    nttf::store_words<C>(a, out, pf, tid);
    __syncthreads();
}

// Reduce an accumulator whose register k holds evaluation slot k T + tid
// and store pi_g of it in evaluation order: one LDS round trip, coalesced
// stores.
template <class C>
__device__ __forceinline__ void permute_store_eval(const double (&a)[C::E], uint32_t g, double* lds,
                                                   const PrimeF& pf, uint64_t* __restrict__ out, int tid) {
#pragma unroll
    for (int k = 0; k < C::E; ++k) lds[ntt::pad<C>(k * C::T + tid)] = rmod(a[k], pf.qinv, pf.nq);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int j = e * C::T + tid;
        out[j] = d2u(canon(lds[ntt::pad<C>(eval_perm(j, g, C::LOGN))], pf.qinv, pf.nq));
    }
    __syncthreads();
}

// permute_inverse for both accumulators at once: two LDS buffers, one
// gather each, one transform pair.
template <class C>
__device__ __forceinline__ void permute_inverse2(double (&a0)[C::E], double (&a1)[C::E], uint32_t g, double* lds,
                                                 const double* __restrict__ itw, const PrimeF& pf,
                                                 uint64_t* __restrict__ out0, uint64_t* __restrict__ out1, int tid) {
    double* lds1 = lds + C::LDS_WORDS;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int slot = ntt::pad<C>(k * C::T + tid);
        lds[slot] = rmod(a0[k], pf.qinv, pf.nq);
        lds1[slot] = rmod(a1[k], pf.qinv, pf.nq);
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int src = ntt::pad<C>(eval_perm(ntt::eval_index<C>(k, tid), g, C::LOGN));
        a0[k] = lds[src];
        a1[k] = lds1[src];
    }
    __syncthreads();
    nttf::inv_ntt2<C>(a0, a1, lds, itw, pf, tid);
    nttf::store_words<C>(a0, out0, pf, tid);
    nttf::store_words<C>(a1, out1, pf, tid);
}

// One workgroup per (target t, ciphertext b, rotation r), placed by
// hoist_slot.  keys: [R][dnum][2][key_limbs][N] key words as doubles,
// permuted by pi_g^-1; acc: [R][B][2][l+1][N] coefficient form, the
// layout ckks_relin's ModDown reads.  acc1 = -(sum D a) directly.
template <int LOGN, int T, bool EVAL>
__global__ void __launch_bounds__(T) k_ks_hoisted(const uint64_t* __restrict__ D, const double* __restrict__ keys,
                                                  size_t key_stride, int key_limbs, GaloisList galois,
                                                  uint64_t* __restrict__ acc, const double* __restrict__ twf_all,
                                                  const PrimeF* __restrict__ pfs, LimbMap map, int l, int pidx_p,
                                                  int batch, int R, HoistOrder order) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds_u[];
    double* lds = reinterpret_cast<double*>(lds_u);
    const int tid = threadIdx.x;
    const TRB s = hoist_slot(blockIdx.x, gridDim.x, R, batch, order);
    const int b = wave_uniform(s.b), r = wave_uniform(s.r), t = wave_uniform(s.t);
    const int pt = t < l ? map.idx[t] : pidx_p;
    const int kt = t < l ? t : key_limbs - 1;
    const PrimeF pf = pfs[pt];
    const double nq = pf.nq;
    const double* key = keys + (size_t)r * key_stride;
    double a0[C::E], a1[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) a0[k] = a1[k] = 0.0;
    // the digit row and its two key rows, coalesced (the sum is slot by slot)
    auto load_digit = [&](int j, uint64_t (&dv)[C::E], double (&bv)[C::E], double (&av)[C::E]) {
        const uint64_t* d = D + (((size_t)b * l + j) * (l + 1) + t) * C::N;
        const double* kb = key + ((size_t)(j * 2 + 0) * key_limbs + kt) * C::N;
        const double* ka = key + ((size_t)(j * 2 + 1) * key_limbs + kt) * C::N;
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            dv[k] = d[k * C::T + tid];
            bv[k] = kb[k * C::T + tid];
            av[k] = ka[k * C::T + tid];
        }
    };
    // Without EVAL's epilogue (and at most 8 words a thread) the registers
    // allow the next digit's rows to be loaded during the current digit's
    // products (the last digit loads its own rows again: a clamped, valid
    // address).
    if constexpr (!EVAL && C::E <= 8) {
        uint64_t nd[C::E];
        double nb[C::E], na[C::E];
        load_digit(0, nd, nb, na);
        for (int j = 0; j < l; ++j) {
            uint64_t dv[C::E];
            double bv[C::E], av[C::E];
#pragma unroll
            for (int k = 0; k < C::E; ++k) {
                dv[k] = nd[k];
                bv[k] = nb[k];
                av[k] = na[k];
            }
            load_digit(j + 1 < l ? j + 1 : j, nd, nb, na);
#pragma unroll
            for (int k = 0; k < C::E; ++k) {
#pragma clang fp contract(off)
                const double x = u2d(dv[k]);
                a0[k] = a0[k] + mulk(x, bv[k], pf.qinv, nq);   // 7q/8 each, l <= 15 digits
                a1[k] = a1[k] - mulk(x, av[k], pf.qinv, nq);
            }
        }
    } else if constexpr (EVAL && C::E == 8) {
        // EVAL's epilogue leaves registers for half a digit ahead: the words
        // in two halves of 4, each summed over the digits in order, the next
        // (digit, half)'s rows loaded during the current one's products
        constexpr int H = C::E / 2;
        auto load_half = [&](int j, int h, uint64_t (&dv)[H], double (&bv)[H], double (&av)[H]) {
            const uint64_t* d = D + (((size_t)b * l + j) * (l + 1) + t) * C::N;
            const double* kb = key + ((size_t)(j * 2 + 0) * key_limbs + kt) * C::N;
            const double* ka = key + ((size_t)(j * 2 + 1) * key_limbs + kt) * C::N;
#pragma unroll
            for (int k = 0; k < H; ++k) {
                const int idx = (h * H + k) * C::T + tid;
                dv[k] = d[idx];
                bv[k] = kb[idx];
                av[k] = ka[idx];
            }
        };
        uint64_t nd[H];
        double nb[H], na[H];
        load_half(0, 0, nd, nb, na);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            for (int j = 0; j < l; ++j) {
                uint64_t dv[H];
                double bv[H], av[H];
#pragma unroll
                for (int k = 0; k < H; ++k) {
                    dv[k] = nd[k];
                    bv[k] = nb[k];
                    av[k] = na[k];
                }
                const bool last = j + 1 == l;
                load_half(last ? (h == 0 ? 0 : j) : j + 1, last && h == 0 ? 1 : h, nd, nb, na);
#pragma unroll
                for (int k = 0; k < H; ++k) {
#pragma clang fp contract(off)
                    const double x = u2d(dv[k]);
                    a0[h * H + k] = a0[h * H + k] + mulk(x, bv[k], pf.qinv, nq);   // 7q/8 each, l <= 15 digits
                    a1[h * H + k] = a1[h * H + k] - mulk(x, av[k], pf.qinv, nq);
                }
            }
        }
    } else {
        for (int j = 0; j < l; ++j) {
            const uint64_t* d = D + (((size_t)b * l + j) * (l + 1) + t) * C::N;
            const double* kb = key + ((size_t)(j * 2 + 0) * key_limbs + kt) * C::N;
            const double* ka = key + ((size_t)(j * 2 + 1) * key_limbs + kt) * C::N;
#pragma unroll
            for (int k = 0; k < C::E; ++k) {
#pragma clang fp contract(off)
                const int idx = k * C::T + tid;
                const double x = u2d(d[idx]);
                a0[k] = a0[k] + mulk(x, kb[idx], pf.qinv, nq);   // 7q/8 each, l <= 15 digits
                a1[k] = a1[k] - mulk(x, ka[idx], pf.qinv, nq);
            }
        }
    }
    const uint32_t g = galois.g[r];
    const double* itw = twf_all + ((size_t)pt * 2 + 1) * C::N;
    uint64_t* out = acc + (((size_t)r * batch + b) * 2 * (l + 1) + t) * C::N;
    if (EVAL && t < l) {
        // chain limbs stay in evaluation form; ModDown finishes there
        permute_store_eval<C>(a0, g, lds, pf, out, tid);
        permute_store_eval<C>(a1, g, lds, pf, out + (size_t)(l + 1) * C::N, tid);
        return;
    }
    if constexpr (ntt::pair_pays<C>()) {
        permute_inverse2<C>(a0, a1, g, lds, itw, pf, out, out + (size_t)(l + 1) * C::N, tid);
    } else {
        permute_inverse<C>(a0, g, lds, itw, pf, out, tid);
        permute_inverse<C>(a1, g, lds, itw, pf, out + (size_t)(l + 1) * C::N, tid);
    }
}

// Reduce (and negate) an accumulator whose register k holds evaluation slot
// k T + tid (the coalesced order the key products are summed in), apply
// pi_g through LDS, inverse-transform and store it in coefficient form.
template <class C>
__device__ __forceinline__ void permute_inverse_i(uint64_t (&a)[C::E], bool negate, uint32_t g, uint64_t* lds,
                                                const ulonglong2* __restrict__ itw, const PrimeConst& pc,
                                                uint64_t* __restrict__ out, int tid) {
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const uint64_t v = reduce64(a[k], pc);
        lds[ntt::pad<C>(k * C::T + tid)] = (negate && v) ? pc.q - v : v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) a[k] = lds[ntt::pad<C>(eval_perm(ntt::eval_index<C>(k, tid), g, C::LOGN))];
    __syncthreads();
    ntt::inv_ntt<C>(a, lds, itw, pc, tid);
#pragma unroll
    for (int k = 0; k < C::E; ++k) a[k] = csub(a[k], pc.q);
    ntt::store_coalesced<C>(a, out, tid);
    __syncthreads();
}

// Reduce (and negate) an accumulator whose register k holds evaluation slot
// k T + tid and
// store pi_g of it in evaluation order: one LDS round trip, coalesced stores.
template <class C>
__device__ __forceinline__ void permute_store_eval_i(uint64_t (&a)[C::E], bool negate, uint32_t g, uint64_t* lds,
                                                   const PrimeConst& pc, uint64_t* __restrict__ out, int tid) {
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const uint64_t v = reduce64(a[k], pc);
        lds[ntt::pad<C>(k * C::T + tid)] = (negate && v) ? pc.q - v : v;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int j = e * C::T + tid;
        out[j] = lds[ntt::pad<C>(eval_perm(j, g, C::LOGN))];
    }
    __syncthreads();
}

// permute_inverse_i for both accumulators at once (the second negated): two
// LDS buffers, one gather each, one transform pair.
template <class C>
__device__ __forceinline__ void permute_inverse2_i(uint64_t (&a0)[C::E], uint64_t (&a1)[C::E], uint32_t g,
                                                 uint64_t* lds, const ulonglong2* __restrict__ itw,
                                                 const PrimeConst& pc, uint64_t* __restrict__ out0,
                                                 uint64_t* __restrict__ out1, int tid) {
    uint64_t* lds1 = lds + C::LDS_WORDS;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int slot = ntt::pad<C>(k * C::T + tid);
        lds[slot] = reduce64(a0[k], pc);
        const uint64_t v = reduce64(a1[k], pc);
        lds1[slot] = v ? pc.q - v : 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int src = ntt::pad<C>(eval_perm(ntt::eval_index<C>(k, tid), g, C::LOGN));
        a0[k] = lds[src];
        a1[k] = lds1[src];
    }
    __syncthreads();
    ntt::inv_ntt2<C>(a0, a1, lds, itw, pc, tid);
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        a0[k] = csub(a0[k], pc.q);
        a1[k] = csub(a1[k], pc.q);
    }
    ntt::store_coalesced<C>(a0, out0, tid);
    ntt::store_coalesced<C>(a1, out1, tid);
}

// k_ks_hoisted on 64-bit integer words (keys of (value, Shoup companion)
// pairs), for contexts set to CKKS_ARITH_INT64.  One workgroup per (target
// t, ciphertext b, rotation r), placed by hoist_slot.  keys: [R][dnum][2][key_limbs][N]
// (value, companion) pairs permuted by pi_g^-1; acc: [R][B][2][l+1][N]
// coefficient form, the layout ckks_relin's ModDown reads.
template <int LOGN, int T, bool EVAL>
__global__ void __launch_bounds__(T) k_ks_hoisted_i(const uint64_t* __restrict__ D, const ulonglong2* __restrict__ keys,
                                                  size_t key_stride, int key_limbs, GaloisList galois,
                                                  uint64_t* __restrict__ acc, const ulonglong2* __restrict__ tw_all,
                                                  const PrimeConst* __restrict__ pcs, LimbMap map, int l, int pidx_p,
                                                  int batch, int R, HoistOrder order) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    const int tid = threadIdx.x;
    const TRB s = hoist_slot(blockIdx.x, gridDim.x, R, batch, order);
    const int b = wave_uniform(s.b), r = wave_uniform(s.r), t = wave_uniform(s.t);
    const int pt = t < l ? map.idx[t] : pidx_p;
    const int kt = t < l ? t : key_limbs - 1;
    const PrimeConst pc = pcs[pt];
    const uint64_t nq = (uint64_t)0 - pc.q;
    const ulonglong2* key = keys + (size_t)r * key_stride;
    uint64_t a0[C::E], a1[C::E];
#pragma unroll
    for (int k = 0; k < C::E; ++k) a0[k] = a1[k] = 0;
    for (int j = 0; j < l; ++j) {
        const uint64_t* d = D + (((size_t)b * l + j) * (l + 1) + t) * C::N;
        const ulonglong2* kb = key + ((size_t)(j * 2 + 0) * key_limbs + kt) * C::N;
        const ulonglong2* ka = key + ((size_t)(j * 2 + 1) * key_limbs + kt) * C::N;
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            const int idx = k * C::T + tid;          // coalesced: the sum is slot by slot
            const uint64_t x = d[idx];
            const ulonglong2 wb = kb[idx], wa = ka[idx];
            a0[k] += mul_shoup4_nq(x, wb.x, wb.y, nq);   // < 4q each, l <= 15 digits
            a1[k] += mul_shoup4_nq(x, wa.x, wa.y, nq);
        }
    }
    const uint32_t g = galois.g[r];
    const ulonglong2* itw = tw_all + ((size_t)pt * 2 + 1) * C::N;
    uint64_t* out = acc + (((size_t)r * batch + b) * 2 * (l + 1) + t) * C::N;
    if (EVAL && t < l) {
        // chain limbs stay in evaluation form; ModDown finishes there
        permute_store_eval_i<C>(a0, false, g, lds, pc, out, tid);
        permute_store_eval_i<C>(a1, true, g, lds, pc, out + (size_t)(l + 1) * C::N, tid);
        return;
    }
    if constexpr (ntt::pair_pays<C>()) {
        permute_inverse2_i<C>(a0, a1, g, lds, itw, pc, out, out + (size_t)(l + 1) * C::N, tid);
    } else {
        permute_inverse_i<C>(a0, false, g, lds, itw, pc, out, tid);
        permute_inverse_i<C>(a1, true, g, lds, itw, pc, out + (size_t)(l + 1) * C::N, tid);
    }
}

// ModDown in the evaluation domain for hoisted rotations, one workgroup per
// (rotation r, ciphertext b, poly, limb i):
//   out_i = (X_i - NTT_i(w)) p^-1 (+ pi_g(NTT_i(c0)) for poly 0),
//   w = [(x_p + h) mod p] - h (mod q_i), h = (p - 1)/2,
// the evaluation-domain image of the coefficient ModDown
// ((x_i + h) - [(x_p + h) mod p]) p^-1 (round half up of x / p): X_i is the
// chain-limb accumulator left in evaluation form and x_p the special limb in
// coefficient form.  acc: [R][B][2][l+1][N]; Cn: the ciphertexts in
// evaluation form [B][2][l][N]; out: [R][B][2][l][N] in evaluation order.
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_moddown_eval(const uint64_t* __restrict__ acc, const uint64_t* __restrict__ Cn,
                                                    uint64_t* __restrict__ out, const ulonglong2* __restrict__ tw_all,
                                                    const PrimeConst* __restrict__ pcs, LimbMap map, int l, int pidx_p,
                                                    ModDownConst mc, GaloisList galois, int batch, int narrow) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    const int tid = threadIdx.x;
    const int i = wave_uniform(blockIdx.x % l);
    const int row = wave_uniform(blockIdx.x / l);                 // ((r batch + b) 2 + poly)
    const int poly = row & 1, rb = row >> 1;
    const int b = rb % batch, r = rb / batch;
    const int p = map.idx[i];
    const PrimeConst pc = pcs[p];
    const uint64_t qp = pcs[pidx_p].q, hp = mc.hp[i];
    const uint64_t* arow = acc + (size_t)row * (l + 1) * C::N;
    uint64_t x[C::E];
    ntt::load_coalesced<C>(x, arow + (size_t)l * C::N, tid);
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const uint64_t v = csub(x[e] + (qp >> 1), qp);
        x[e] = sub_mod(narrow ? csub(v, pc.q) : reduce64(v, pc), hp, pc.q);
    }
    ntt::fwd_ntt<C>(x, lds, tw_all + (size_t)p * 2 * C::N, pc.q, pc.two_q, tid);
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = reduce64(x[e], pc);
    __syncthreads();
    ntt::to_lds<C, C::NP - 1>(x, lds, tid);        // NTT_i(w) at its evaluation slots
    __syncthreads();
    const uint64_t* X = arow + (size_t)i * C::N;
    const uint64_t* c0 = Cn + ((size_t)b * 2 * l + i) * C::N;
    const uint32_t g = galois.g[r];
    uint64_t* o = out + ((size_t)row * l + i) * C::N;
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int j = e * C::T + tid;
        uint64_t v = sub_mod(X[j], lds[ntt::pad<C>(j)], pc.q);
        v = csub(mul_shoup(v, mc.pinv[i], mc.pinv_s[i], pc.q), pc.q);
        if (poly == 0) v = add_mod(v, c0[eval_perm(j, g, C::LOGN)], pc.q);
        o[j] = v;
    }
}

// k_moddown_eval on the FP64 pipeline: the same words.  w enters the
// transform reduced to (q + 1) / 2; NTT_i(w) (below 16 q) is read back from
// LDS into registers, and for poly 0 the LDS then stages the row of c0 so
// that its permuted gather reads the LDS instead of scattered words of
// memory.  (X - NTT_i(w)) is below 15 q, its product by p^-1 below 5q/8, and
// the sum with c0 below 2 q: one canonical reduction at the store.
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_moddown_eval_f(const uint64_t* __restrict__ acc,
                                                      const uint64_t* __restrict__ Cn, uint64_t* __restrict__ out,
                                                      const double* __restrict__ twf_all,
                                                      const PrimeF* __restrict__ pfs, LimbMap map, int l, int pidx_p,
                                                      ModDownConst mc, GaloisList galois, int batch) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds_u[];
    double* lds = reinterpret_cast<double*>(lds_u);
    const int tid = threadIdx.x;
    const int i = wave_uniform(blockIdx.x % l);
    const int row = wave_uniform(blockIdx.x / l);                 // ((r batch + b) 2 + poly)
    const int poly = row & 1, rb = row >> 1;
    const int b = rb % batch, r = rb / batch;
    const int p = map.idx[i];
    const PrimeF pf = pfs[p];
    const double qp = pfs[pidx_p].q, hp = __builtin_floor(0.5 * qp);
    const uint64_t* arow = acc + (size_t)row * (l + 1) * C::N;
    double x[C::E];
    nttf::load_words<C>(x, arow + (size_t)l * C::N, tid);
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
#pragma clang fp contract(off)
        double v = x[e] + hp;                                      // (x_p + h) mod p, in [0, p)
        v = v >= qp ? v - qp : v;
        x[e] = rmod(v - mc.hpf[i], pf.qinv, pf.nq);               // w mod q_i
    }
    nttf::fwd_ntt<C>(x, lds, twf_all + (size_t)p * 2 * C::N, pf, tid);
    __syncthreads();
    ntt::to_lds<C, C::NP - 1>(x, lds, tid);        // NTT_i(w) at its evaluation slots
    __syncthreads();
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = lds[ntt::pad<C>(e * C::T + tid)];
    const uint64_t* X = arow + (size_t)i * C::N;
    uint64_t* o = out + ((size_t)row * l + i) * C::N;
    double y[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
#pragma clang fp contract(off)
        y[e] = mulw(u2d(X[e * C::T + tid]) - x[e], mc.pinvf[i], mc.pinvf_q[i], pf.nq);
    }
    if (poly == 0) {
        const uint64_t* c0 = Cn + ((size_t)b * 2 * l + i) * C::N;
        const uint32_t g = galois.g[r];
        __syncthreads();
#pragma unroll
        for (int e = 0; e < C::E; ++e) lds[ntt::pad<C>(e * C::T + tid)] = u2d(c0[e * C::T + tid]);
        __syncthreads();
#pragma unroll
        for (int e = 0; e < C::E; ++e) {
#pragma clang fp contract(off)
            y[e] = y[e] + lds[ntt::pad<C>(eval_perm(e * C::T + tid, g, C::LOGN))];
        }
    }
#pragma unroll
    for (int e = 0; e < C::E; ++e) o[e * C::T + tid] = d2u(canon(y[e], pf.qinv, pf.nq));
}

// out[row][j(i)] = in[row][i]: the key rows permuted by pi_g^-1.
template <typename W>
__global__ void k_key_permute(const W* __restrict__ in, W* __restrict__ out, uint32_t g, int logN, size_t total) {
    const size_t gi = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total) return;
    const size_t N = (size_t)1 << logN;
    const int i = (int)(gi & (N - 1));
    out[(gi - i) + eval_perm(i, g, logN)] = in[gi];
}

template <class K>
hipError_t set_lds(K kernel, int bytes) {
    return hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

template <int LOGN, int T>
hipError_t run_modup(const DevTables& t, const uint64_t* ct, uint64_t* D, int batch, int l, const LimbMap& map,
                     int pidx_p, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    if (t.fp64) {
        hipError_t e = set_lds(k_modup_f<LOGN, T>, C::LDS_BYTES);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_modup_f<LOGN, T>), dim3(batch * l * (l + 1)), dim3(T), C::LDS_BYTES, s, ct, D, t.twf,
                           t.pf, map, l, pidx_p, t.narrow ? 0 : 1);
        return hipGetLastError();
    }
    hipError_t e = set_lds(k_modup<LOGN, T>, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_modup<LOGN, T>), dim3(batch * l * (l + 1)), dim3(T), C::LDS_BYTES, s, ct, D, t.tw, t.pc,
                       map, l, pidx_p, t.narrow ? 0 : 1);
    return hipGetLastError();
}

template <int LOGN, int T>
hipError_t run_ks_hoisted(const DevTables& t, const uint64_t* D, const void* keys, size_t key_stride,
                          int key_limbs, const GaloisList& galois, uint64_t* acc, int batch, int l, const LimbMap& map,
                          int pidx_p, int R, int eval_out, HoistOrder order, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    constexpr int lds = ntt::pair_pays<C>() ? 2 * C::LDS_BYTES : C::LDS_BYTES;
    if (!t.fp64) {
        auto k = eval_out ? k_ks_hoisted_i<LOGN, T, true> : k_ks_hoisted_i<LOGN, T, false>;
        hipError_t e = set_lds(k, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k, dim3((l + 1) * batch * R), dim3(T), lds, s, D, (const ulonglong2*)keys, key_stride,
                           key_limbs, galois, acc, t.tw, t.pc, map, l, pidx_p, batch, R, order);
        return hipGetLastError();
    }
    auto k = eval_out ? k_ks_hoisted<LOGN, T, true> : k_ks_hoisted<LOGN, T, false>;
    hipError_t e = set_lds(k, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k, dim3((l + 1) * batch * R), dim3(T), lds, s, D, (const double*)keys, key_stride, key_limbs,
                       galois, acc, t.twf, t.pf, map, l, pidx_p, batch, R, order);
    return hipGetLastError();
}

template <int LOGN, int T>
hipError_t run_moddown_eval(const DevTables& t, const uint64_t* acc, const uint64_t* Cn, uint64_t* out, int batch,
                            int l, const LimbMap& map, int pidx_p, const ModDownConst& mc, const GaloisList& galois,
                            int R, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    if (t.fp64) {
        hipError_t e = set_lds(k_moddown_eval_f<LOGN, T>, C::LDS_BYTES);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_moddown_eval_f<LOGN, T>), dim3(R * batch * 2 * l), dim3(T), C::LDS_BYTES, s, acc, Cn,
                           out, t.twf, t.pf, map, l, pidx_p, mc, galois, batch);
        return hipGetLastError();
    }
    hipError_t e = set_lds(k_moddown_eval<LOGN, T>, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_moddown_eval<LOGN, T>), dim3(R * batch * 2 * l), dim3(T), C::LDS_BYTES, s, acc, Cn, out,
                       t.tw, t.pc, map, l, pidx_p, mc, galois, batch, t.narrow);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_modup(const DevTables& t, const uint64_t* ct, uint64_t* D, int batch, int l, const LimbMap& map,
                        int pidx_p, int threads, hipStream_t s) {
    if (!batch) return hipSuccess;
    switch (t.logN * 10000 + (threads ? threads : (t.logN == 12 ? 256 : t.logN == 13 ? 1024 : 512))) {
        case 120256: return run_modup<12, 256>(t, ct, D, batch, l, map, pidx_p, s);
        case 120512: return run_modup<12, 512>(t, ct, D, batch, l, map, pidx_p, s);
        case 130256: return run_modup<13, 256>(t, ct, D, batch, l, map, pidx_p, s);
        case 130512: return run_modup<13, 512>(t, ct, D, batch, l, map, pidx_p, s);
        case 131024: return run_modup<13, 1024>(t, ct, D, batch, l, map, pidx_p, s);
        case 140512: return run_modup<14, 512>(t, ct, D, batch, l, map, pidx_p, s);
        case 141024: return run_modup<14, 1024>(t, ct, D, batch, l, map, pidx_p, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_ks_hoisted(const DevTables& t, const uint64_t* D, const void* keys, size_t key_stride,
                             int key_limbs, const GaloisList& galois, uint64_t* acc, int batch, int l,
                             const LimbMap& map, int pidx_p, int R, int eval_out, HoistOrder order, int threads,
                             hipStream_t s) {
    if (!batch || !R) return hipSuccess;
    if (order.btile == 0) {
        // Measured at N = 8192, l = 4, 64 ciphertexts: with the inverse
        // transforms in the workgroup (coefficient form), tiles of 4
        // ciphertexts dealt out per XCD beat one key slice across the batch
        // by 8-10% (24 and 15 rotations); without them (evaluation form)
        // the key slices dominate and the whole-batch walk is best.
        order = eval_out ? HoistOrder{batch, 0} : HoistOrder{4, 1};
    }
    if (order.btile > batch) order.btile = batch;
    switch (t.logN * 10000 + (threads ? threads : (t.logN == 12 ? 256 : t.logN == 13 ? 1024 : 512))) {
        case 120256: return run_ks_hoisted<12, 256>(t, D, keys, key_stride, key_limbs, galois, acc, batch, l, map, pidx_p, R, eval_out, order, s);
        case 120512: return run_ks_hoisted<12, 512>(t, D, keys, key_stride, key_limbs, galois, acc, batch, l, map, pidx_p, R, eval_out, order, s);
        case 130256: return run_ks_hoisted<13, 256>(t, D, keys, key_stride, key_limbs, galois, acc, batch, l, map, pidx_p, R, eval_out, order, s);
        case 130512: return run_ks_hoisted<13, 512>(t, D, keys, key_stride, key_limbs, galois, acc, batch, l, map, pidx_p, R, eval_out, order, s);
        case 131024: return run_ks_hoisted<13, 1024>(t, D, keys, key_stride, key_limbs, galois, acc, batch, l, map, pidx_p, R, eval_out, order, s);
        case 140512: return run_ks_hoisted<14, 512>(t, D, keys, key_stride, key_limbs, galois, acc, batch, l, map, pidx_p, R, eval_out, order, s);
        case 141024: return run_ks_hoisted<14, 1024>(t, D, keys, key_stride, key_limbs, galois, acc, batch, l, map, pidx_p, R, eval_out, order, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_moddown_eval(const DevTables& t, const uint64_t* acc, const uint64_t* Cn, uint64_t* out, int batch,
                               int l, const LimbMap& map, int pidx_p, const ModDownConst& mc,
                               const GaloisList& galois, int R, int threads, hipStream_t s) {
    if (!batch || !R) return hipSuccess;
    switch (t.logN * 10000 + (threads ? threads : (t.logN == 12 ? 256 : t.logN == 13 ? 1024 : 512))) {
        case 120256: return run_moddown_eval<12, 256>(t, acc, Cn, out, batch, l, map, pidx_p, mc, galois, R, s);
        case 120512: return run_moddown_eval<12, 512>(t, acc, Cn, out, batch, l, map, pidx_p, mc, galois, R, s);
        case 130256: return run_moddown_eval<13, 256>(t, acc, Cn, out, batch, l, map, pidx_p, mc, galois, R, s);
        case 130512: return run_moddown_eval<13, 512>(t, acc, Cn, out, batch, l, map, pidx_p, mc, galois, R, s);
        case 131024: return run_moddown_eval<13, 1024>(t, acc, Cn, out, batch, l, map, pidx_p, mc, galois, R, s);
        case 140512: return run_moddown_eval<14, 512>(t, acc, Cn, out, batch, l, map, pidx_p, mc, galois, R, s);
        case 141024: return run_moddown_eval<14, 1024>(t, acc, Cn, out, batch, l, map, pidx_p, mc, galois, R, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_key_permute(const DevTables& t, const void* in, void* out, int rows, uint32_t g, hipStream_t s) {
    const size_t total = (size_t)rows << t.logN;
    if (!total) return hipSuccess;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (t.fp64)
        hipLaunchKernelGGL(k_key_permute<double>, grid, dim3(256), 0, s, (const double*)in, (double*)out, g, t.logN,
                           total);
    else
        hipLaunchKernelGGL(k_key_permute<ulonglong2>, grid, dim3(256), 0, s, (const ulonglong2*)in, (ulonglong2*)out,
                           g, t.logN, total);
    return hipGetLastError();
}
