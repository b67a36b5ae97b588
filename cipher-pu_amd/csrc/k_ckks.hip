// CKKS evaluation kernels: tensor product, hybrid key switch (ModUp fused
// with the digit NTTs and the key inner product, then the inverse
// transforms), ModDown + rescale, standalone rescale, Galois automorphism
// and the plain noise monitor.
//
// Layouts (uint64, row-major, limb i reduced mod prime map.idx[i]):
//   coefficient / evaluation tensors  [batch][poly][limb][N]
//   evaluation order = bit-reversed NTT output order (as written by the
//   forward transform)
//   key  [digit][2][key_limbs][N] of (value, Shoup companion) pairs,
//        evaluation order; digit j = (b_j, a_j), last limb = special prime
#include <string.h>

#include "launch.h"
#include "ntt_core.h"
#include "ntt_f64.h"

namespace {

using ntt::Cfg;

// Tensor product in the evaluation domain (verify_e2e_mult.py:78-92 /
// rtl/m4/e2e_ckks_mult_ctrl_m4.sv S_MUL_*):
//   d0 = a0 b0, d1 = a0 b1 + a1 b0, d2 = a1 b1.
// Writes p*d0, p*d1 (evaluation form, ready to be folded into the key
// switch accumulators), d2 in evaluation form, and d2 back in coefficient
// form (the key-switch digits).  FP64 pipeline (modf64.h); every word
// written is canonical.  Operands below q, so each data product is exact.
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_tensor_intt(const uint64_t* __restrict__ ea, const uint64_t* __restrict__ eb,
                                                   uint64_t* __restrict__ d01, uint64_t* __restrict__ d2e,
                                                   uint64_t* __restrict__ d2c, const double* __restrict__ twf_all,
                                                   const PrimeF* __restrict__ pfs, LimbMap map, int l,
                                                   ModDownConst mc) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds_u[];
    double* lds = reinterpret_cast<double*>(lds_u);
    const int tid = threadIdx.x;
    const int b = wave_uniform(blockIdx.x / l), i = wave_uniform(blockIdx.x % l);
    const int p = map.idx[i];
    const PrimeF pf = pfs[p];
    const double qi = pf.qinv, nq = pf.nq;
    const size_t LN = (size_t)l * C::N;
    const size_t ct0 = (size_t)b * 2 * LN + (size_t)i * C::N;  // poly 0, limb i
    const double pm = mc.pmf[i], pmq = mc.pmf_q[i];
    double x[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const size_t j = (size_t)e * C::T + tid;
        const double a0 = u2d(ea[ct0 + j]), a1 = u2d(ea[ct0 + LN + j]);
        const double b0 = u2d(eb[ct0 + j]), b1 = u2d(eb[ct0 + LN + j]);
        const double d1 = muld(a0, b1, qi, nq) + muld(a1, b0, qi, nq);
        const double d2 = muld(a1, b1, qi, nq);
        d01[ct0 + j] = d2u(canon(mulw(muld(a0, b0, qi, nq), pm, pmq, nq), qi, nq));
        d01[ct0 + LN + j] = d2u(canon(mulw(d1, pm, pmq, nq), qi, nq));
        d2e[(size_t)b * LN + (size_t)i * C::N + j] = d2u(canon(d2, qi, nq));
        x[e] = d2;
    }
    // to the last forward pass's layout, through LDS
#pragma unroll
    for (int e = 0; e < C::E; ++e) lds[ntt::pad<C>(e * C::T + tid)] = x[e];
    __syncthreads();
    ntt::from_lds<C, C::NP - 1>(x, lds, tid);
    __syncthreads();
    nttf::inv_ntt<C>(x, lds, twf_all + ((size_t)p * 2 + 1) * C::N, pf, tid);
    nttf::store_words<C>(x, d2c + (size_t)b * LN + (size_t)i * C::N, pf, tid);
}

// One (pair, limb) of the tensor step from the four forward transforms
// (bounded by 16 q): p d0, p d1 and d2 written in evaluation order at
// register k's slot; d2 returned (bounded by 3q/4) for the inverse transform.
template <class C>
__device__ __forceinline__ double tensor_one(double x0, double x1, double y0, double y1, uint64_t* __restrict__ o0,
                                             uint64_t* __restrict__ o2, size_t LN, int idx, const PrimeF& pf,
                                             double pm, double pmq) {
#pragma clang fp contract(off)
    const double qi = pf.qinv, nq = pf.nq;
    const double a0 = rmod(x0, qi, nq), a1 = rmod(x1, qi, nq);
    const double b0 = rmod(y0, qi, nq), b1 = rmod(y1, qi, nq);
    o0[idx] = d2u(canon(mulw(muld(a0, b0, qi, nq), pm, pmq, nq), qi, nq));
    const double d1 = muld(a1, b0, qi, nq) + muld(a0, b1, qi, nq);
    o0[LN + idx] = d2u(canon(mulw(d1, pm, pmq, nq), qi, nq));
    const double d2 = muld(a1, b1, qi, nq);
    o2[idx] = d2u(canon(d2, qi, nq));
    return d2;
}

// The whole front of HMult for one (pair, limb i), coefficient form in:
// forward transforms of a0, a1, b0, b1, the tensor product, p*d0 and p*d1
// and d2 written in evaluation form, d2 transformed back to coefficient form.
// FP64 pipeline: the forward transforms leave values below
// (1 + 7 log2(N)/8) q, reduced before the data products.
// Items it = limb * batch + pair, limb-major (items in flight together share
// one prime's twiddles).  With paired transforms each workgroup loops over
// items with a stride of the grid, which run_front sizes to the workgroups
// resident at once: the next item's a0 is then loaded during the current
// item's last transform instead of stalling its start.  (All workgroups
// start their items together, so loads issued at an item's start arrive as
// one burst over the whole device.)  SQ (squaring, B == A; paired shapes):
// the transforms of b0, b1 are those of a0, a1, so they are not repeated and
// the tensor product reads the same values twice: the same words as the
// general kernel with B = A.
template <int LOGN, int T, bool SQ = false>
__global__ void __launch_bounds__(T) k_front(const uint64_t* __restrict__ A, const uint64_t* __restrict__ B,
                                             uint64_t* __restrict__ d01, uint64_t* __restrict__ d2e,
                                             uint64_t* __restrict__ d2c, const double* __restrict__ twf_all,
                                             const PrimeF* __restrict__ pfs, LimbMap map, int l, int batch,
                                             ModDownConst mc) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds_u[];
    double* lds = reinterpret_cast<double*>(lds_u);
    const int tid = threadIdx.x;
    const size_t LN = (size_t)l * C::N;
    if constexpr (ntt::pair_pays<C>()) {
        // a0 with a1 and b0 with b1 as transform pairs
        const int items = l * batch;
        double x0[C::E], x1[C::E], y0[C::E], y1[C::E];
        uint64_t na0[C::E], nb0[C::E];
        int it = (int)blockIdx.x;
        if (it < items) {
            const int i = wave_uniform(it / batch), b = wave_uniform(it % batch);
            nttf::issue_row<C>(na0, A + (size_t)b * 2 * LN + (size_t)i * C::N, tid);
        }
        for (; it < items; it += (int)gridDim.x) {
            const int i = wave_uniform(it / batch), b = wave_uniform(it % batch);
            const int p = map.idx[i];
            const PrimeF pf = pfs[p];
            const double* tw = twf_all + (size_t)p * 2 * C::N;
            const size_t c0 = (size_t)b * 2 * LN + (size_t)i * C::N;   // poly 0 of pair b, limb i
            const double pm = mc.pmf[i], pmq = mc.pmf_q[i];
            nttf::load_words<C>(x1, A + c0 + LN, tid);
#pragma unroll
            for (int k = 0; k < C::E; ++k) x0[k] = u2d(na0[k]);
            // b0 is loaded ahead, in flight while the first passes (scalar
            // twiddle loads only) transform a0 and a1
            if constexpr (!SQ) nttf::issue_row<C>(nb0, B + c0, tid);
            nttf::fwd_ntt2<C>(x0, x1, lds, tw, pf, tid);
            if constexpr (SQ) {
#pragma unroll
                for (int k = 0; k < C::E; ++k) {
                    y0[k] = x0[k];
                    y1[k] = x1[k];
                }
            } else {
#pragma unroll
                for (int k = 0; k < C::E; ++k) y0[k] = u2d(nb0[k]);
                nttf::load_words<C>(y1, B + c0 + LN, tid);
                nttf::fwd_ntt2<C>(y0, y1, lds, tw, pf, tid);
            }
            // the inverse transform's first twiddles, loaded before the stores
            const auto pre = nttf::inv_tw<C, C::NP - 1>(tw + C::N, tid);
            uint64_t* o0 = d01 + c0;
            uint64_t* o2 = d2e + (size_t)b * LN + (size_t)i * C::N;
#pragma unroll
            for (int k = 0; k < C::E; ++k)
                x0[k] = tensor_one<C>(x0[k], x1[k], y0[k], y1[k], o0, o2, LN, ntt::eval_index<C>(k, tid), pf, pm,
                                      pmq);
            // the next item's a0, issued where the inverse transform has no
            // vector twiddle load left to wait behind it
            const int next = it + (int)gridDim.x;
            auto prefetch = [&]() {
                if (next < items) {
                    const int ni = wave_uniform(next / batch), nb = wave_uniform(next % batch);
                    nttf::issue_row<C>(na0, A + (size_t)nb * 2 * LN + (size_t)ni * C::N, tid);
                }
            };
            nttf::inv_ntt_db<C>(x0, lds, tw + C::N, pf, tid, pre, prefetch);
            nttf::store_words<C>(x0, d2c + (size_t)b * LN + (size_t)i * C::N, pf, tid);
            // other waves may still read the LDS: the next item's transform
            // pair starts with a workgroup barrier before its first write
        }
        return;
    }
    const int i = wave_uniform(blockIdx.x / batch), b = wave_uniform(blockIdx.x % batch);
    const int p = map.idx[i];
    const PrimeF pf = pfs[p];
    const double* tw = twf_all + (size_t)p * 2 * C::N;
    const size_t c0 = (size_t)b * 2 * LN + (size_t)i * C::N;   // poly 0 of pair b, limb i
    const double pm = mc.pmf[i], pmq = mc.pmf_q[i];
    uint64_t* o0 = d01 + c0;
    uint64_t* o2 = d2e + (size_t)b * LN + (size_t)i * C::N;
    double x0[C::E], x1[C::E], y[C::E], z[C::E];
    const double qi = pf.qinv, nq = pf.nq;
    nttf::load_words<C>(x0, A + c0, tid);
    nttf::fwd_ntt<C>(x0, lds, tw, pf, tid);
    nttf::load_words<C>(x1, A + c0 + LN, tid);
    nttf::fwd_ntt<C>(x1, lds, tw, pf, tid);
    nttf::load_words<C>(y, B + c0, tid);
    nttf::fwd_ntt<C>(y, lds, tw, pf, tid);
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const double a0 = rmod(x0[k], qi, nq), a1 = rmod(x1[k], qi, nq), b0 = rmod(y[k], qi, nq);
        o0[ntt::eval_index<C>(k, tid)] = d2u(canon(mulw(muld(a0, b0, qi, nq), pm, pmq, nq), qi, nq));
        z[k] = muld(a1, b0, qi, nq);
        x0[k] = a0;
        x1[k] = a1;
    }
    nttf::load_words<C>(y, B + c0 + LN, tid);
    nttf::fwd_ntt<C>(y, lds, tw, pf, tid);
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
#pragma clang fp contract(off)
        const double b1 = rmod(y[k], qi, nq);
        const int idx = ntt::eval_index<C>(k, tid);
        const double d1 = z[k] + muld(x0[k], b1, qi, nq);
        o0[LN + idx] = d2u(canon(mulw(d1, pm, pmq, nq), qi, nq));
        const double d2 = muld(x1[k], b1, qi, nq);
        o2[idx] = d2u(canon(d2, qi, nq));
        x0[k] = d2;
    }
    nttf::inv_ntt<C>(x0, lds, tw + C::N, pf, tid);
    nttf::store_words<C>(x0, d2c + (size_t)b * LN + (size_t)i * C::N, pf, tid);
}

// a0 += v * b_j,t and a1 -= v * a_j,t for digit j in evaluation order (key
// rows of the words as doubles, 8 bytes a word: the products form their
// quotients, mulk).  v below 2^51; each product is bounded by 7q/8, so with
// at most 15 digits and a start below q the sums stay below 15 q.
template <class C>
__device__ __forceinline__ void key_mac(double (&a0)[C::E], double (&a1)[C::E], const double (&v)[C::E],
                                        const double* __restrict__ key, int j, int key_limbs, int kt, double qinv,
                                        double nq, int tid) {
    const double* kb = key + ((size_t)(j * 2 + 0) * key_limbs + kt) * C::N;
    const double* ka = key + ((size_t)(j * 2 + 1) * key_limbs + kt) * C::N;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
#pragma clang fp contract(off)
        const int idx = ntt::eval_index<C>(k, tid);
        a0[k] = a0[k] + mulk(v[k], kb[idx], qinv, nq);
        a1[k] = a1[k] - mulk(v[k], ka[idx], qinv, nq);
    }
}

// ModDown by p and rescale by q_(l-1) of one coefficient of chain limb t <
// l - 1 on the FP64 pipeline: the per-limb steps of k_moddown_rescale_l,
//   y_i = ((x_i + h_p) - [(x_p + h_p) mod p]) p^-1 mod q_i,  i = t and l - 1,
//   out = ((y_t + h_q) - [(y_(l-1) + h_q) mod q_(l-1)]) q_(l-1)^-1 mod q_t,
// with h_p = (p - 1) / 2, h_q = (q_(l-1) - 1) / 2 (and their residues).  The
// two bracketed lifts are exact canonical integers; every other value is a
// signed residue, so no reduction is needed before the products (each
// product input is below 3 * 2^45).  xt: the target's accumulator (bounded
// by 5q/8); xl, xp: the canonical accumulators of limb l - 1 and of p.
__device__ __forceinline__ uint64_t moddown_rescale_one(double xt, double xl, double xp, int t, int lm1,
                                                        const PrimeF& pc, const PrimeF& pl, const PrimeF& pp,
                                                        const ModDownConst& mc) {
#pragma clang fp contract(off)
    double tp = xp + __builtin_floor(0.5 * pp.q);       // (x_p + h_p) mod p in [0, p)
    tp = tp >= pp.q ? tp - pp.q : tp;
    const double yl = mulw(xl + mc.hpf[lm1] - tp, mc.pinvf[lm1], mc.pinvf_q[lm1], pl.nq);
    const double tl = canon(yl + __builtin_floor(0.5 * pl.q), pl.qinv, pl.nq);   // (y_l + h_q) mod q_l
    const double yt = mulw(xt + mc.hpf[t] - tp, mc.pinvf[t], mc.pinvf_q[t], pc.nq);
    return d2u(canon(mulw(yt + mc.hqf[t] - tl, mc.qinvf[t], mc.qinvf_q[t], pc.nq), pc.qinv, pc.nq));
}

// The fused epilogue for poly c of target t: x holds the accumulator in
// coefficient form (coalesced layout, bounded by 5q/8); the special-prime
// and limb l - 1 accumulators come from an earlier launch through acc_c
// ([l+1][N], canonical words).
template <class C>
__device__ __forceinline__ void fused_out(const double (&x)[C::E], const uint64_t* __restrict__ acc_c,
                                          uint64_t* __restrict__ out_c, int t, int l,
                                          const PrimeF* __restrict__ pfs, const LimbMap& map, int pidx_p,
                                          const ModDownConst& mc, int tid) {
    const PrimeF pc = pfs[map.idx[t]], pl = pfs[map.idx[l - 1]], pp = pfs[pidx_p];
    const uint64_t* sp = acc_c + (size_t)l * C::N;
    const uint64_t* sl = acc_c + (size_t)(l - 1) * C::N;
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int j = e * C::T + tid;
        out_c[j] = moddown_rescale_one(x[e], u2d(sl[j]), u2d(sp[j]), t, l - 1, pc, pl, pp, mc);
    }
}

// fused_out with the rows of limb l - 1 and of p already loaded (coalesced
// layout, register e holding word e T + tid).
template <class C>
__device__ __forceinline__ void fused_out_rows(const double (&x)[C::E], const uint64_t (&rl)[C::E],
                                               const uint64_t (&rp)[C::E], uint64_t* __restrict__ out_c, int t,
                                               int l, const PrimeF* __restrict__ pfs, const LimbMap& map,
                                               int pidx_p, const ModDownConst& mc, int tid) {
    const PrimeF pc = pfs[map.idx[t]], pl = pfs[map.idx[l - 1]], pp = pfs[pidx_p];
#pragma unroll
    for (int e = 0; e < C::E; ++e)
        out_c[e * C::T + tid] = moddown_rescale_one(x[e], u2d(rl[e]), u2d(rp[e]), t, l - 1, pc, pl, pp, mc);
}

// Hybrid key switch for one (ciphertext, target limb t) pair, t = 0..l where
// t = l is the special prime p.  Generalises the single-digit relinearisation
// f0 = e0 + e2 b, f1 = e1 - e2 a (verify_keyswitch.py:62-68,
// rtl/m3/keyswitch_pipe.sv:90-100) to one digit per RNS prime:
//   acc0_t = [p d0]_t + sum_j NTT_t([d2]_(q_j)) * b_j,t
//   acc1_t = [p d1]_t - sum_j NTT_t([d2]_(q_j)) * a_j,t
// (FP64 pipeline: signed residues, key rows of the words as doubles)
// then both accumulators go back to coefficient form.  The digit lift
// ("ModUp") is the reduction of the digit's coefficients mod the target
// prime on load; the digit NTTs, the key products and the inverse
// transforms never leave the workgroup.
// KEYTAB: ciphertext b switches with its own key, key_table[b] (a device
// address), instead of the one key of the launch (rotations by different
// steps in one launch).
// Targets t0 .. t0 + gridDim.x / batch - 1 in one launch.  FUSED (HMult,
// chain limbs t < l - 1 only): instead of writing its accumulators, the
// workgroup finishes ModDown and the rescale for its limb, reading the
// coefficient-form accumulators of limb l - 1 and of the special prime
// that an earlier launch (targets l - 1 and l) left in acc, and writes the
// output ciphertext limb [batch][2][l-1][N] to out.
template <int LOGN, int T, bool KEYTAB = false, bool FUSED = false>
__global__ void __launch_bounds__(T) k_keyswitch(const uint64_t* __restrict__ d2c, size_t d2c_bstride,
                                                 const uint64_t* __restrict__ d2e,
                                                 const uint64_t* __restrict__ d01, const double* __restrict__ key0,
                                                 int key_limbs, uint64_t* __restrict__ acc,
                                                 const double* __restrict__ twf_all, const PrimeF* __restrict__ pfs,
                                                 LimbMap map, int l, int pidx_p, int batch, int reduce_digits,
                                                 const uint64_t* __restrict__ key_table, int t0,
                                                 uint64_t* __restrict__ out_ct, ModDownConst mc, int pair_major,
                                                 int items) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds_u[];
    double* lds = reinterpret_cast<double*>(lds_u);
    const int tid = threadIdx.x;
    const int nt = items / batch;
    const bool own = d2e != nullptr;   // digit t comes in evaluation form
    const size_t LN = (size_t)l * C::N;
    // The coefficient-form digits are loaded one ahead: the next digit's
    // loads are issued as the current digit's transform starts, whose first
    // passes take their twiddles through scalar loads, so the loads are in
    // flight for those passes before any vector load must wait for them.
    // (16 coefficients per thread: the extra row would spill.)
    constexpr bool PF = C::E <= 8;
    // Items v = 0 .. items - 1, one (target, ciphertext) each; with paired
    // transforms a workgroup loops over them with a stride of the grid
    // (run_ks sizes it to the workgroups resident at once), and the next
    // item's first digit is loaded during the current item's scalar-twiddle
    // inverse rounds.  Target-major order: items in flight together share
    // one target limb, so each XCD's L2 holds one limb's key slice and one
    // prime's twiddles.  Ciphertext-major (pair_major): item v runs on XCD
    // v mod 8 (the grid is a multiple of 8), which takes logical index
    // (v mod 8) items/8 + v/8, so each XCD walks one contiguous range of
    // ciphertexts, all targets of one ciphertext back to back.
    auto slot = [&](int v, int& t, int& b) {
        if (pair_major) {
            const int L = (items & 7) == 0 && (gridDim.x & 7) == 0 ? (v & 7) * (items >> 3) + (v >> 3) : v;
            t = wave_uniform(t0 + L % nt);
            b = wave_uniform(L / nt);
        } else {
            t = wave_uniform(t0 + v / batch);
            b = wave_uniform(v % batch);
        }
    };
    uint64_t nx[C::E];
    bool have = false;   // nx holds this item's first digit
    for (int v = (int)blockIdx.x; v < items; v += (int)gridDim.x) {
        int t, b;
        slot(v, t, b);
        const double* key = key0;
        if constexpr (KEYTAB) key = reinterpret_cast<const double*>(key_table[b]);
        const int pt = t < l ? map.idx[t] : pidx_p;
        const int kt = t < l ? t : key_limbs - 1;
        const PrimeF pf = pfs[pt];
        const double nq = pf.nq;
        const double* tw = twf_all + (size_t)pt * 2 * C::N;
        double a0[C::E], a1[C::E], x[C::E];
        const uint64_t* dig = d2c + (size_t)b * d2c_bstride;
        int jn = (own && t == 0) ? 1 : 0;
        if (PF && jn < l && !have) nttf::issue_row<C>(nx, dig + (size_t)jn * C::N, tid);   // else loaded ahead
        if (t < l && d01) {
            const uint64_t* s0 = d01 + (size_t)b * 2 * LN + (size_t)t * C::N;
#pragma unroll
            for (int k = 0; k < C::E; ++k) {
                const int j = ntt::eval_index<C>(k, tid);
                a0[k] = u2d(s0[j]);
                a1[k] = u2d(s0[LN + j]);   // acc1 = p d1 - sum d a
            }
        } else {
#pragma unroll
            for (int k = 0; k < C::E; ++k) a0[k] = a1[k] = 0.0;
        }
        for (int j = 0; j < l; ++j) {
            if (j == t && own) {
                // the digit of the target's own prime is already in evaluation form
                const uint64_t* s = d2e + (size_t)b * LN + (size_t)j * C::N;
#pragma unroll
                for (int k = 0; k < C::E; ++k) x[k] = u2d(s[ntt::eval_index<C>(k, tid)]);
            } else {
                if (!PF) nttf::issue_row<C>(nx, dig + (size_t)j * C::N, tid);
#pragma unroll
                for (int k = 0; k < C::E; ++k) x[k] = u2d(nx[k]);
                // the digit's residues mod q_j enter the transform mod q_t as they
                // are when q_j < 2 q_t (reduce_digits = 0); else reduce them first
                if (reduce_digits) {
#pragma unroll
                    for (int k = 0; k < C::E; ++k) x[k] = rmod(x[k], pf.qinv, nq);
                }
                jn = j + 1 + ((own && j + 1 == t) ? 1 : 0);
                if (PF && jn < l) nttf::issue_row<C>(nx, dig + (size_t)jn * C::N, tid);
                // left unreduced (below 15 q): the key product is exact for any such input
                if constexpr (ntt::pair_pays<C>()) nttf::fwd_ntt_db<C>(x, lds, tw, pf, tid);
                else nttf::fwd_ntt<C>(x, lds, tw, pf, tid);
            }
            key_mac<C>(a0, a1, x, key, j, key_limbs, kt, pf.qinv, nq, tid);
        }
        // the double-buffered digit transforms leave reads of the LDS in flight
        if constexpr (ntt::pair_pays<C>()) __syncthreads();
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            a0[k] = rmod(a0[k], pf.qinv, nq);
            a1[k] = rmod(a1[k], pf.qinv, nq);
        }
        uint64_t* out = acc + ((size_t)b * 2 * (l + 1) + t) * C::N;
        const uint64_t* acc_b = acc + (size_t)b * 2 * (l + 1) * C::N;
        uint64_t* out_b = FUSED ? out_ct + ((size_t)b * 2 * (l - 1) + t) * C::N : nullptr;
        const double* itw = tw + C::N;
        const int vn = v + (int)gridDim.x;
        have = PF && vn < items;
        auto next_digit = [&]() {
            if (have) {
                int tn, bn;
                slot(vn, tn, bn);
                const int j0 = (own && tn == 0) ? 1 : 0;
                nttf::issue_row<C>(nx, d2c + (size_t)bn * d2c_bstride + (size_t)j0 * C::N, tid);
            }
        };
        if constexpr (ntt::pair_pays<C>()) {
            // both accumulators as one transform pair
            if constexpr (FUSED) {
                // The epilogue's rows (limb l - 1 and p) of poly 0 are loaded
                // during the pair's last rounds, with the next item's first
                // digit; those of poly 1 before poly 0 is finished: their latency
                // overlaps work instead of following the transform.
                uint64_t rl[C::E], rp[C::E];
                auto prefetch = [&]() {
                    nttf::issue_row<C>(rl, acc_b + (size_t)(l - 1) * C::N, tid);
                    nttf::issue_row<C>(rp, acc_b + (size_t)l * C::N, tid);
                    next_digit();
                };
                nttf::inv_ntt2<C>(a0, a1, lds, itw, pf, tid, prefetch);
                const uint64_t* acc_b1 = acc_b + (size_t)(l + 1) * C::N;
                uint64_t rl1[C::E], rp1[C::E];
                nttf::issue_row<C>(rl1, acc_b1 + (size_t)(l - 1) * C::N, tid);
                nttf::issue_row<C>(rp1, acc_b1 + (size_t)l * C::N, tid);
                fused_out_rows<C>(a0, rl, rp, out_b, t, l, pfs, map, pidx_p, mc, tid);
                fused_out_rows<C>(a1, rl1, rp1, out_b + (size_t)(l - 1) * C::N, t, l, pfs, map, pidx_p, mc, tid);
            } else {
                nttf::inv_ntt2<C>(a0, a1, lds, itw, pf, tid, next_digit);
                nttf::store_words<C>(a0, out, pf, tid);
                nttf::store_words<C>(a1, out + (size_t)(l + 1) * C::N, pf, tid);
            }
            // the next item's first transform starts with a workgroup barrier
            // before its first LDS write
            continue;
        }
        have = false;
        nttf::inv_ntt<C>(a0, lds, itw, pf, tid);
        if constexpr (FUSED) fused_out<C>(a0, acc_b, out_b, t, l, pfs, map, pidx_p, mc, tid);
        else nttf::store_words<C>(a0, out, pf, tid);
        __syncthreads();
        nttf::inv_ntt<C>(a1, lds, itw, pf, tid);
        if constexpr (FUSED)
            fused_out<C>(a1, acc_b + (size_t)(l + 1) * C::N, out_b + (size_t)(l - 1) * C::N, t, l, pfs, map, pidx_p, mc,
                         tid);
        else nttf::store_words<C>(a1, out + (size_t)(l + 1) * C::N, pf, tid);
        __syncthreads();
    }
}

// ModDown by p (round-half-up of X/p), optional addition of coefficient-form
// (d0, d1), then optional rescale by the last prime (verify_rescale.py:38-53
// on the CRT value, in the per-limb form
//   ((y_i + h) - [(y_L + h) mod q_L]) q_L^-1 mod q_i,  h = (q_L - 1)/2).
// narrow = 1: every prime is below twice every other one, so a residue mod
// one prime is reduced mod another by one conditional subtraction.  dadd
// (optional, [batch][2][l][N] like dcoef) is added to both polynomials: a
// rotation's input, for out = ct + rot(ct) without rescale.
__global__ void k_moddown_rescale(const uint64_t* __restrict__ acc, const uint64_t* __restrict__ dcoef, int dmask,
                                  const uint64_t* __restrict__ dadd, uint64_t* __restrict__ out,
                                  const PrimeConst* __restrict__ pcs, LimbMap map, int l, int pidx_p,
                                  ModDownConst mc, int rescale, int narrow, int logN, size_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const size_t N = (size_t)1 << logN;
    const size_t n = g & (N - 1);
    const size_t bp = g >> logN;              // (batch, poly) row
    const uint64_t* in = acc + bp * (l + 1) * N + n;
    const PrimeConst& pp = pcs[pidx_p];
    const uint64_t tp = csub(in[(size_t)l * N] + (pp.q >> 1), pp.q);
    const int lo = rescale ? l - 1 : l;
    uint64_t* o = out + bp * lo * N + n;
    const uint64_t* dc = (dcoef && ((dmask >> (bp & 1)) & 1)) ? dcoef + bp * l * N + n : nullptr;
    const uint64_t* da = dadd ? dadd + bp * l * N + n : nullptr;
    auto moddown = [&](int i) -> uint64_t {
        const PrimeConst& pc = pcs[map.idx[i]];
        const uint64_t tpi = narrow ? csub(tp, pc.q) : reduce64(tp, pc);
        uint64_t v = sub_mod(add_mod(in[(size_t)i * N], mc.hp[i], pc.q), tpi, pc.q);
        v = csub(mul_shoup(v, mc.pinv[i], mc.pinv_s[i], pc.q), pc.q);
        if (dc) v = add_mod(v, dc[(size_t)i * N], pc.q);
        if (da) v = add_mod(v, da[(size_t)i * N], pc.q);
        return v;
    };
    if (!rescale) {
        for (int i = 0; i < l; ++i) o[(size_t)i * N] = moddown(i);
        return;
    }
    const PrimeConst& pl = pcs[map.idx[l - 1]];
    const uint64_t tl = csub(moddown(l - 1) + (pl.q >> 1), pl.q);
    for (int i = 0; i < l - 1; ++i) {
        const PrimeConst& pc = pcs[map.idx[i]];
        const uint64_t tli = narrow ? csub(tl, pc.q) : reduce64(tl, pc);
        uint64_t v = sub_mod(add_mod(moddown(i), mc.hq[i], pc.q), tli, pc.q);
        o[(size_t)i * N] = csub(mul_shoup(v, mc.qinv[i], mc.qinv_s[i], pc.q), pc.q);
    }
}

// The same ModDown (+ rescale) with the limb count L fixed at compile time,
// for the chains in use (L <= 8).  Each thread takes coefficients n and
// n + N/2 of one row, so the per-limb constants (uniform, in SGPRs) and the
// index arithmetic serve two coefficients, and the loops unroll.  The
// modular steps keep lazy sums: y + h + q - t < 3q enters the Shoup product
// (exact for any input below 2^62), whose result in [0, 2q) is brought to
// [0, q) by one subtraction -- the same canonical words as the general
// kernel above.
template <int L, bool RESCALE>
__global__ void __launch_bounds__(256) k_moddown_rescale_l(const uint64_t* __restrict__ acc,
                                                           const uint64_t* __restrict__ dcoef, int dmask,
                                                           const uint64_t* __restrict__ dadd,
                                                           uint64_t* __restrict__ out,
                                                           const PrimeConst* __restrict__ pcs, LimbMap map,
                                                           int pidx_p, ModDownConst mc, int narrow, int logN,
                                                           size_t total_half) {
    constexpr int V = 2;
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total_half) return;
    const size_t N = (size_t)1 << logN, H = N >> 1;
    const size_t n = g & (H - 1);
    const size_t bp = g >> (logN - 1);        // (batch, poly) row
    const uint64_t* in = acc + bp * (L + 1) * N + n;
    const PrimeConst& pp = pcs[pidx_p];
    uint64_t tp[V];
#pragma unroll
    for (int v = 0; v < V; ++v) tp[v] = csub(in[(size_t)L * N + v * H] + (pp.q >> 1), pp.q);
    constexpr int LO = RESCALE ? L - 1 : L;
    uint64_t* o = out + bp * LO * N + n;
    const uint64_t* dc = (dcoef && ((dmask >> (bp & 1)) & 1)) ? dcoef + bp * L * N + n : nullptr;
    const uint64_t* da = dadd ? dadd + bp * L * N + n : nullptr;
    uint64_t md[L][V];
#pragma unroll
    for (int i = 0; i < L; ++i) {
        const PrimeConst& pc = pcs[map.idx[i]];
        const uint64_t q = pc.q;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const uint64_t tpi = narrow ? csub(tp[v], q) : reduce64(tp[v], pc);
            const uint64_t u = in[(size_t)i * N + v * H] + mc.hp[i] + q - tpi;   // < 3q
            uint64_t r = csub(mul_shoup(u, mc.pinv[i], mc.pinv_s[i], q), q);
            if (dc) r = add_mod(r, dc[(size_t)i * N + v * H], q);
            if (da) r = add_mod(r, da[(size_t)i * N + v * H], q);
            md[i][v] = r;
        }
    }
    if constexpr (!RESCALE) {
#pragma unroll
        for (int i = 0; i < L; ++i)
#pragma unroll
            for (int v = 0; v < V; ++v) o[(size_t)i * N + v * H] = md[i][v];
    } else {
        const PrimeConst& pl = pcs[map.idx[L - 1]];
        uint64_t tl[V];
#pragma unroll
        for (int v = 0; v < V; ++v) tl[v] = csub(md[L - 1][v] + (pl.q >> 1), pl.q);
#pragma unroll
        for (int i = 0; i < L - 1; ++i) {
            const PrimeConst& pc = pcs[map.idx[i]];
            const uint64_t q = pc.q;
#pragma unroll
            for (int v = 0; v < V; ++v) {
                const uint64_t tli = narrow ? csub(tl[v], q) : reduce64(tl[v], pc);
                const uint64_t u = md[i][v] + mc.hq[i] + q - tli;                // < 3q
                o[(size_t)i * N + v * H] = csub(mul_shoup(u, mc.qinv[i], mc.qinv_s[i], q), q);
            }
        }
    }
}

// Rescale by the last limb (verify_rescale.py:38-57, RNS per-limb form).
__global__ void k_rescale(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                          const PrimeConst* __restrict__ pcs, LimbMap map, int l, ModDownConst mc, int narrow,
                          int logN, size_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const size_t N = (size_t)1 << logN;
    const size_t n = g & (N - 1), row = g >> logN;
    const uint64_t* src = in + row * l * N + n;
    uint64_t* dst = out + row * (l - 1) * N + n;
    const PrimeConst& pl = pcs[map.idx[l - 1]];
    const uint64_t tl = csub(src[(size_t)(l - 1) * N] + (pl.q >> 1), pl.q);
    for (int i = 0; i < l - 1; ++i) {
        const PrimeConst& pc = pcs[map.idx[i]];
        const uint64_t tli = narrow ? csub(tl, pc.q) : reduce64(tl, pc);
        uint64_t v = sub_mod(add_mod(src[(size_t)i * N], mc.hq[i], pc.q), tli, pc.q);
        dst[(size_t)i * N] = csub(mul_shoup(v, mc.qinv[i], mc.qinv_s[i], pc.q), pc.q);
    }
}

// f(X) -> f(X^g) mod X^N + 1 for odd g, as a gather: output coefficient d
// takes input k = g^-1 d mod 2N, negated when k >= N
// (exp_e24_rotation_bootstrap.py:143-162 accumulates the same map).
__global__ void k_automorphism(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                               const PrimeConst* __restrict__ pcs, LimbMap map, int n_limbs, uint64_t ginv,
                               int logN, size_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const uint64_t N = 1ull << logN;
    const uint64_t d = g & (N - 1);
    const size_t row = g >> logN;
    const uint64_t k = (ginv * d) & (2 * N - 1);
    const uint64_t* src = in + row * N;
    uint64_t v;
    if (k < N) {
        v = src[k];
    } else {
        const uint64_t q = pcs[map.idx[row % n_limbs]].q;
        const uint64_t c = src[k - N];
        v = c ? q - c : 0;
    }
    out[g] = v;
}

// The same map, one workgroup per (poly, limb) row: the row is read into LDS
// with coalesced loads and written back with coalesced stores in the
// permuted order, so neither side of HBM sees the gather's stride.  (It is
// also safe in place: a row is fully staged before any of it is written.)
// Rows are (group, limb) with group g's limb i at g group_stride + i N in
// both tensors.
// gal_table (optional): group g uses the Galois element gal_table[g / gal_div]
// instead of the launch's, inverted here.
__device__ __forceinline__ uint64_t galois_inverse(uint64_t g, int logN) {
    uint64_t inv = 1;   // Newton's iteration for the inverse of an odd g modulo 2^64
    for (int k = 0; k < 6; ++k) inv = inv * (2 - g * inv);
    return inv & ((2ull << logN) - 1);
}

__global__ void __launch_bounds__(1024) k_automorphism_row(const uint64_t* in, uint64_t* out,
                                                           const PrimeConst* __restrict__ pcs, LimbMap map,
                                                           int n_limbs, size_t group_stride, uint64_t ginv,
                                                           int logN, const uint64_t* __restrict__ gal_table,
                                                           int gal_div) {
    extern __shared__ uint64_t lds[];
    const int N = 1 << logN;
    const int r = blockIdx.x;
    if (gal_table) ginv = galois_inverse(gal_table[r / n_limbs / gal_div], logN);
    const size_t off = (size_t)(r / n_limbs) * group_stride + (size_t)(r % n_limbs) * N;
    const uint64_t* src = in + off;
    for (int k = threadIdx.x; k < N; k += blockDim.x) lds[k] = src[k];
    __syncthreads();
    const uint64_t q = pcs[map.idx[r % n_limbs]].q;
    const uint64_t mask = 2 * (uint64_t)N - 1;
    uint64_t* dst = out + off;
    for (int d = threadIdx.x; d < N; d += blockDim.x) {
        const uint64_t k = (ginv * (uint64_t)d) & mask;
        uint64_t v;
        if (k < (uint64_t)N) {
            v = lds[k];
        } else {
            const uint64_t c = lds[k - N];
            v = c ? q - c : 0;
        }
        dst[d] = v;
    }
}

// Plain noise monitor per (poly, limb) row: with c' = min(c, q - c)
// (rtl/m4/noise_monitor.sv:44-78, verify_m4_h_multilayer.py:162-174) the row
// reports sum c', sum c'^2 (128-bit), max c' and the sum of the centered
// values, all exact.  stats: [row][5] = {sum_abs, sumsq_lo, sumsq_hi,
// max_abs, sum_centered (two's complement)}.
// One row of the monitor: block-wide exact sums over n coefficients mod q.
__device__ __forceinline__ void noise_row(const uint64_t* __restrict__ src, int n, uint64_t q,
                                          uint64_t* __restrict__ o) {
    __shared__ uint64_t s_abs[256], s_lo[256], s_hi[256], s_max[256], s_sgn[256];
    uint64_t sa = 0, lo = 0, hi = 0, mx = 0, sg = 0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const uint64_t c = src[k];
        const bool neg = c > q - c;
        const uint64_t a = neg ? q - c : c;
        sa += a;
        mx = a > mx ? a : mx;
        sg += neg ? (uint64_t)0 - a : a;
        const uint64_t plo = a * a, phi = __umul64hi(a, a);
        lo += plo;
        hi += phi + (lo < plo ? 1 : 0);
    }
    const int t = threadIdx.x;
    s_abs[t] = sa; s_lo[t] = lo; s_hi[t] = hi; s_max[t] = mx; s_sgn[t] = sg;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            s_abs[t] += s_abs[t + w];
            const uint64_t l2 = s_lo[t] + s_lo[t + w];
            s_hi[t] += s_hi[t + w] + (l2 < s_lo[t] ? 1 : 0);
            s_lo[t] = l2;
            s_max[t] = s_max[t] > s_max[t + w] ? s_max[t] : s_max[t + w];
            s_sgn[t] += s_sgn[t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        o[0] = s_abs[0]; o[1] = s_lo[0]; o[2] = s_hi[0]; o[3] = s_max[0]; o[4] = s_sgn[0];
    }
}

__global__ void __launch_bounds__(256) k_noise_stats(const uint64_t* __restrict__ in, uint64_t* __restrict__ stats,
                                                     const PrimeConst* __restrict__ pcs, LimbMap map, int n_limbs,
                                                     int logN) {
    const int N = 1 << logN;
    const size_t row = blockIdx.x;
    noise_row(in + row * N, N, pcs[map.idx[row % n_limbs]].q, stats + row * 5);
}

// The same monitor for rows of any length n under one modulus q, with no
// transform context: the statistics need only q (the reference's call sites
// use toy moduli such as q = 113 on 32 coefficients).
__global__ void __launch_bounds__(256) k_noise_stats_q(const uint64_t* __restrict__ in, uint64_t* __restrict__ stats,
                                                       size_t n, uint64_t q) {
    const size_t row = blockIdx.x;
    noise_row(in + row * n, (int)n, q, stats + row * 5);
}

// The reference's toy rescale (scripts/verify_e2e_mult.py:70-75, the same
// formula in gen_tb_e2e_vec.py:59-63 and gen_tb_rescale_vec.py:28-33):
//   out = floor((c q_out + floor(q_in / 2)) / (q_in 2^sb)) mod q_out
// for any 64-bit c.  floor(X / (q_in 2^sb)) = floor(floor(X / 2^sb) / q_in),
// and the 128-bit Y = floor(X / 2^sb) is divided by q_in sixteen bits at a
// time: the running remainder stays below q_in < 2^45, so (r << 16) | digit
// fits 64 bits, and each quotient digit (< 2^16) is folded into the result
// modulo q_out (Q 2^16 + d < 2^62).  Exact for 2 <= q_in, q_out < 2^45.
__global__ void k_mod_switch_toy(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, size_t n,
                                 uint64_t q_in, uint64_t q_out, int sb) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    const uint64_t c = in[g];
    uint64_t lo = c * q_out, hi = __umul64hi(c, q_out);
    const uint64_t lo2 = lo + (q_in >> 1);
    hi += lo2 < lo ? 1 : 0;
    lo = lo2;
    if (sb >= 64) {
        lo = sb >= 128 ? 0 : hi >> (sb - 64);
        hi = 0;
    } else if (sb > 0) {
        lo = (lo >> sb) | (hi << (64 - sb));
        hi >>= sb;
    }
    uint64_t r = 0, Q = 0;
    for (int k = 7; k >= 0; --k) {
        const uint64_t digit = ((k >= 4 ? hi : lo) >> (16 * (k & 3))) & 0xffffu;
        const uint64_t cur = (r << 16) | digit;
        const uint64_t d = cur / q_in;
        r = cur - d * q_in;
        Q = ((Q << 16) + d) % q_out;
    }
    out[g] = Q;
}

// Exact product modulo any q through three transform primes p0, p1, p2:
// the lift takes x mod q (any 64-bit x) and writes its residues modulo the
// three primes, [rows][N] -> [rows][3][N].
__global__ void k_lift3(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                        const PrimeConst* __restrict__ pcs, PrimeConst Pq, int logN, size_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const size_t N = (size_t)1 << logN;
    const size_t n = g & (N - 1), row = g >> logN;
    const uint64_t x = reduce64(in[g], Pq);
    uint64_t* o = out + row * 3 * N + n;
    for (int i = 0; i < 3; ++i) o[(size_t)i * N] = reduce64(x, pcs[i]);
}

// Garner's recombination of the three residues of a product over Z whose
// coefficients lie in (-P/4, P/4), P = p0 p1 p2:
//   t1 = (r1 - r0) p0^-1 mod p1,  t2 = (r2 - r0 - p0 t1) (p0 p1)^-1 mod p2,
//   x = r0 + p0 t1 + p0 p1 t2 in [0, P),  the value is x - P when t2 > p2 / 2,
// and the output is that value modulo q, [rows][3][N] -> [rows][N].
__global__ void k_crt3_modq(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                            const PrimeConst* __restrict__ pcs, Crt3Const cc, int logN, size_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const size_t N = (size_t)1 << logN;
    const size_t n = g & (N - 1), row = g >> logN;
    const uint64_t* src = in + row * 3 * N + n;
    const uint64_t p1 = pcs[1].q, p2 = pcs[2].q;
    const uint64_t r0 = src[0], r1 = src[N], r2 = src[2 * N];
    const uint64_t t1 = csub(mul_shoup(sub_mod(r1, reduce64(r0, pcs[1]), p1), cc.p0inv_p1, cc.p0inv_p1_s, p1), p1);
    const uint64_t u = csub(mul_shoup(t1, cc.p0_p2, cc.p0_p2_s, p2), p2);            // p0 t1 mod p2
    const uint64_t v = sub_mod(sub_mod(r2, reduce64(r0, pcs[2]), p2), u, p2);
    const uint64_t t2 = csub(mul_shoup(v, cc.p01inv_p2, cc.p01inv_p2_s, p2), p2);
    const PrimeConst& Q = cc.Pq;
    uint64_t acc = reduce64(r0, Q);
    acc = add_mod(acc, mul_barrett(reduce64(t1, Q), cc.p0_q, Q), Q.q);
    acc = add_mod(acc, mul_barrett(reduce64(t2, Q), cc.p01_q, Q), Q.q);
    if (t2 > (p2 >> 1)) acc = sub_mod(acc, cc.P_q, Q.q);
    out[g] = acc;
}

// Single-prime modulus switch round-half-up(c q_out / q_in) mod q_out
// (scripts/verify_rescale.py:38-53).  c q_out < 2^(2l), l = bit length of
// q_in, so Barrett (HAC 14.42) with mu = floor(2^(2l) / q_in) gives the
// quotient to within 2; the remainder decides the rounding.
__global__ void k_mod_switch(const uint64_t* __restrict__ in, uint64_t* __restrict__ out, size_t n, uint64_t q_in,
                             uint64_t q_out, uint64_t mu, int ell) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n) return;
    uint64_t c = in[g];
    // c = c1 q_in + c0 gives the same result mod q_out as c0 alone
    if (c >= q_in) c %= q_in;
    const uint64_t lo = c * q_out, hi = __umul64hi(c, q_out);
    const uint64_t q1 = (hi << (65 - ell)) | (lo >> (ell - 1));
    uint64_t quo = __umul64hi(q1 << (63 - ell), mu);
    uint64_t r = lo - quo * q_in;   // in [0, 3 q_in)
    for (int it = 0; it < 2 && r >= q_in; ++it) {
        r -= q_in;
        ++quo;
    }
    if (2 * r >= q_in) ++quo;
    out[g] = quo >= q_out ? quo - q_out : quo;
}

// Plaintext linear combination in the evaluation domain, the inner loop of
// the diagonal (Halevi-Shoup) matrix-vector product:
//   out[g][c][poly][i][n] = sum_b E[b][c][poly][i][n] * D[g][b][i][n]  (mod q_i)
// E: [nb][B][2][l][N] evaluation-form ciphertext rows, D: [ng][nb][l][N]
// (value, RN(value / q)) FP64 plaintext pairs shared by the whole batch.
// With scalar = 1, D is [ng][nb][l]: one integer per (g, b, limb), the
// same product in either domain (a constant plaintext), so E may then be in
// coefficient form.
// Blocked for reuse: one thread per (block of R rows, limb i, coefficient n)
// keeps NG x R accumulators, so each input value is
// loaded once per call and each plaintext value once per R rows (rows are
// the (ciphertext, poly) pairs, 2 B of them).  ng <= NG; rows past the end
// of the last block are not stored.
// Workgroup order: a workgroup covers a chunk of 64 coefficients (N when
// smaller) of one limb, each of its waves one row block, so the waves read
// the same plaintext values (the second and later from the CU's L1).  The
// row-block groups of one (limb, chunk) are adjacent logical indices, and
// logical indices are dealt to the XCDs in contiguous ranges (workgroup w
// runs on XCD w mod 8), so a chunk's plaintext rows are read from HBM once
// and then from that XCD's L2.  Each baby step's input rows are loaded one
// step ahead, and no load sits behind a branch: the kernel is bound by the
// latency of its loads, which a wave would otherwise wait out one by one.
template <int NG, int R>
__global__ void __launch_bounds__(256) k_lincomb_blk(const uint64_t* __restrict__ E,
                                                     const double2* __restrict__ D, uint64_t* __restrict__ out,
                                                     const PrimeF* __restrict__ pfs, LimbMap map, int nb, int ng,
                                                     int rows, int l, int logN, int scalar) {
    const int G = (int)gridDim.x, w = (int)blockIdx.x;
    const int L = (G & 7) == 0 ? (w & 7) * (G >> 3) + (w >> 3) : w;
    const int lanes = blockDim.x < 64 ? (int)blockDim.x : 64;   // coefficients of a chunk
    const int W = (int)blockDim.x / lanes;                      // row blocks (waves) per workgroup
    const int lane = (int)threadIdx.x % lanes, wv = (int)threadIdx.x / lanes;
    const int nrb = (rows + R - 1) / R;
    const int nrbg = (nrb + W - 1) / W;
    const int chunks = (1 << logN) / lanes;
    const int rbg = L % nrbg, rest = L / nrbg;
    const size_t N = (size_t)1 << logN, LN = (size_t)l * N;
    const size_t n = (size_t)(rest % chunks) * lanes + lane;
    const int i = wave_uniform(rest / chunks);
    const int r0 = (rbg * W + wv) * R;
    if (r0 >= rows) return;   // a wave past the last row block (no barrier follows)
    const PrimeF& pf = pfs[map.idx[i]];
    const double nq = pf.nq;
    const size_t dn = scalar ? 1 : N;
    const size_t dstride = (size_t)l * dn;
    const double2* dbase = D + (size_t)i * dn + (scalar ? 0 : n);
    const uint64_t* ebase = E + (size_t)r0 * LN + (size_t)i * N + n;
    const size_t brow = (size_t)rows * LN;   // one b (or g) block of rows
    double acc[NG][R];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int r = 0; r < R; ++r) acc[g][r] = 0.0;
    // Every load of a baby step is issued before the first product: rows
    // past the end and giant steps past ng read clamped (valid) addresses
    // whose products are never stored, so no load sits behind a branch that
    // would make it wait alone.
    const int rlast = rows - 1 - r0;
    uint64_t en[R];   // the next baby step's rows, loaded one step ahead
#pragma unroll
    for (int r = 0; r < R; ++r) en[r] = ebase[(size_t)(r < rlast ? r : rlast) * LN];
    for (int b = 0; b < nb; ++b) {
        double e[R];
        double2 w[NG];
#pragma unroll
        for (int r = 0; r < R; ++r) e[r] = u2d(en[r]);
#pragma unroll
        for (int g = 0; g < NG; ++g) w[g] = dbase[((size_t)(g < ng ? g : ng - 1) * nb + b) * dstride];
        const int bn = b + 1 < nb ? b + 1 : b;
#pragma unroll
        for (int r = 0; r < R; ++r) en[r] = ebase[(size_t)bn * brow + (size_t)(r < rlast ? r : rlast) * LN];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
#pragma unroll
            for (int r = 0; r < R; ++r) {
#pragma clang fp contract(off)
                acc[g][r] = acc[g][r] + mulw(e[r], w[g].x, w[g].y, nq);   // nb <= 32: below 20 q
            }
        }
    }
    uint64_t* obase = out + (size_t)r0 * LN + (size_t)i * N + n;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        if (g < ng) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (r0 + r < rows)
                    obase[(size_t)g * brow + (size_t)r * LN] = d2u(canon(rmod(acc[g][r], pf.qinv, nq), pf.qinv, nq));
        }
    }
}

// Re-express a decrypted message in L limbs (the DARE refresh of
// exp_e24_rotation_bootstrap.py:287-304 keeps the scale-carrying integer
// message as is).  While |m| < q_0 / 2 the centred residue in limb 0 is the
// integer m itself; it is reduced into every target prime.
__global__ void k_lift_small(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                             const PrimeConst* __restrict__ pcs, LimbMap map, int l_in, int l_out, int logN,
                             size_t total) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    const size_t N = (size_t)1 << logN;
    const size_t n = g & (N - 1), row = g >> logN;
    const uint64_t q0 = pcs[map.idx[0]].q;
    const uint64_t c = in[row * l_in * N + n];
    const bool neg = c > (q0 >> 1);
    const uint64_t a = neg ? q0 - c : c;
    uint64_t* o = out + row * l_out * N + n;
    for (int i = 0; i < l_out; ++i) {
        const PrimeConst& pc = pcs[map.idx[i]];
        const uint64_t r = reduce64(a, pc);
        o[(size_t)i * N] = (neg && r) ? pc.q - r : r;
    }
}

template <typename K>
hipError_t set_lds(K kernel, int bytes) {
    return hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

template <int LOGN, int T>
hipError_t run_tensor(const DevTables& t, const uint64_t* ea, const uint64_t* eb, uint64_t* d01, uint64_t* d2e,
                      uint64_t* d2c, int batch, int l, const LimbMap& map, const ModDownConst& mc, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    hipError_t e = set_lds(k_tensor_intt<LOGN, T>, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_tensor_intt<LOGN, T>), dim3(batch * l), dim3(T), C::LDS_BYTES, s, ea, eb, d01, d2e, d2c,
                       t.twf, t.pf, map, l, mc);
    return hipGetLastError();
}

template <int LOGN, int T>
hipError_t run_front(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* d01, uint64_t* d2e,
                     uint64_t* d2c, int batch, int l, const LimbMap& map, const ModDownConst& mc, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    constexpr int lds = ntt::pair_pays<C>() ? 2 * C::LDS_BYTES : C::LDS_BYTES;
    // the squaring kernel where both operands are one ciphertext batch
    const bool sq = ntt::pair_pays<C>() && a == b;
    auto k = sq ? k_front<LOGN, T, true> : k_front<LOGN, T>;
    hipError_t e = set_lds(k, lds);
    if (e != hipSuccess) return e;
    // paired shapes loop over items: one workgroup per resident slot
    int grid = batch * l;
    if (ntt::pair_pays<C>()) {
        static int per_cu[2];   // per variant, workgroups a CU holds at once: a property of the compiled kernel
        if (!per_cu[sq]) {
            int n = 0;
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, T, lds);
            if (e != hipSuccess) return e;
            per_cu[sq] = n > 0 ? n : 1;
        }
        grid = grid < per_cu[sq] * t.n_cu ? grid : per_cu[sq] * t.n_cu;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(T), lds, s, a, b, d01, d2e, d2c, t.twf, t.pf, map, l, batch, mc);
    return hipGetLastError();
}

template <int LOGN, int T>
hipError_t run_ks(const DevTables& t, const uint64_t* d2c, size_t d2s, const uint64_t* d2e, const uint64_t* d01,
                  const double* key, int key_limbs, uint64_t* acc, int batch, int l, const LimbMap& map,
                  int pidx_p, const uint64_t* key_table, const KsTargets& tg, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    constexpr int lds = ntt::pair_pays<C>() ? 2 * C::LDS_BYTES : C::LDS_BYTES;
    auto k = tg.out ? k_keyswitch<LOGN, T, false, true>
                    : (key_table ? k_keyswitch<LOGN, T, true> : k_keyswitch<LOGN, T, false>);
    hipError_t e = set_lds(k, lds);
    if (e != hipSuccess) return e;
    const int nt = tg.count ? tg.count : l + 1;
    ModDownConst mc;
    if (tg.mc) mc = *tg.mc;
    else memset(&mc, 0, sizeof(mc));
    // paired shapes loop over items: one workgroup per resident slot
    const int items = batch * nt;
    int grid = items;
    if (ntt::pair_pays<C>()) {
        static int per_cu[3];   // per kernel variant: workgroups a CU holds at once
        const int kv = tg.out ? 2 : (key_table ? 1 : 0);
        if (!per_cu[kv]) {
            int n = 0;
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k, T, lds);
            if (e != hipSuccess) return e;
            per_cu[kv] = n > 0 ? n : 1;
        }
        grid = grid < per_cu[kv] * t.n_cu ? grid : per_cu[kv] * t.n_cu;
    }
    hipLaunchKernelGGL(k, dim3(grid), dim3(T), lds, s, d2c, d2s, d2e, d01, key, key_limbs, acc, t.twf,
                       t.pf, map, l, pidx_p, batch, t.narrow ? 0 : 1, key_table, tg.first, tg.out, mc,
                       tg.pair_major, items);
    return hipGetLastError();
}

// Untuned launch shapes: the tuner's usual choice on MI355X at N = 8192
// (one coefficient group of 8 per thread, paired transforms).
int default_threads(int logN) { return logN >= 14 ? 512 : (logN == 13 ? 1024 : 256); }

}  // namespace

hipError_t launch_tensor_intt(const DevTables& t, const uint64_t* ea, const uint64_t* eb, uint64_t* d01,
                              uint64_t* d2e, uint64_t* d2c, int batch, int l, const LimbMap& map,
                              const ModDownConst& mc, int threads, hipStream_t s) {
    if (batch * l == 0) return hipSuccess;
    if (!threads) threads = default_threads(t.logN);
    if (!t.fp64) return launch_tensor_intt_int(t, ea, eb, d01, d2e, d2c, batch, l, map, mc, threads, s);
    switch (t.logN * 10000 + threads) {
        case 120256: return run_tensor<12, 256>(t, ea, eb, d01, d2e, d2c, batch, l, map, mc, s);
        case 120512: return run_tensor<12, 512>(t, ea, eb, d01, d2e, d2c, batch, l, map, mc, s);
        case 130256: return run_tensor<13, 256>(t, ea, eb, d01, d2e, d2c, batch, l, map, mc, s);
        case 130512: return run_tensor<13, 512>(t, ea, eb, d01, d2e, d2c, batch, l, map, mc, s);
        case 131024: return run_tensor<13, 1024>(t, ea, eb, d01, d2e, d2c, batch, l, map, mc, s);
        case 140512: return run_tensor<14, 512>(t, ea, eb, d01, d2e, d2c, batch, l, map, mc, s);
        case 141024: return run_tensor<14, 1024>(t, ea, eb, d01, d2e, d2c, batch, l, map, mc, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_front(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* d01, uint64_t* d2e,
                        uint64_t* d2c, int batch, int l, const LimbMap& map, const ModDownConst& mc, int threads,
                        hipStream_t s) {
    if (batch * l == 0) return hipSuccess;
    if (!threads) threads = default_threads(t.logN);
    if (!t.fp64) return launch_front_int(t, a, b, d01, d2e, d2c, batch, l, map, mc, threads, s);
    switch (t.logN * 10000 + threads) {
        case 120256: return run_front<12, 256>(t, a, b, d01, d2e, d2c, batch, l, map, mc, s);
        case 120512: return run_front<12, 512>(t, a, b, d01, d2e, d2c, batch, l, map, mc, s);
        case 130256: return run_front<13, 256>(t, a, b, d01, d2e, d2c, batch, l, map, mc, s);
        case 130512: return run_front<13, 512>(t, a, b, d01, d2e, d2c, batch, l, map, mc, s);
        case 131024: return run_front<13, 1024>(t, a, b, d01, d2e, d2c, batch, l, map, mc, s);
        case 140512: return run_front<14, 512>(t, a, b, d01, d2e, d2c, batch, l, map, mc, s);
        case 141024: return run_front<14, 1024>(t, a, b, d01, d2e, d2c, batch, l, map, mc, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_keyswitch(const DevTables& t, const uint64_t* d2c, size_t d2c_bstride, const uint64_t* d2e,
                            const uint64_t* d01,
                            const void* key_pairs, int key_limbs, uint64_t* acc, int batch, int l,
                            const LimbMap& map, int pidx_p, int threads, hipStream_t s, const uint64_t* key_table,
                            const KsTargets& tg) {
    if (batch * l == 0) return hipSuccess;
    if (tg.out && (key_table || tg.first < 0 || tg.first + tg.count > l - 1 || !tg.mc)) return hipErrorInvalidValue;
    if (tg.count < 0 || tg.first < 0 || tg.first + (tg.count ? tg.count : l + 1) > l + 1) return hipErrorInvalidValue;
    if (!threads) threads = default_threads(t.logN);
    if (!t.fp64)
        return launch_keyswitch_int(t, d2c, d2c_bstride, d2e, d01, (const ulonglong2*)key_pairs, key_limbs, acc,
                                    batch, l, map, pidx_p, threads, s, key_table, tg);
    const double* key = (const double*)key_pairs;
    switch (t.logN * 10000 + threads) {
        case 120256: return run_ks<12, 256>(t, d2c, d2c_bstride, d2e, d01, key, key_limbs, acc, batch, l, map, pidx_p, key_table, tg, s);
        case 120512: return run_ks<12, 512>(t, d2c, d2c_bstride, d2e, d01, key, key_limbs, acc, batch, l, map, pidx_p, key_table, tg, s);
        case 130256: return run_ks<13, 256>(t, d2c, d2c_bstride, d2e, d01, key, key_limbs, acc, batch, l, map, pidx_p, key_table, tg, s);
        case 130512: return run_ks<13, 512>(t, d2c, d2c_bstride, d2e, d01, key, key_limbs, acc, batch, l, map, pidx_p, key_table, tg, s);
        case 131024: return run_ks<13, 1024>(t, d2c, d2c_bstride, d2e, d01, key, key_limbs, acc, batch, l, map, pidx_p, key_table, tg, s);
        case 140512: return run_ks<14, 512>(t, d2c, d2c_bstride, d2e, d01, key, key_limbs, acc, batch, l, map, pidx_p, key_table, tg, s);
        case 141024: return run_ks<14, 1024>(t, d2c, d2c_bstride, d2e, d01, key, key_limbs, acc, batch, l, map, pidx_p, key_table, tg, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_moddown_rescale(const DevTables& t, const uint64_t* acc, const uint64_t* dcoef, int dmask,
                                  uint64_t* out,
                                  int batch, int l, const LimbMap& map, int pidx_special, const ModDownConst& mc,
                                  bool rescale, hipStream_t s, const uint64_t* dadd) {
    const size_t total = ((size_t)batch * 2) << t.logN;
    if (!total) return hipSuccess;
    if (l >= 1 && l <= 8 && t.logN >= 1) {
        const size_t half = total >> 1;
        const dim3 grid((unsigned)((half + 255) / 256));
#define CKKS_MD_CASE(LL)                                                                                          \
    case LL:                                                                                                      \
        if (rescale)                                                                                              \
            hipLaunchKernelGGL((k_moddown_rescale_l<LL, true>), grid, dim3(256), 0, s, acc, dcoef, dmask, dadd,    \
                               out, t.pc, map, pidx_special, mc, t.narrow, t.logN, half);                         \
        else                                                                                                      \
            hipLaunchKernelGGL((k_moddown_rescale_l<LL, false>), grid, dim3(256), 0, s, acc, dcoef, dmask, dadd,   \
                               out, t.pc, map, pidx_special, mc, t.narrow, t.logN, half);                         \
        return hipGetLastError();
        switch (l) {
            CKKS_MD_CASE(1)
            CKKS_MD_CASE(2)
            CKKS_MD_CASE(3)
            CKKS_MD_CASE(4)
            CKKS_MD_CASE(5)
            CKKS_MD_CASE(6)
            CKKS_MD_CASE(7)
            CKKS_MD_CASE(8)
            default: break;
        }
#undef CKKS_MD_CASE
    }
    hipLaunchKernelGGL(k_moddown_rescale, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, acc, dcoef, dmask,
                       dadd, out, t.pc, map, l, pidx_special, mc, rescale ? 1 : 0, t.narrow, t.logN, total);
    return hipGetLastError();
}

hipError_t launch_rescale(const DevTables& t, const uint64_t* in, uint64_t* out, int n_polys, int l,
                          const LimbMap& map, const ModDownConst& mc, hipStream_t s) {
    const size_t total = (size_t)n_polys << t.logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_rescale, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, out, t.pc, map, l, mc,
                       t.narrow, t.logN, total);
    return hipGetLastError();
}

hipError_t launch_automorphism_groups(const DevTables& t, const uint64_t* in, uint64_t* out, int n_groups,
                                     int n_limbs, size_t group_stride, const LimbMap& map, uint64_t galois,
                                     hipStream_t s, const uint64_t* gal_table, int gal_div) {
    if (n_groups == 0 || n_limbs == 0) return hipSuccess;
    const int N = 1 << t.logN;
    const uint64_t M = 2ull << t.logN;
    // inverse of the odd Galois element modulo 2N (a power of two)
    uint64_t inv = 1;
    for (int k = 0; k < 7; ++k) inv = inv * (2 - galois * inv);
    inv &= M - 1;
    const int lds = N * 8;
    if (N >= 128 && lds <= 128 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)k_automorphism_row,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_automorphism_row, dim3((unsigned)(n_groups * n_limbs)), dim3(N >= 2048 ? 1024 : N / 2),
                           lds, s, in, out, t.pc, map, n_limbs, group_stride, inv, t.logN, gal_table,
                           gal_div > 0 ? gal_div : 1);
        return hipGetLastError();
    }
    if (gal_table) return hipErrorInvalidValue;   // per-group elements need the LDS kernel
    // rows too long for LDS: the gather kernel, one launch per group
    const size_t total = (size_t)n_limbs << t.logN;
    for (int g = 0; g < n_groups; ++g) {
        hipLaunchKernelGGL(k_automorphism, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                           in + g * group_stride, out + g * group_stride, t.pc, map, n_limbs, inv, t.logN, total);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_automorphism(const DevTables& t, const uint64_t* in, uint64_t* out, int n_polys, int n_limbs,
                               const LimbMap& map, uint64_t galois, hipStream_t s) {
    return launch_automorphism_groups(t, in, out, n_polys, n_limbs, (size_t)n_limbs << t.logN, map, galois, s);
}

hipError_t launch_noise_stats(const DevTables& t, const uint64_t* in, uint64_t* stats, int n_polys, int n_limbs,
                              const LimbMap& map, hipStream_t s) {
    const int rows = n_polys * n_limbs;
    if (!rows) return hipSuccess;
    hipLaunchKernelGGL(k_noise_stats, dim3(rows), dim3(256), 0, s, in, stats, t.pc, map, n_limbs, t.logN);
    return hipGetLastError();
}

hipError_t launch_noise_stats_q(const uint64_t* in, uint64_t* stats, int rows, size_t n, uint64_t q, hipStream_t s) {
    if (!rows || !n) return hipSuccess;
    hipLaunchKernelGGL(k_noise_stats_q, dim3(rows), dim3(256), 0, s, in, stats, n, q);
    return hipGetLastError();
}

hipError_t launch_mod_switch_toy(const uint64_t* in, uint64_t* out, size_t n, uint64_t q_in, uint64_t q_out, int sb,
                                 hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_mod_switch_toy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n, q_in, q_out,
                       sb);
    return hipGetLastError();
}

hipError_t launch_lift3(const DevTables& t, const uint64_t* in, uint64_t* out, int rows, const PrimeConst& Pq,
                        hipStream_t s) {
    const size_t total = (size_t)rows << t.logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_lift3, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, out, t.pc, Pq, t.logN,
                       total);
    return hipGetLastError();
}

hipError_t launch_crt3_modq(const DevTables& t, const uint64_t* in, uint64_t* out, int rows, const Crt3Const& cc,
                            hipStream_t s) {
    const size_t total = (size_t)rows << t.logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_crt3_modq, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, out, t.pc, cc,
                       t.logN, total);
    return hipGetLastError();
}

hipError_t launch_mod_switch(const uint64_t* in, uint64_t* out, size_t n, uint64_t q_in, uint64_t q_out, uint64_t mu,
                             int ell, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_mod_switch, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n, q_in, q_out, mu,
                       ell);
    return hipGetLastError();
}

namespace {

template <int NG, int R>
hipError_t run_lincomb_blk(const DevTables& t, const uint64_t* E, const double2* D, uint64_t* out, int nb, int ng,
                           int rows, int l, const LimbMap& map, int scalar, hipStream_t s) {
    const int width = t.logN < 8 ? 1 << t.logN : 256;   // threads per workgroup
    const int lanes = width < 64 ? width : 64, W = width / lanes;
    const int nrbg = ((rows + R - 1) / R + W - 1) / W;
    const size_t blocks = (size_t)nrbg * l * (((size_t)1 << t.logN) / lanes);
    hipLaunchKernelGGL((k_lincomb_blk<NG, R>), dim3((unsigned)blocks), dim3(width), 0, s, E, D, out, t.pf, map, nb,
                       ng, rows, l, t.logN, scalar);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_lincomb(const DevTables& t, const uint64_t* E, const void* D_pairs, uint64_t* out, int nb, int ng,
                          int batch, int l, const LimbMap& map, int scalar, hipStream_t s) {
    const int rows = batch * 2;
    if (!rows || !ng || !l) return hipSuccess;
    if (!t.fp64) return launch_lincomb_int(t, E, (const ulonglong2*)D_pairs, out, nb, ng, batch, l, map, scalar, s);
    const double2* D = (const double2*)D_pairs;
    const size_t dn = scalar ? 1 : ((size_t)1 << t.logN);
    const size_t LN = (size_t)l << t.logN;
    // giant steps in chunks of at most 8 (ng = 10 as 8 + 2: an input read
    // twice costs less than 16 accumulators a thread)
    for (int g0 = 0; g0 < ng; g0 += 8) {
        const int gc = ng - g0 < 8 ? ng - g0 : 8;
        const double2* Dg = D + (size_t)g0 * nb * l * dn;
        uint64_t* og = out + (size_t)g0 * rows * LN;
        hipError_t e;
        if (gc <= 4) e = run_lincomb_blk<4, 4>(t, E, Dg, og, nb, gc, rows, l, map, scalar, s);
        else e = run_lincomb_blk<8, 4>(t, E, Dg, og, nb, gc, rows, l, map, scalar, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_lift_small(const DevTables& t, const uint64_t* in, uint64_t* out, int rows, int l_in, int l_out,
                             const LimbMap& map, hipStream_t s) {
    const size_t total = (size_t)rows << t.logN;
    if (!total) return hipSuccess;
    hipLaunchKernelGGL(k_lift_small, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, in, out, t.pc, map, l_in,
                       l_out, t.logN, total);
    return hipGetLastError();
}
