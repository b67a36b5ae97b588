// The HMult front, key switch and plaintext combination on 64-bit integer
// words: Shoup products for fixed operands (twiddles, key rows, constants)
// and Barrett products for data, lazy butterflies (modarith.h, ntt_core.h).
// The same words as the FP64 kernels of k_ckks.hip; a context selects them
// with ckks_ctx_set_arith(ctx, CKKS_ARITH_INT64), its prepared keys and
// plaintexts then being (value, floor(value 2^63 / q)) pairs.
//
// Layouts as in k_ckks.hip:
//   coefficient / evaluation tensors  [batch][poly][limb][N]
//   key  [digit][2][key_limbs][N] of (value, Shoup companion) pairs,
//        evaluation order; digit j = (b_j, a_j), last limb = special prime
#include <string.h>

#include "launch.h"
#include "ntt_core.h"

namespace {

using ntt::Cfg;

// Tensor product in the evaluation domain (verify_e2e_mult.py:78-92 /
// rtl/m4/e2e_ckks_mult_ctrl_m4.sv S_MUL_*):
//   d0 = a0 b0, d1 = a0 b1 + a1 b0, d2 = a1 b1.
// Writes p*d0, p*d1 (evaluation form, ready to be folded into the key
// switch accumulators), d2 in evaluation form, and d2 back in coefficient
// form (the key-switch digits).
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_tensor_intt_i(const uint64_t* __restrict__ ea, const uint64_t* __restrict__ eb,
                                                   uint64_t* __restrict__ d01, uint64_t* __restrict__ d2e,
                                                   uint64_t* __restrict__ d2c, const ulonglong2* __restrict__ tw_all,
                                                   const PrimeConst* __restrict__ pcs, LimbMap map, int l,
                                                   ModDownConst mc) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    const int tid = threadIdx.x;
    const int b = wave_uniform(blockIdx.x / l), i = wave_uniform(blockIdx.x % l);
    const int p = map.idx[i];
    const PrimeConst pc = pcs[p];
    const size_t LN = (size_t)l * C::N;
    const size_t ct0 = (size_t)b * 2 * LN + (size_t)i * C::N;  // poly 0, limb i
    const uint64_t pm = mc.p_mod[i], pms = mc.p_mod_s[i];
    uint64_t x[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const size_t j = (size_t)e * C::T + tid;
        const uint64_t a0 = ea[ct0 + j], a1 = ea[ct0 + LN + j];
        const uint64_t b0 = eb[ct0 + j], b1 = eb[ct0 + LN + j];
        const uint64_t d0 = mul_barrett(a0, b0, pc);
        const uint64_t d1 = add_mod(mul_barrett(a0, b1, pc), mul_barrett(a1, b0, pc), pc.q);
        const uint64_t d2 = mul_barrett(a1, b1, pc);
        d01[ct0 + j] = csub(mul_shoup(d0, pm, pms, pc.q), pc.q);
        d01[ct0 + LN + j] = csub(mul_shoup(d1, pm, pms, pc.q), pc.q);
        d2e[(size_t)b * LN + (size_t)i * C::N + j] = d2;
        x[e] = d2;
    }
    ntt::coalesced_to_last<C>(x, lds, tid);
    ntt::inv_ntt<C>(x, lds, tw_all + ((size_t)p * 2 + 1) * C::N, pc, tid);
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = csub(x[e], pc.q);
    ntt::store_coalesced<C>(x, d2c + (size_t)b * LN + (size_t)i * C::N, tid);
}

// The whole front of HMult for one (pair, limb i), coefficient form in:
// forward transforms of a0, a1, b0, b1, the tensor product, p*d0 and p*d1
// and d2 written in evaluation form, d2 transformed back to coefficient form.
// Limb-major grid: workgroups in flight share one prime's twiddles.
template <int LOGN, int T>
__global__ void __launch_bounds__(T) k_front_i(const uint64_t* __restrict__ A, const uint64_t* __restrict__ B,
                                             uint64_t* __restrict__ d01, uint64_t* __restrict__ d2e,
                                             uint64_t* __restrict__ d2c, const ulonglong2* __restrict__ tw_all,
                                             const PrimeConst* __restrict__ pcs, LimbMap map, int l, int batch,
                                             ModDownConst mc) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    const int tid = threadIdx.x;
    const int i = wave_uniform(blockIdx.x / batch), b = wave_uniform(blockIdx.x % batch);
    const int p = map.idx[i];
    const PrimeConst pc = pcs[p];
    const ulonglong2* tw = tw_all + (size_t)p * 2 * C::N;
    const size_t LN = (size_t)l * C::N;
    const size_t c0 = (size_t)b * 2 * LN + (size_t)i * C::N;   // poly 0 of pair b, limb i
    const uint64_t pm = mc.p_mod[i], pms = mc.p_mod_s[i];
    if constexpr (ntt::pair_pays<C>()) {
        // a0 with a1 and b0 with b1 as transform pairs
        uint64_t x0[C::E], x1[C::E], y0[C::E], y1[C::E];
        ntt::load_coalesced<C>(x0, A + c0, tid);
        ntt::load_coalesced<C>(x1, A + c0 + LN, tid);
        ntt::fwd_ntt2<C>(x0, x1, lds, tw, pc.q, pc.two_q, tid);
        ntt::load_coalesced<C>(y0, B + c0, tid);
        ntt::load_coalesced<C>(y1, B + c0 + LN, tid);
        ntt::fwd_ntt2<C>(y0, y1, lds, tw, pc.q, pc.two_q, tid);
        uint64_t* o0 = d01 + c0;
        uint64_t* o2 = d2e + (size_t)b * LN + (size_t)i * C::N;
        // the inverse transform's first twiddles, loaded before the stores
        const auto pre = ntt::inv_first<C, C::NP - 1>(tw + C::N, tid);
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            const uint64_t a0 = reduce64(x0[k], pc), a1 = reduce64(x1[k], pc);
            const uint64_t b0 = reduce64(y0[k], pc), b1 = reduce64(y1[k], pc);
            const int idx = ntt::eval_index<C>(k, tid);
            o0[idx] = csub(mul_shoup(mul_barrett(a0, b0, pc), pm, pms, pc.q), pc.q);
            const uint64_t d1 = add_mod(mul_barrett(a1, b0, pc), mul_barrett(a0, b1, pc), pc.q);
            o0[LN + idx] = csub(mul_shoup(d1, pm, pms, pc.q), pc.q);
            const uint64_t d2 = mul_barrett(a1, b1, pc);
            o2[idx] = d2;
            x0[k] = d2;
        }
        ntt::inv_ntt_db<C>(x0, lds, tw + C::N, pc, tid, pre);
#pragma unroll
        for (int k = 0; k < C::E; ++k) x0[k] = csub(x0[k], pc.q);
        ntt::store_coalesced<C>(x0, d2c + (size_t)b * LN + (size_t)i * C::N, tid);
        return;
    }
    uint64_t x0[C::E], x1[C::E], y[C::E], z[C::E];
    ntt::load_coalesced<C>(x0, A + c0, tid);
    ntt::fwd_ntt<C>(x0, lds, tw, pc.q, pc.two_q, tid);
    ntt::load_coalesced<C>(x1, A + c0 + LN, tid);
    ntt::fwd_ntt<C>(x1, lds, tw, pc.q, pc.two_q, tid);
    ntt::load_coalesced<C>(y, B + c0, tid);
    ntt::fwd_ntt<C>(y, lds, tw, pc.q, pc.two_q, tid);
    uint64_t* o0 = d01 + c0;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const uint64_t a0 = reduce64(x0[k], pc), a1 = reduce64(x1[k], pc), b0 = reduce64(y[k], pc);
        o0[ntt::eval_index<C>(k, tid)] = csub(mul_shoup(mul_barrett(a0, b0, pc), pm, pms, pc.q), pc.q);
        z[k] = mul_barrett(a1, b0, pc);
        x0[k] = a0;
        x1[k] = a1;
    }
    ntt::load_coalesced<C>(y, B + c0 + LN, tid);
    ntt::fwd_ntt<C>(y, lds, tw, pc.q, pc.two_q, tid);
    uint64_t* o2 = d2e + (size_t)b * LN + (size_t)i * C::N;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const uint64_t b1 = reduce64(y[k], pc);
        const int idx = ntt::eval_index<C>(k, tid);
        const uint64_t d1 = add_mod(z[k], mul_barrett(x0[k], b1, pc), pc.q);
        o0[LN + idx] = csub(mul_shoup(d1, pm, pms, pc.q), pc.q);
        const uint64_t d2 = mul_barrett(x1[k], b1, pc);
        o2[idx] = d2;
        x0[k] = d2;
    }
    ntt::inv_ntt<C>(x0, lds, tw + C::N, pc, tid);
#pragma unroll
    for (int k = 0; k < C::E; ++k) x0[k] = csub(x0[k], pc.q);
    ntt::store_coalesced<C>(x0, d2c + (size_t)b * LN + (size_t)i * C::N, tid);
}

// a0 += v * b_j,t and a1 += v * a_j,t for digit j in evaluation order (key
// rows of (value, companion) pairs).  Each product is below 4q: with at most
// 15 digits and a start below q the sums stay below 2^51.
template <class C>
__device__ __forceinline__ void key_mac_i(uint64_t (&a0)[C::E], uint64_t (&a1)[C::E], const uint64_t (&v)[C::E],
                                        const ulonglong2* __restrict__ key, int j, int key_limbs, int kt,
                                        uint64_t nq, int tid) {
    const ulonglong2* kb = key + ((size_t)(j * 2 + 0) * key_limbs + kt) * C::N;
    const ulonglong2* ka = key + ((size_t)(j * 2 + 1) * key_limbs + kt) * C::N;
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        const int idx = ntt::eval_index<C>(k, tid);
        const ulonglong2 wb = kb[idx], wa = ka[idx];
        a0[k] += mul_shoup4_nq(v[k], wb.x, wb.y, nq);
        a1[k] += mul_shoup4_nq(v[k], wa.x, wa.y, nq);
    }
}

// ModDown by p and rescale by q_(l-1) of one coefficient of chain limb t <
// l - 1, from the coefficient-form accumulators of limb t (xt), of limb
// l - 1 (xl) and of the special prime (xp), all reduced: the per-limb steps
// of k_moddown_rescale_l, word for word.
__device__ __forceinline__ uint64_t moddown_rescale_one_i(uint64_t xt, uint64_t xl, uint64_t xp, int t, int lm1,
                                                        const PrimeConst& pc, const PrimeConst& pl,
                                                        const PrimeConst& pp, const ModDownConst& mc, int narrow) {
    const uint64_t tp = csub(xp + (pp.q >> 1), pp.q);
    const uint64_t ql = pl.q, q = pc.q;
    const uint64_t tpl = narrow ? csub(tp, ql) : reduce64(tp, pl);
    const uint64_t yl = csub(mul_shoup(xl + mc.hp[lm1] + ql - tpl, mc.pinv[lm1], mc.pinv_s[lm1], ql), ql);
    const uint64_t tl = csub(yl + (ql >> 1), ql);
    const uint64_t tpt = narrow ? csub(tp, q) : reduce64(tp, pc);
    const uint64_t yt = csub(mul_shoup(xt + mc.hp[t] + q - tpt, mc.pinv[t], mc.pinv_s[t], q), q);
    const uint64_t tlt = narrow ? csub(tl, q) : reduce64(tl, pc);
    return csub(mul_shoup(yt + mc.hq[t] + q - tlt, mc.qinv[t], mc.qinv_s[t], q), q);
}

// The fused epilogue for poly c of target t: x holds the accumulator in
// coefficient form (coalesced layout, reduced); the special-prime and limb
// l - 1 accumulators come from an earlier launch through acc_c ([l+1][N]).
template <class C>
__device__ __forceinline__ void fused_out_i(const uint64_t (&x)[C::E], const uint64_t* __restrict__ acc_c,
                                          uint64_t* __restrict__ out_c, int t, int l,
                                          const PrimeConst* __restrict__ pcs, const LimbMap& map, int pidx_p,
                                          const ModDownConst& mc, int narrow, int tid) {
    const PrimeConst pc = pcs[map.idx[t]], pl = pcs[map.idx[l - 1]], pp = pcs[pidx_p];
    const uint64_t* sp = acc_c + (size_t)l * C::N;
    const uint64_t* sl = acc_c + (size_t)(l - 1) * C::N;
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int j = e * C::T + tid;
        out_c[j] = moddown_rescale_one_i(x[e], sl[j], sp[j], t, l - 1, pc, pl, pp, mc, narrow);
    }
}

// Hybrid key switch for one (ciphertext, target limb t) pair, t = 0..l where
// t = l is the special prime p.  Generalises the single-digit relinearisation
// f0 = e0 + e2 b, f1 = e1 - e2 a (verify_keyswitch.py:62-68,
// rtl/m3/keyswitch_pipe.sv:90-100) to one digit per RNS prime:
//   acc0_t = [p d0]_t + sum_j NTT_t([d2]_(q_j)) * b_j,t
//   acc1_t = [p d1]_t - sum_j NTT_t([d2]_(q_j)) * a_j,t
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
__global__ void __launch_bounds__(T) k_keyswitch_i(const uint64_t* __restrict__ d2c, size_t d2c_bstride,
                                                 const uint64_t* __restrict__ d2e,
                                                 const uint64_t* __restrict__ d01, const ulonglong2* __restrict__ key,
                                                 int key_limbs, uint64_t* __restrict__ acc,
                                                 const ulonglong2* __restrict__ tw_all,
                                                 const PrimeConst* __restrict__ pcs, LimbMap map, int l, int pidx_p,
                                                 int batch, int reduce_digits, const uint64_t* __restrict__ key_table,
                                                 int t0, uint64_t* __restrict__ out_ct, ModDownConst mc, int narrow,
                                                 int pair_major) {
    using C = Cfg<LOGN, T>;
    extern __shared__ uint64_t lds[];
    const int tid = threadIdx.x;
    // Target-major order: workgroups in flight together share one target
    // limb, so each XCD's L2 holds one limb's key slice (2 dnum N 16 B) and
    // one prime's twiddles rather than the whole key.
    // ciphertext-major placement: as in the FP64 kernel (k_ckks.hip)
    const int G = (int)gridDim.x, nt = G / batch;
    int t, b;
    if (pair_major) {
        const int i = (int)blockIdx.x;
        const int L = (G & 7) == 0 ? (i & 7) * (G >> 3) + (i >> 3) : i;
        t = wave_uniform(t0 + L % nt);
        b = wave_uniform(L / nt);
    } else {
        t = wave_uniform(t0 + (int)(blockIdx.x / batch));
        b = wave_uniform(blockIdx.x % batch);
    }
    if constexpr (KEYTAB) key = reinterpret_cast<const ulonglong2*>(key_table[b]);
    const int pt = t < l ? map.idx[t] : pidx_p;
    const int kt = t < l ? t : key_limbs - 1;
    const PrimeConst pc = pcs[pt];
    const uint64_t nq = (uint64_t)0 - pc.q;
    const ulonglong2* tw = tw_all + (size_t)pt * 2 * C::N;
    const size_t LN = (size_t)l * C::N;
    uint64_t a0[C::E], a1[C::E], x[C::E];
    if (t < l && d01) {
        const uint64_t* s0 = d01 + (size_t)b * 2 * LN + (size_t)t * C::N;
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            const int j = ntt::eval_index<C>(k, tid);
            a0[k] = s0[j];
            const uint64_t v = s0[LN + j];
            a1[k] = v ? pc.q - v : 0;   // acc1 holds -(p d1) + sum d a, negated at the end
        }
    } else {
#pragma unroll
        for (int k = 0; k < C::E; ++k) a0[k] = a1[k] = 0;
    }
    for (int j = 0; j < l; ++j) {
        if (j == t && d2e) {
            // the digit of the target's own prime is already in evaluation form
            const uint64_t* s = d2e + (size_t)b * LN + (size_t)j * C::N;
#pragma unroll
            for (int k = 0; k < C::E; ++k) x[k] = s[ntt::eval_index<C>(k, tid)];
        } else {
            ntt::load_coalesced<C>(x, d2c + (size_t)b * d2c_bstride + (size_t)j * C::N, tid);
            // the digit's residues mod q_j enter the transform mod q_t as they
            // are when q_j < 2 q_t (reduce_digits = 0); else reduce them first
            if (reduce_digits) {
#pragma unroll
                for (int k = 0; k < C::E; ++k) x[k] = reduce64(x[k], pc);
            }
            // left unreduced (below 2^51): the key product is exact for any such input
            if constexpr (ntt::pair_pays<C>()) ntt::fwd_ntt_db<C>(x, lds, tw, pc.q, pc.two_q, tid);
            else ntt::fwd_ntt<C>(x, lds, tw, pc.q, pc.two_q, tid);
        }
        key_mac_i<C>(a0, a1, x, key, j, key_limbs, kt, nq, tid);
    }
    // the double-buffered digit transforms leave reads of the LDS in flight
    if constexpr (ntt::pair_pays<C>()) __syncthreads();
#pragma unroll
    for (int k = 0; k < C::E; ++k) {
        a0[k] = reduce64(a0[k], pc);
        const uint64_t v = reduce64(a1[k], pc);
        a1[k] = v ? pc.q - v : 0;
    }
    uint64_t* out = acc + ((size_t)b * 2 * (l + 1) + t) * C::N;
    const uint64_t* acc_b = acc + (size_t)b * 2 * (l + 1) * C::N;
    uint64_t* out_b = FUSED ? out_ct + ((size_t)b * 2 * (l - 1) + t) * C::N : nullptr;
    const ulonglong2* itw = tw + C::N;
    if constexpr (ntt::pair_pays<C>()) {
        // both accumulators as one transform pair
        ntt::inv_ntt2<C>(a0, a1, lds, itw, pc, tid);
#pragma unroll
        for (int k = 0; k < C::E; ++k) {
            a0[k] = csub(a0[k], pc.q);
            a1[k] = csub(a1[k], pc.q);
        }
        if constexpr (FUSED) {
            fused_out_i<C>(a0, acc_b, out_b, t, l, pcs, map, pidx_p, mc, narrow, tid);
            fused_out_i<C>(a1, acc_b + (size_t)(l + 1) * C::N, out_b + (size_t)(l - 1) * C::N, t, l, pcs, map,
                         pidx_p, mc, narrow, tid);
        } else {
            ntt::store_coalesced<C>(a0, out, tid);
            ntt::store_coalesced<C>(a1, out + (size_t)(l + 1) * C::N, tid);
        }
        return;
    }
    ntt::inv_ntt<C>(a0, lds, itw, pc, tid);
#pragma unroll
    for (int k = 0; k < C::E; ++k) a0[k] = csub(a0[k], pc.q);
    if constexpr (FUSED) fused_out_i<C>(a0, acc_b, out_b, t, l, pcs, map, pidx_p, mc, narrow, tid);
    else ntt::store_coalesced<C>(a0, out, tid);
    __syncthreads();
    ntt::inv_ntt<C>(a1, lds, itw, pc, tid);
#pragma unroll
    for (int k = 0; k < C::E; ++k) a1[k] = csub(a1[k], pc.q);
    if constexpr (FUSED)
        fused_out_i<C>(a1, acc_b + (size_t)(l + 1) * C::N, out_b + (size_t)(l - 1) * C::N, t, l, pcs, map, pidx_p, mc,
                     narrow, tid);
    else ntt::store_coalesced<C>(a1, out + (size_t)(l + 1) * C::N, tid);
}

// Plaintext linear combination in the evaluation domain, the inner loop of
// the diagonal (Halevi-Shoup) matrix-vector product:
//   out[g][c][poly][i][n] = sum_b E[b][c][poly][i][n] * D[g][b][i][n]  (mod q_i)
// E: [nb][B][2][l][N] evaluation-form ciphertext rows, D: [ng][nb][l][N]
// (value, Shoup companion) plaintext rows shared by the whole batch.
// With scalar = 1, D is [ng][nb][l]: one integer per (g, b, limb), the
// same product in either domain (a constant plaintext), so E may then be in
// coefficient form.
// Blocked for reuse: one thread per (block of R rows, limb i, coefficient n)
// keeps NG x R accumulators, so each input value is
// loaded once per call and each plaintext value once per R rows (rows are
// the (ciphertext, poly) pairs, 2 B of them).  ng <= NG; rows past the end
// of the last block are not stored.  Each baby step's input rows are loaded
// one step ahead, and no load sits behind a branch (rows past the end and
// giant steps past ng read clamped, valid addresses).
template <int NG, int R>
__global__ void __launch_bounds__(256) k_lincomb_blk_i(const uint64_t* __restrict__ E,
                                                     const ulonglong2* __restrict__ D, uint64_t* __restrict__ out,
                                                     const PrimeConst* __restrict__ pcs, LimbMap map, int nb, int ng,
                                                     int rows, int l, int logN, size_t total, int scalar) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= total) return;
    const size_t N = (size_t)1 << logN, LN = (size_t)l * N;
    const size_t n = gid & (N - 1), rest = gid >> logN;
    const int i = (int)(rest % l);
    const int r0 = (int)(rest / l) * R;
    const PrimeConst& pc = pcs[map.idx[i]];
    const uint64_t nq = (uint64_t)0 - pc.q;
    const size_t dn = scalar ? 1 : N;
    const size_t dstride = (size_t)l * dn;
    const ulonglong2* dbase = D + (size_t)i * dn + (scalar ? 0 : n);
    const uint64_t* ebase = E + (size_t)r0 * LN + (size_t)i * N + n;
    const size_t brow = (size_t)rows * LN;   // one b (or g) block of rows
    uint64_t acc[NG][R];
#pragma unroll
    for (int g = 0; g < NG; ++g)
#pragma unroll
        for (int r = 0; r < R; ++r) acc[g][r] = 0;
    const int rlast = rows - 1 - r0;
    uint64_t en[R];
#pragma unroll
    for (int r = 0; r < R; ++r) en[r] = ebase[(size_t)(r < rlast ? r : rlast) * LN];
    for (int b = 0; b < nb; ++b) {
        uint64_t e[R];
        ulonglong2 w[NG];
#pragma unroll
        for (int r = 0; r < R; ++r) e[r] = en[r];
#pragma unroll
        for (int g = 0; g < NG; ++g) w[g] = dbase[((size_t)(g < ng ? g : ng - 1) * nb + b) * dstride];
        const int bn = b + 1 < nb ? b + 1 : b;
#pragma unroll
        for (int r = 0; r < R; ++r) en[r] = ebase[(size_t)bn * brow + (size_t)(r < rlast ? r : rlast) * LN];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
#pragma unroll
            for (int r = 0; r < R; ++r) acc[g][r] += mul_shoup4_nq(e[r], w[g].x, w[g].y, nq);   // nb <= 32: < 2^52
        }
    }
    uint64_t* obase = out + (size_t)r0 * LN + (size_t)i * N + n;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        if (g < ng) {
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (r0 + r < rows) obase[(size_t)g * brow + (size_t)r * LN] = reduce64(acc[g][r], pc);
        }
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
    hipError_t e = set_lds(k_tensor_intt_i<LOGN, T>, C::LDS_BYTES);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_tensor_intt_i<LOGN, T>), dim3(batch * l), dim3(T), C::LDS_BYTES, s, ea, eb, d01, d2e, d2c,
                       t.tw, t.pc, map, l, mc);
    return hipGetLastError();
}

template <int LOGN, int T>
hipError_t run_front(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* d01, uint64_t* d2e,
                     uint64_t* d2c, int batch, int l, const LimbMap& map, const ModDownConst& mc, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    constexpr int lds = ntt::pair_pays<C>() ? 2 * C::LDS_BYTES : C::LDS_BYTES;
    hipError_t e = set_lds(k_front_i<LOGN, T>, lds);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((k_front_i<LOGN, T>), dim3(batch * l), dim3(T), lds, s, a, b, d01, d2e, d2c, t.tw, t.pc, map,
                       l, batch, mc);
    return hipGetLastError();
}

template <int LOGN, int T>
hipError_t run_ks(const DevTables& t, const uint64_t* d2c, size_t d2s, const uint64_t* d2e, const uint64_t* d01,
                  const ulonglong2* key, int key_limbs, uint64_t* acc, int batch, int l, const LimbMap& map,
                  int pidx_p, const uint64_t* key_table, const KsTargets& tg, hipStream_t s) {
    using C = Cfg<LOGN, T>;
    constexpr int lds = ntt::pair_pays<C>() ? 2 * C::LDS_BYTES : C::LDS_BYTES;
    auto k = tg.out ? k_keyswitch_i<LOGN, T, false, true>
                    : (key_table ? k_keyswitch_i<LOGN, T, true> : k_keyswitch_i<LOGN, T, false>);
    hipError_t e = set_lds(k, lds);
    if (e != hipSuccess) return e;
    const int nt = tg.count ? tg.count : l + 1;
    ModDownConst mc;
    if (tg.mc) mc = *tg.mc;
    else memset(&mc, 0, sizeof(mc));
    hipLaunchKernelGGL(k, dim3(batch * nt), dim3(T), lds, s, d2c, d2s, d2e, d01, key, key_limbs, acc, t.tw, t.pc, map,
                       l, pidx_p, batch, t.narrow ? 0 : 1, key_table, tg.first, tg.out, mc, t.narrow, tg.pair_major);
    return hipGetLastError();
}

template <int NG, int R>
hipError_t run_lincomb_blk(const DevTables& t, const uint64_t* E, const ulonglong2* D, uint64_t* out, int nb, int ng,
                           int rows, int l, const LimbMap& map, int scalar, hipStream_t s) {
    const size_t total = ((size_t)((rows + R - 1) / R) * l) << t.logN;
    hipLaunchKernelGGL((k_lincomb_blk_i<NG, R>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, E, D, out,
                       t.pc, map, nb, ng, rows, l, t.logN, total, scalar);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_tensor_intt_int(const DevTables& t, const uint64_t* ea, const uint64_t* eb, uint64_t* d01,
                                  uint64_t* d2e, uint64_t* d2c, int batch, int l, const LimbMap& map,
                                  const ModDownConst& mc, int threads, hipStream_t s) {
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

hipError_t launch_front_int(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* d01, uint64_t* d2e,
                            uint64_t* d2c, int batch, int l, const LimbMap& map, const ModDownConst& mc, int threads,
                            hipStream_t s) {
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

hipError_t launch_keyswitch_int(const DevTables& t, const uint64_t* d2c, size_t d2c_bstride, const uint64_t* d2e,
                                const uint64_t* d01, const ulonglong2* key, int key_limbs, uint64_t* acc, int batch,
                                int l, const LimbMap& map, int pidx_p, int threads, hipStream_t s,
                                const uint64_t* key_table, const KsTargets& tg) {
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

hipError_t launch_lincomb_int(const DevTables& t, const uint64_t* E, const ulonglong2* D, uint64_t* out, int nb,
                              int ng, int batch, int l, const LimbMap& map, int scalar, hipStream_t s) {
    const int rows = batch * 2;
    const size_t dn = scalar ? 1 : ((size_t)1 << t.logN);
    const size_t LN = (size_t)l << t.logN;
    // giant steps in chunks of at most 8, as on the FP64 path
    for (int g0 = 0; g0 < ng; g0 += 8) {
        const int gc = ng - g0 < 8 ? ng - g0 : 8;
        const ulonglong2* Dg = D + (size_t)g0 * nb * l * dn;
        uint64_t* og = out + (size_t)g0 * rows * LN;
        hipError_t e;
        if (gc <= 4) e = run_lincomb_blk<4, 4>(t, E, Dg, og, nb, gc, rows, l, map, scalar, s);
        else e = run_lincomb_blk<8, 4>(t, E, Dg, og, nb, gc, rows, l, map, scalar, s);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}
