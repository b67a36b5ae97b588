// LDS-resident negacyclic (or cyclic) transforms on the FP64 pipeline: the
// register layouts, pass plan, exchanges and LDS padding of ntt_core.h, with
// residues held as exact FP64 integers (modf64.h) instead of 64-bit words.
//
// Behavioural spec: the reference's radix-2 transform
// (tb/ref/gen_poly_mult_vectors_n12.py:54-100; rtl/m1/ntt_unit.sv:154-382).
// The outputs are those of the integer transforms of ntt_core.h, residue for
// residue; only the representative of each residue differs (signed, lazy).
//
//  * Forward (Cooley-Tukey, natural order in, bit-reversed out): each stage
//    adds at most 7q/8 to the bound, so from inputs bounded by B q the
//    outputs stay below (B + 7 log2(N) / 8) q; every value entering a product
//    must be below 2^51, so B + 7 log2(N) / 8 < 2^51 / q (at least 64).
//  * Inverse (Gentleman-Sande, N^-1 folded into the last stage): the sums
//    double the bound at each stage.  From inputs bounded by q, and again
//    after each reduction to (q + 1) / 2, at most six stages keep every
//    product input below 2^51; the plan below reduces all registers at the
//    start of the passes where the next pass would exceed that.  Outputs
//    are bounded by 5q/8 (the last stage's N^-1 constants carry their
//    quotients).
//  * Twiddles are the words w as doubles, one 8-byte load each, in the
//    same [prime][forward | inverse][N] order as the integer table; each
//    butterfly forms RN(w qinv) itself (mulk).
#pragma once
#include "modf64.h"
#include "ntt_core.h"

namespace nttf {

using ntt::Cfg;

__device__ __forceinline__ void opaque(const double* __restrict__& tw, int& tid) {
    int zero = 0;
    asm volatile("" : "+v"(tid), "+s"(zero));
    tw += zero;
}

// The twiddles of forward pass P, loaded in one go: for register group g
// and stage b (HI down to LO), the 2^(HI-b) entries 2^(LOGN-1-b) + (fh(g)
// 2^(HI-b) | k), at slot g (M-1) + 2^(HI-b) - 1 + k.  The loads are issued
// between an exchange's LDS writes and its barrier -- the register data is
// dead there -- so their latency overlaps the barrier instead of stalling
// each stage (a wait for one vector load waits for all earlier ones).  A
// pass that needs more than TW_SET twiddles per thread (16 coefficients per
// thread) keeps only its first stage's here and loads the rest as it goes,
// so that the set does not crowd out the kernel's other registers.  The
// tables hold the twiddle words alone (8 bytes each); each butterfly forms
// the quotient RN(w qinv) with one multiply (modf64.h:mulk).
constexpr int TW_SET = 8;
template <class C, int P>
struct FwdTw {
    static constexpr int M = 1 << C::rsize(P), G = C::E >> C::rsize(P);
    static constexpr bool FULL = G * (M - 1) <= TW_SET;
    // slot of stage b's k-th twiddle of group g (first stage only: g)
    static constexpr int slot(int g, int b, int k) { return FULL ? g * (M - 1) + (1 << (C::hi(P) - b)) - 1 + k : g; }
    double w[FULL ? G * (M - 1) : G];
};

template <class C, int P>
__device__ __forceinline__ FwdTw<C, P> fwd_tw(const double* __restrict__ tw, int tid) {
    constexpr int LO = C::lo(P), HI = C::hi(P);
    using S = FwdTw<C, P>;
    S f;
#pragma unroll
    for (int g = 0; g < S::G; ++g) {
        const int fh = ntt::group_of<C, P>(g, tid);
#pragma unroll
        for (int b = HI; b >= (S::FULL ? LO : HI); --b)
#pragma unroll
            for (int k = 0; k < (1 << (HI - b)); ++k)
                f.w[S::slot(g, b, k)] = tw[(1 << (C::LOGN - 1 - b)) + ((fh << (HI - b)) | k)];
    }
    return f;
}

// Twiddle of stage b for butterfly k = e >> (b - LO + 1) of group g: from the
// pass's set, or loaded here when the set holds the first stage only.
template <class C, int P>
__device__ __forceinline__ double fwd_w(const FwdTw<C, P>& set, const double* __restrict__ tw, int g, int b, int k,
                                         int tid) {
    constexpr int HI = C::hi(P);
    if (FwdTw<C, P>::FULL || b == HI) return set.w[FwdTw<C, P>::slot(g, b, k)];
    const int fh = ntt::group_of<C, P>(g, tid);
    return tw[(1 << (C::LOGN - 1 - b)) + ((fh << (HI - b)) | k)];
}

template <class C, int P>
__device__ __forceinline__ void fwd_pass(double (&x)[C::E], const FwdTw<C, P>& set, const double* __restrict__ tw,
                                         const PrimeF& pf, int tid) {
    constexpr int RP = C::rsize(P), LO = C::lo(P), HI = C::hi(P);
    constexpr int M = 1 << RP, G = C::E >> RP;
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int b = HI; b >= LO; --b) {
            const int h = 1 << (b - LO);
#pragma unroll
            for (int e = 0; e < M; ++e) {
                if (e & h) continue;
                const double w = fwd_w<C, P>(set, tw, g, b, e >> (b - LO + 1), tid);
                ct_bfly_f(x[g * M + e], x[g * M + e + h], w, w * pf.qinv, pf.nq);
            }
        }
    }
}

template <class C, int P>
__device__ __forceinline__ void fwd_pass2(double (&x)[C::E], double (&y)[C::E], const FwdTw<C, P>& set,
                                          const double* __restrict__ tw, const PrimeF& pf, int tid) {
    constexpr int RP = C::rsize(P), LO = C::lo(P), HI = C::hi(P);
    constexpr int M = 1 << RP, G = C::E >> RP;
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int b = HI; b >= LO; --b) {
            const int h = 1 << (b - LO);
#pragma unroll
            for (int e = 0; e < M; ++e) {
                if (e & h) continue;
                const double w = fwd_w<C, P>(set, tw, g, b, e >> (b - LO + 1), tid);
                const double wq = w * pf.qinv;
                ct_bfly_f(x[g * M + e], x[g * M + e + h], w, wq, pf.nq);
                ct_bfly_f(y[g * M + e], y[g * M + e + h], w, wq, pf.nq);
            }
        }
    }
}

// Inverse passes run from NP - 1 down to 0.  True when every register is
// reduced at the start of inverse pass P: greedy segments of at most six
// stages, the first starting from inputs bounded by q.
template <class C>
constexpr bool inv_reduce_at(int P) {
    int since = 0;
    for (int p = C::NP - 1; p >= 0; --p) {
        const bool red = since + C::rsize(p) > 6;
        if (red) since = 0;
        if (p == P) return red;
        since += C::rsize(p);
    }
    return false;
}

// The twiddles of inverse pass P (as FwdTw): for group g and stage b (LO up
// to HI, the last stage of the transform excepted: it uses N^-1 constants),
// the 2^(HI-b) entries 2^(LOGN-1-b) + (fh(g) 2^(HI-b) | k) of the inverse
// table; a set of the first stage (b = LO) only when the pass needs more
// than TW_SET.
template <class C, int P>
struct InvTw {
    static constexpr int M = 1 << C::rsize(P), G = C::E >> C::rsize(P);
    static constexpr bool FULL = G * (M - 1) <= TW_SET;
    static constexpr int slot(int g, int b, int k) {
        return FULL ? g * (M - 1) + (1 << (C::hi(P) - b)) - 1 + k : g * (M / 2) + k;
    }
    double w[FULL ? G * (M - 1) : G * (M / 2)];
};

template <class C, int P>
__device__ __forceinline__ InvTw<C, P> inv_tw(const double* __restrict__ itw, int tid) {
    constexpr int LO = C::lo(P), HI = C::hi(P);
    using S = InvTw<C, P>;
    S f;
#pragma unroll
    for (int g = 0; g < S::G; ++g) {
        const int fh = ntt::group_of<C, P>(g, tid);
#pragma unroll
        for (int b = LO; b <= (S::FULL ? HI : LO); ++b) {
            if (b == C::LOGN - 1) continue;
#pragma unroll
            for (int k = 0; k < (1 << (HI - b)); ++k)
                f.w[S::slot(g, b, k)] = itw[(1 << (C::LOGN - 1 - b)) + ((fh << (HI - b)) | k)];
        }
    }
    return f;
}

template <class C, int P>
__device__ __forceinline__ double inv_w(const InvTw<C, P>& set, const double* __restrict__ itw, int g, int b, int k,
                                         int tid) {
    constexpr int LO = C::lo(P), HI = C::hi(P);
    if (InvTw<C, P>::FULL || b == LO) return set.w[InvTw<C, P>::slot(g, b, k)];
    const int fh = ntt::group_of<C, P>(g, tid);
    return itw[(1 << (C::LOGN - 1 - b)) + ((fh << (HI - b)) | k)];
}

template <class C, int P>
__device__ __forceinline__ void inv_pass(double (&x)[C::E], const InvTw<C, P>& set, const double* __restrict__ itw,
                                         const PrimeF& pf, int tid) {
    static_assert(C::rsize(P) <= 6, "an inverse pass longer than one reduction segment");
    constexpr int RP = C::rsize(P), LO = C::lo(P), HI = C::hi(P);
    constexpr int M = 1 << RP, G = C::E >> RP;
    const double nq = pf.nq;
    if constexpr (inv_reduce_at<C>(P)) {
#pragma unroll
        for (int e = 0; e < C::E; ++e) x[e] = rmod(x[e], pf.qinv, nq);
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int b = LO; b <= HI; ++b) {
            const int h = 1 << (b - LO);
#pragma unroll
            for (int e = 0; e < M; ++e) {
                if (e & h) continue;
                double& X = x[g * M + e];
                double& Y = x[g * M + e + h];
                if (b == C::LOGN - 1) {
#pragma clang fp contract(off)
                    const double u = X, v = Y;
                    X = mulw(u + v, pf.ninv, pf.ninv_q, nq);
                    Y = mulw(u - v, pf.ninv_w1, pf.ninv_w1_q, nq);
                } else {
                    const double w = inv_w<C, P>(set, itw, g, b, e >> (b - LO + 1), tid);
                    gs_bfly_f(X, Y, w, w * pf.qinv, nq);
                }
            }
        }
    }
}

template <class C, int P>
__device__ __forceinline__ void inv_pass2(double (&x)[C::E], double (&y)[C::E], const InvTw<C, P>& set,
                                          const double* __restrict__ itw, const PrimeF& pf, int tid) {
    static_assert(C::rsize(P) <= 6, "an inverse pass longer than one reduction segment");
    constexpr int RP = C::rsize(P), LO = C::lo(P), HI = C::hi(P);
    constexpr int M = 1 << RP, G = C::E >> RP;
    const double nq = pf.nq;
    if constexpr (inv_reduce_at<C>(P)) {
#pragma unroll
        for (int e = 0; e < C::E; ++e) {
            x[e] = rmod(x[e], pf.qinv, nq);
            y[e] = rmod(y[e], pf.qinv, nq);
        }
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
        for (int b = LO; b <= HI; ++b) {
            const int h = 1 << (b - LO);
#pragma unroll
            for (int e = 0; e < M; ++e) {
                if (e & h) continue;
                double& X = x[g * M + e];
                double& Y = x[g * M + e + h];
                double& U = y[g * M + e];
                double& V = y[g * M + e + h];
                if (b == C::LOGN - 1) {
#pragma clang fp contract(off)
                    const double x0 = X, x1 = Y, u0 = U, u1 = V;
                    X = mulw(x0 + x1, pf.ninv, pf.ninv_q, nq);
                    Y = mulw(x0 - x1, pf.ninv_w1, pf.ninv_w1_q, nq);
                    U = mulw(u0 + u1, pf.ninv, pf.ninv_q, nq);
                    V = mulw(u0 - u1, pf.ninv_w1, pf.ninv_w1_q, nq);
                } else {
                    const double w = inv_w<C, P>(set, itw, g, b, e >> (b - LO + 1), tid);
                    const double wq = w * pf.qinv;
                    gs_bfly_f(X, Y, w, wq, nq);
                    gs_bfly_f(U, V, w, wq, nq);
                }
            }
        }
    }
}

// ---- single transforms, one exchange buffer --------------------------------

template <class C, int P>
__device__ __forceinline__ void fwd_rest(double (&x)[C::E], double* lds, const double* __restrict__ tw,
                                         const PrimeF& pf, int tid) {
    if constexpr (P < C::NP) {
        constexpr bool W = ntt::wave_local<C, P>();
        if constexpr (P == 1) __syncthreads();
        else if constexpr (ntt::relaid_after_wave_local<C, P - 1, P>()) __syncthreads();
        ntt::to_lds<C, P - 1, ntt::xshift<C>(P)>(x, lds, tid);
        const FwdTw<C, P> w = fwd_tw<C, P>(tw, tid);
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, ntt::xshift<C>(P)>(x, lds, tid);
        ntt::exchange_sync<W>();
        fwd_pass<C, P>(x, w, tw, pf, tid);
        fwd_rest<C, P + 1>(x, lds, tw, pf, tid);
    }
}

// x holds pass-0 layout on entry (x[e] = a[e T + tid]), the last pass's
// layout on exit (bit-reversed evaluation order, see ntt::eval_index).
template <class C>
__device__ __forceinline__ void fwd_ntt(double (&x)[C::E], double* lds, const double* __restrict__ tw, const PrimeF& pf,
                                        int tid) {
    opaque(tw, tid);
    fwd_pass<C, 0>(x, fwd_tw<C, 0>(tw, tid), tw, pf, tid);
    fwd_rest<C, 1>(x, lds, tw, pf, tid);
}

template <class C, int P>
__device__ __forceinline__ void inv_rest(double (&x)[C::E], double* lds, const double* __restrict__ itw,
                                         const PrimeF& pf, int tid) {
    if constexpr (P >= 0) {
        constexpr bool W = ntt::wave_local<C, P + 1>();
        if constexpr (ntt::relaid_after_wave_local<C, P + 2, P + 1>()) __syncthreads();
        ntt::to_lds<C, P + 1, ntt::xshift<C>(P + 1)>(x, lds, tid);
        const InvTw<C, P> w = inv_tw<C, P>(itw, tid);
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, ntt::xshift<C>(P + 1)>(x, lds, tid);
        ntt::exchange_sync<W>();
        inv_pass<C, P>(x, w, itw, pf, tid);
        inv_rest<C, P - 1>(x, lds, itw, pf, tid);
    }
}

// x holds the last forward pass's layout on entry (bounded by q), pass-0
// (coalesced) layout on exit, bounded by 5q/8.
template <class C>
__device__ __forceinline__ void inv_ntt(double (&x)[C::E], double* lds, const double* __restrict__ itw,
                                        const PrimeF& pf, int tid) {
    opaque(itw, tid);
    inv_pass<C, C::NP - 1>(x, inv_tw<C, C::NP - 1>(itw, tid), itw, pf, tid);
    inv_rest<C, C::NP - 2>(x, lds, itw, pf, tid);
}

// ---- two alternating exchange buffers (2 LDS_WORDS of LDS) ------------------
// The ordering argument of ntt::fwd_rest_db / inv_rest_db: an exchange needs
// only its write -> read ordering; a workgroup barrier precedes a write into
// a buffer whose previous exchange was wave-local.  On exit other waves may
// still be reading the LDS: the caller puts a workgroup barrier before any
// other LDS write.

template <class C, int P>
__device__ __forceinline__ void fwd_rest_db(double (&x)[C::E], double* lds, const double* __restrict__ tw,
                                            const PrimeF& pf, int tid) {
    if constexpr (P < C::NP) {
        constexpr bool W = ntt::wave_local<C, P>();
        if constexpr (P == 1) __syncthreads();
        else if constexpr (P >= 3 && ntt::wave_local<C, P - 1>()) __syncthreads();
        double* buf = lds + ((P & 1) ? C::LDS_WORDS : 0);
        ntt::to_lds<C, P - 1, ntt::xshift<C>(P)>(x, buf, tid);
        const FwdTw<C, P> w = fwd_tw<C, P>(tw, tid);
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, ntt::xshift<C>(P)>(x, buf, tid);
        fwd_pass<C, P>(x, w, tw, pf, tid);
        fwd_rest_db<C, P + 1>(x, lds, tw, pf, tid);
    }
}

template <class C>
__device__ __forceinline__ void fwd_ntt_db(double (&x)[C::E], double* lds, const double* __restrict__ tw,
                                           const PrimeF& pf, int tid) {
    static_assert(2 * C::LDS_BYTES <= 160 * 1024, "two exchange buffers exceed the LDS");
    opaque(tw, tid);
    fwd_pass<C, 0>(x, fwd_tw<C, 0>(tw, tid), tw, pf, tid);
    fwd_rest_db<C, 1>(x, lds, tw, pf, tid);
}

template <class C, int P>
__device__ __forceinline__ void inv_rest_db(double (&x)[C::E], double* lds, const double* __restrict__ itw,
                                            const PrimeF& pf, int tid) {
    if constexpr (P >= 0) {
        constexpr bool W = ntt::wave_local<C, P + 1>();
        if constexpr (P == C::NP - 2) __syncthreads();
        else if constexpr (P <= C::NP - 4 && ntt::wave_local<C, P + 2>()) __syncthreads();
        double* buf = lds + ((P & 1) ? C::LDS_WORDS : 0);
        ntt::to_lds<C, P + 1, ntt::xshift<C>(P + 1)>(x, buf, tid);
        const InvTw<C, P> w = inv_tw<C, P>(itw, tid);
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, ntt::xshift<C>(P + 1)>(x, buf, tid);
        inv_pass<C, P>(x, w, itw, pf, tid);
        inv_rest_db<C, P - 1>(x, lds, itw, pf, tid);
    }
}

template <class C>
__device__ __forceinline__ void inv_ntt_db(double (&x)[C::E], double* lds, const double* __restrict__ itw,
                                           const PrimeF& pf, int tid) {
    static_assert(2 * C::LDS_BYTES <= 160 * 1024, "two exchange buffers exceed the LDS");
    opaque(itw, tid);
    inv_pass<C, C::NP - 1>(x, inv_tw<C, C::NP - 1>(itw, tid), itw, pf, tid);
    inv_rest_db<C, C::NP - 2>(x, lds, itw, pf, tid);
}

// The first inverse pass (counting down) from which every remaining pass
// takes its twiddles through scalar loads (ntt::group_of: lowest index bit
// of 6 or more); -1 if the last pass does not.  A vector load issued there
// is waited for by no vector twiddle load of the rest of the transform.
template <class C>
constexpr int inv_scalar_from() {
    int p = -1;
    for (int q = 0; q < C::NP && C::lo(q) >= 6; ++q) p = q;
    return p;
}

// inv_rest_db calling hook() once, at pass inv_scalar_from (after that
// pass's twiddle loads are issued).
template <class C, int P, class H>
__device__ __forceinline__ void inv_rest_db_h(double (&x)[C::E], double* lds, const double* __restrict__ itw,
                                              const PrimeF& pf, int tid, H& hook) {
    if constexpr (P >= 0) {
        constexpr bool W = ntt::wave_local<C, P + 1>();
        if constexpr (P == C::NP - 2) __syncthreads();
        else if constexpr (P <= C::NP - 4 && ntt::wave_local<C, P + 2>()) __syncthreads();
        double* buf = lds + ((P & 1) ? C::LDS_WORDS : 0);
        ntt::to_lds<C, P + 1, ntt::xshift<C>(P + 1)>(x, buf, tid);
        const InvTw<C, P> w = inv_tw<C, P>(itw, tid);
        if constexpr (P == inv_scalar_from<C>()) hook();
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, ntt::xshift<C>(P + 1)>(x, buf, tid);
        inv_pass<C, P>(x, w, itw, pf, tid);
        inv_rest_db_h<C, P - 1>(x, lds, itw, pf, tid, hook);
    }
}

// inv_ntt_db with the first pass's twiddles already loaded: a caller that
// stores results just before the transform loads them first, since a wait
// for a load also waits for every memory operation issued before it.
// hook() runs once, where every later twiddle load is a scalar one (at the
// start if all are, at the end if none is): the place for the caller's
// loads of its next work item.
struct NoHook {
    __device__ void operator()() const {}
};

template <class C, class H>
__device__ __forceinline__ void inv_ntt_db(double (&x)[C::E], double* lds, const double* __restrict__ itw,
                                           const PrimeF& pf, int tid, const InvTw<C, C::NP - 1>& first, H& hook) {
    static_assert(2 * C::LDS_BYTES <= 160 * 1024, "two exchange buffers exceed the LDS");
    constexpr int PS = inv_scalar_from<C>();
    opaque(itw, tid);
    if constexpr (PS >= C::NP - 1) hook();
    inv_pass<C, C::NP - 1>(x, first, itw, pf, tid);
    inv_rest_db_h<C, C::NP - 2>(x, lds, itw, pf, tid, hook);
    if constexpr (PS < 0) hook();
}

template <class C>
__device__ __forceinline__ void inv_ntt_db(double (&x)[C::E], double* lds, const double* __restrict__ itw,
                                           const PrimeF& pf, int tid, const InvTw<C, C::NP - 1>& first) {
    NoHook none;
    inv_ntt_db<C>(x, lds, itw, pf, tid, first, none);
}

// ---- two transforms of one prime, software-pipelined -----------------------
// y runs half a pass behind x, each polynomial in its own exchange buffer
// with its own barrier, so the LDS traffic of one overlaps the butterflies
// of the other:
//   for P = 1 .. NP-1:  write x (layout P-1); y's pass P-1; load pass P's
//                       twiddles; sync x; read x (layout P); write y (layout
//                       P-1); x's pass P; sync y; read y (layout P)
//   then y's last pass.
// One twiddle set is live at a time: pass P's serves x's pass P and then y's
// (in the next round), and is replaced only after that.  Buffer reuse: a
// wave writes a buffer in round P into the slots of the layout it read from
// that buffer in round P-1, under round P's padding.  With the same padding
// those are the very slots it read itself; with a different one they may be
// slots another wave has still to read, so a workgroup sync must separate
// the two.  For x the sync in between is round P-1's y sync (a workgroup
// barrier unless exchange P-1 is wave-local); for y it is round P's x sync
// (unless exchange P is wave-local).  Where the sync in between is
// wave-level and the padding changes, a workgroup barrier comes first.
template <class C, int X>
constexpr bool y_relaid() {   // forward round X: y's write needs a workgroup barrier first
    if constexpr (X < 2 || X >= C::NP) return false;
    else return ntt::wave_local<C, X>() && ntt::xshift<C>(X) != ntt::xshift<C>(X - 1);
}
template <class C, int X>
constexpr bool y_relaid_inv() {   // inverse round through exchange X
    if constexpr (X < 1 || X + 1 >= C::NP) return false;
    else return ntt::wave_local<C, X>() && ntt::xshift<C>(X) != ntt::xshift<C>(X + 1);
}

template <class C, int P>
__device__ __forceinline__ void fwd_rest2(double (&x)[C::E], double (&y)[C::E], double* lds,
                                          const double* __restrict__ tw, const FwdTw<C, P - 1>& prev, const PrimeF& pf,
                                          int tid) {
    if constexpr (P < C::NP) {
        constexpr bool W = ntt::wave_local<C, P>();
        constexpr int K = ntt::xshift<C>(P);
        double* bx = lds;
        double* by = lds + C::LDS_WORDS;
        if constexpr (P == 1) __syncthreads();
        else if constexpr (ntt::relaid_after_wave_local<C, P - 1, P>()) __syncthreads();
        ntt::to_lds<C, P - 1, K>(x, bx, tid);
        fwd_pass<C, P - 1>(y, prev, tw, pf, tid);
        const FwdTw<C, P> w = fwd_tw<C, P>(tw, tid);
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, K>(x, bx, tid);
        if constexpr (y_relaid<C, P>()) __syncthreads();
        ntt::to_lds<C, P - 1, K>(y, by, tid);
        fwd_pass<C, P>(x, w, tw, pf, tid);
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, K>(y, by, tid);
        fwd_rest2<C, P + 1>(x, y, lds, tw, w, pf, tid);
    } else {
        fwd_pass<C, C::NP - 1>(y, prev, tw, pf, tid);
    }
}

// fwd_ntt on x and y at once (same entry and exit conditions).  On exit
// other waves may still be reading y's buffer: a workgroup barrier precedes
// any further LDS write (the transforms here start with one).
template <class C>
__device__ __forceinline__ void fwd_ntt2(double (&x)[C::E], double (&y)[C::E], double* lds,
                                         const double* __restrict__ tw, const PrimeF& pf, int tid) {
    static_assert(ntt::pair_fits<C>(), "two exchange buffers exceed the LDS");
    opaque(tw, tid);
    const FwdTw<C, 0> w0 = fwd_tw<C, 0>(tw, tid);
    fwd_pass<C, 0>(x, w0, tw, pf, tid);
    fwd_rest2<C, 1>(x, y, lds, tw, w0, pf, tid);
}

template <class C, int P>
__device__ __forceinline__ void inv_rest2(double (&x)[C::E], double (&y)[C::E], double* lds,
                                          const double* __restrict__ itw, const InvTw<C, P + 1>& prev,
                                          const PrimeF& pf, int tid) {
    if constexpr (P >= 0) {
        constexpr bool W = ntt::wave_local<C, P + 1>();
        constexpr int K = ntt::xshift<C>(P + 1);
        double* bx = lds;
        double* by = lds + C::LDS_WORDS;
        if constexpr (P == C::NP - 2) __syncthreads();
        else if constexpr (ntt::relaid_after_wave_local<C, P + 2, P + 1>()) __syncthreads();
        ntt::to_lds<C, P + 1, K>(x, bx, tid);
        inv_pass<C, P + 1>(y, prev, itw, pf, tid);
        const InvTw<C, P> w = inv_tw<C, P>(itw, tid);
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, K>(x, bx, tid);
        if constexpr (y_relaid_inv<C, P + 1>()) __syncthreads();
        ntt::to_lds<C, P + 1, K>(y, by, tid);
        inv_pass<C, P>(x, w, itw, pf, tid);
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, K>(y, by, tid);
        inv_rest2<C, P - 1>(x, y, lds, itw, w, pf, tid);
    } else {
        inv_pass<C, 0>(y, prev, itw, pf, tid);
    }
}

// inv_rest2 calling hook() once, at round inv_scalar_from (after that
// round's twiddle loads are issued).
template <class C, int P, class H>
__device__ __forceinline__ void inv_rest2_h(double (&x)[C::E], double (&y)[C::E], double* lds,
                                            const double* __restrict__ itw, const InvTw<C, P + 1>& prev,
                                            const PrimeF& pf, int tid, H& hook) {
    if constexpr (P >= 0) {
        constexpr bool W = ntt::wave_local<C, P + 1>();
        constexpr int K = ntt::xshift<C>(P + 1);
        double* bx = lds;
        double* by = lds + C::LDS_WORDS;
        if constexpr (P == C::NP - 2) __syncthreads();
        else if constexpr (ntt::relaid_after_wave_local<C, P + 2, P + 1>()) __syncthreads();
        ntt::to_lds<C, P + 1, K>(x, bx, tid);
        inv_pass<C, P + 1>(y, prev, itw, pf, tid);
        const InvTw<C, P> w = inv_tw<C, P>(itw, tid);
        if constexpr (P == inv_scalar_from<C>()) hook();
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, K>(x, bx, tid);
        if constexpr (y_relaid_inv<C, P + 1>()) __syncthreads();
        ntt::to_lds<C, P + 1, K>(y, by, tid);
        inv_pass<C, P>(x, w, itw, pf, tid);
        ntt::exchange_sync<W>();
        ntt::from_lds<C, P, K>(y, by, tid);
        inv_rest2_h<C, P - 1>(x, y, lds, itw, w, pf, tid, hook);
    } else {
        inv_pass<C, 0>(y, prev, itw, pf, tid);
    }
}

// inv_ntt on x and y at once (same entry and exit conditions; the first
// exchange starts with a workgroup barrier).  On exit other waves may still
// be reading y's buffer: a workgroup barrier precedes any further LDS write.
template <class C>
__device__ __forceinline__ void inv_ntt2(double (&x)[C::E], double (&y)[C::E], double* lds,
                                         const double* __restrict__ itw, const PrimeF& pf, int tid) {
    static_assert(ntt::pair_fits<C>(), "two exchange buffers exceed the LDS");
    opaque(itw, tid);
    const InvTw<C, C::NP - 1> w = inv_tw<C, C::NP - 1>(itw, tid);
    inv_pass<C, C::NP - 1>(x, w, itw, pf, tid);
    inv_rest2<C, C::NP - 2>(x, y, lds, itw, w, pf, tid);
}

// inv_ntt2 with hook() run once where every later twiddle load is a scalar
// one (as the single-transform inv_ntt_db above).
template <class C, class H>
__device__ __forceinline__ void inv_ntt2(double (&x)[C::E], double (&y)[C::E], double* lds,
                                         const double* __restrict__ itw, const PrimeF& pf, int tid, H& hook) {
    static_assert(ntt::pair_fits<C>(), "two exchange buffers exceed the LDS");
    constexpr int PS = inv_scalar_from<C>();
    opaque(itw, tid);
    const InvTw<C, C::NP - 1> w = inv_tw<C, C::NP - 1>(itw, tid);
    if constexpr (PS >= C::NP - 1) hook();
    inv_pass<C, C::NP - 1>(x, w, itw, pf, tid);
    inv_rest2_h<C, C::NP - 2>(x, y, lds, itw, w, pf, tid, hook);
    if constexpr (PS < 0) hook();
}

// The built paired shapes: at N = 4096 with 512 threads exchanges 2 and 3
// are wave-local and exchange 2 pads differently from exchange 1 (forward
// pass 2 needs the barrier); at N = 8192 with 1024 threads the one
// wave-local exchange (4) keeps exchange 3's padding.
static_assert(y_relaid<Cfg<12, 512>, 2>() && !y_relaid<Cfg<12, 512>, 3>());
static_assert(!y_relaid_inv<Cfg<12, 512>, 2>());
static_assert(!y_relaid<Cfg<13, 1024>, 4>() && !y_relaid<Cfg<13, 1024>, 3>());

// ---- HBM <-> registers -----------------------------------------------------

// Issue every load of a row before any of its words is used: the empty
// asm (a compiler memory barrier) keeps the loads above it, so the waits for
// them come after the last one is issued instead of between small groups.
template <class C>
__device__ __forceinline__ void issue_row(uint64_t (&u)[C::E], const uint64_t* __restrict__ src, int tid) {
#pragma unroll
    for (int e = 0; e < C::E; ++e) u[e] = src[e * C::T + tid];
    asm volatile("" ::: "memory");
}

// uint64 words below 2^52 in pass-0 (coalesced) layout.
template <class C>
__device__ __forceinline__ void load_words(double (&x)[C::E], const uint64_t* __restrict__ src, int tid) {
    uint64_t u[C::E];
    issue_row<C>(u, src, tid);
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = u2d(u[e]);
}

// Canonical words (bounded by 16 q on entry) in pass-0 layout.
template <class C>
__device__ __forceinline__ void store_words(const double (&x)[C::E], uint64_t* __restrict__ dst, const PrimeF& pf,
                                            int tid) {
#pragma unroll
    for (int e = 0; e < C::E; ++e) dst[e * C::T + tid] = d2u(canon(x[e], pf.qinv, pf.nq));
}

}  // namespace nttf
