// LDS-resident negacyclic (or cyclic) NTT building blocks for one transform
// per workgroup.
//
// Behavioural spec: the reference's radix-2 transform
// (tb/ref/gen_poly_mult_vectors_n12.py:54-100; rtl/m1/ntt_unit.sv:154-382),
// re-laid out for a 64-wide wavefront machine:
//
//  * T threads, each holding E = N/T coefficients in registers.
//  * The log2(N) radix-2 stages are grouped into passes of at most log2(E)
//    consecutive bits.  Inside a pass every butterfly is register-to-register;
//    between passes the coefficients are exchanged once through LDS.
//  * Pass 0 covers the top log2(E) index bits, so its register layout
//    x[e] = a[e*T + tid] is also the coalesced HBM layout.  The inverse
//    transform runs the same passes in reverse order and ends in that layout.
//  * Forward: Cooley-Tukey, natural order in, bit-reversed order out, twiddle
//    for stage s / group i at tw[2^s + i] (psi^bitrev(2^s+i) for the
//    negacyclic ring).  Inverse: Gentleman-Sande with the inverse table and
//    N^-1 folded into the last stage.  Twiddles are (w, floor(w 2^63/q))
//    pairs, one 16-byte load each.
//  * LDS index pad(j) = j + (j >> PS) keeps the small-stride pass
//    conflict-free (PS = size of the last forward pass).
#pragma once
#include "modarith.h"

namespace ntt {

constexpr int ilog2c(int n) { return n <= 1 ? 0 : 1 + ilog2c(n >> 1); }

template <int LOGN_, int T_>
struct Cfg {
    static constexpr int LOGN = LOGN_;
    static constexpr int N = 1 << LOGN;
    static constexpr int T = T_;
    static constexpr int E = N / T;
    static constexpr int R = ilog2c(E);
    static constexpr int REST = LOGN - R;
    static constexpr int NREST = (REST + R - 1) / R;
    static constexpr int NP = 1 + NREST;
    static_assert(E >= 2 && (E & (E - 1)) == 0, "E must be a power of two");
    static_assert(REST > 0, "need at least two passes");
    static constexpr int rsize(int p) {
        return p == 0 ? R : (REST / NREST + ((p - 1) < (REST % NREST) ? 1 : 0));
    }
    static constexpr int hi(int p) {
        int h = LOGN - 1;
        for (int k = 0; k < p; ++k) h -= rsize(k);
        return h;
    }
    static constexpr int lo(int p) { return hi(p) - rsize(p) + 1; }
    static constexpr int PS = rsize(NP - 1);
    static constexpr int LDS_WORDS = N + (N >> PS);
    static constexpr int LDS_BYTES = LDS_WORDS * 8;
};

template <class C, int K = C::PS>
__host__ __device__ __forceinline__ int pad(int j) { return j + (j >> K); }

// Padding shift of the exchange between the layouts of passes P - 1 and P
// (either direction).  Every exchange may pad differently: its two sides
// are its own.  Where a pass's lanes cover a long contiguous index run, the
// default pad(j) = j + (j >> PS) spreads 32 lanes of 8-byte words over more
// than the LDS's banks and doubles the cycles of every access; a coarser
// shift keeps the run compact.  The shifts below come from a bank model of
// every layout pair of the hot shapes (32 lanes of 8-byte words per cycle
// over 64 four-byte banks, and 16 over 32); others keep PS.  The shift must
// leave each layout's register field wholly above or below it, so that the
// slots stay affine in the registers (Lay).
template <class C>
constexpr int xshift(int P) {
    if (C::LOGN == 13 && C::T == 1024) return P == 1 ? 5 : P == 2 ? 7 : C::PS;
    if ((C::LOGN == 13 || C::LOGN == 12) && C::T == 512) return P == 1 ? 5 : C::PS;
    if (C::LOGN == 14 && C::T == 1024) return P == 1 ? 5 : C::PS;
    if (C::LOGN == 14 && C::T == 512) return P == 1 ? 9 : C::PS;
    return C::PS;
}

// Hide the twiddle table (a workgroup-uniform pointer) and the thread index
// from the optimiser at the start of every transform.  Without this, a kernel
// that runs several transforms with one table (key switching: one per digit)
// gets every twiddle address -- or the twiddles themselves -- hoisted and
// kept live across all of them, and spills.  The statement emits no code.
__device__ __forceinline__ void opaque(const ulonglong2* __restrict__& tw, int& tid) {
    int zero = 0;   // an opaque offset keeps the table a global (not flat) pointer
    asm volatile("" : "+v"(tid), "+s"(zero));
    tw += zero;
}

// Global index of register (g, e) in the layout of pass P.
template <class C, int P>
__host__ __device__ __forceinline__ int index_of(int g, int e, int tid) {
    constexpr int LO = C::lo(P), HI = C::hi(P);
    const int f = g * C::T + tid;
    return ((f >> LO) << (HI + 1)) | (e << LO) | (f & ((1 << LO) - 1));
}

// Twiddle group of register group g in pass P's layout.  When the pass's
// low bit LO is at least 6, all 64 lanes of a wave share it (T is a
// multiple of 64, and lanes differ only in the low 6 bits of g T + tid):
// taking it from the first lane makes the pass's twiddle addresses
// wave-uniform, so they are fetched by scalar loads into SGPRs (several
// twiddles per load) instead of one vector load per butterfly.
template <class C, int P>
__device__ __forceinline__ int group_of(int g, int tid) {
    constexpr int LO = C::lo(P);
    static_assert(C::T % 64 == 0, "workgroups are whole wavefronts");
    const int fh = (g * C::T + tid) >> LO;
    if constexpr (LO >= 6) return __builtin_amdgcn_readfirstlane(fh);
    else return fh;
}

// The twiddles of forward pass P's first stage (bit HI), one per register
// group: every butterfly of group g in that stage uses entry
// 2^(LOGN-1-HI) + fh(g).  Loaded before the exchange into the pass, their
// latency overlaps the exchange instead of stalling the pass's first stage.
template <class C, int P>
struct FwdFirst {
    ulonglong2 w[C::E >> C::rsize(P)];
};

template <class C, int P>
__device__ __forceinline__ FwdFirst<C, P> fwd_first(const ulonglong2* __restrict__ tw, int tid) {
    constexpr int G = C::E >> C::rsize(P);
    constexpr int base = 1 << (C::LOGN - 1 - C::hi(P));
    FwdFirst<C, P> f;
#pragma unroll
    for (int g = 0; g < G; ++g) f.w[g] = tw[base + group_of<C, P>(g, tid)];
    return f;
}

template <class C, int P, bool PRE>
__device__ __forceinline__ void fwd_pass_t(uint64_t (&x)[C::E], const ulonglong2* __restrict__ tw,
                                           uint64_t q, uint64_t two_q, int tid, const FwdFirst<C, P>& pre) {
    const uint64_t nq = (uint64_t)0 - q, four_q = two_q << 1;
    constexpr int RP = C::rsize(P), LO = C::lo(P), HI = C::hi(P);
    constexpr int M = 1 << RP, G = C::E >> RP;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int fh = group_of<C, P>(g, tid);
#pragma unroll
        for (int b = HI; b >= LO; --b) {
            const int h = 1 << (b - LO);
            const int base = 1 << (C::LOGN - 1 - b);
#pragma unroll
            for (int e = 0; e < M; ++e) {
                if (e & h) continue;
                const int i = (fh << (HI - b)) | (e >> (b - LO + 1));
                const ulonglong2 w = (PRE && b == HI) ? pre.w[g] : tw[base + i];
                ct_bfly_lazy(x[g * M + e], x[g * M + e + h], w.x, w.y, nq, four_q);
            }
        }
    }
}

template <class C, int P>
__device__ __forceinline__ void fwd_pass(uint64_t (&x)[C::E], const ulonglong2* __restrict__ tw,
                                         uint64_t q, uint64_t two_q, int tid) {
    fwd_pass_t<C, P, false>(x, tw, q, two_q, tid, FwdFirst<C, P>{});
}

// The twiddles of inverse pass P's first stage (bit LO): M / 2 per register
// group, entry 2^(LOGN-1-LO) + (fh(g) 2^(HI-LO) | e / 2) for even e.  As for
// the forward passes, loaded before the exchange into the pass.
template <class C, int P>
struct InvFirst {
    ulonglong2 w[C::E / 2];
};

template <class C, int P>
__device__ __forceinline__ InvFirst<C, P> inv_first(const ulonglong2* __restrict__ itw, int tid) {
    constexpr int RP = C::rsize(P), LO = C::lo(P), HI = C::hi(P);
    constexpr int M = 1 << RP, G = C::E >> RP;
    constexpr int base = 1 << (C::LOGN - 1 - LO);
    InvFirst<C, P> f;
    if constexpr (LO != C::LOGN - 1) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int fh = group_of<C, P>(g, tid);
#pragma unroll
            for (int e = 0; e < M; e += 2) f.w[g * (M / 2) + e / 2] = itw[base + ((fh << (HI - LO)) | (e >> 1))];
        }
    }
    return f;
}

template <class C, int P, bool PRE = false>
__device__ __forceinline__ void inv_pass(uint64_t (&x)[C::E], const ulonglong2* __restrict__ itw,
                                         const PrimeConst& pc, int tid, const InvFirst<C, P>& pre = InvFirst<C, P>{}) {
    constexpr int RP = C::rsize(P), LO = C::lo(P), HI = C::hi(P);
    constexpr int M = 1 << RP, G = C::E >> RP;
    const uint64_t q = pc.q, nq = (uint64_t)0 - q;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int fh = group_of<C, P>(g, tid);
#pragma unroll
        for (int b = LO; b <= HI; ++b) {
            const int h = 1 << (b - LO);
            const int base = 1 << (C::LOGN - 1 - b);
            const uint64_t bound = q << (b + 1);   // every input of stage b is below 2^(b+1) q
#pragma unroll
            for (int e = 0; e < M; ++e) {
                if (e & h) continue;
                uint64_t& X = x[g * M + e];
                uint64_t& Y = x[g * M + e + h];
                if (b == C::LOGN - 1) {
                    // last stage: fold N^-1 into both outputs
                    const uint64_t u = X, v = Y;
                    X = mul_shoup_nq(u + v, pc.ninv, pc.ninv_s, nq);
                    Y = mul_shoup_nq(u + bound - v, pc.ninv_w1, pc.ninv_w1_s, nq);
                } else {
                    const int i = (fh << (HI - b)) | (e >> (b - LO + 1));
                    const ulonglong2 w = (PRE && b == LO) ? pre.w[g * (M / 2) + e / 2] : itw[base + i];
                    gs_bfly_lazy(X, Y, w.x, w.y, nq, bound);
                }
            }
        }
    }
}

// Padded LDS slot of register (g, e) in pass P's layout, as base(tid) +
// g SG + e SE.  The e bits sit at LO..HI of the index and the g bits above
// HI; both are multiples of 2^PS whenever LO >= PS, so the padding adds
// e 2^(LO-PS) and g-part >> PS exactly; in the last pass (LO = 0, e < 2^PS)
// the padding does not see e at all.  (T is a multiple of 2^LO in every
// built configuration.)  Affine slots let the LDS instructions carry the
// per-register part as an immediate offset.
template <class C, int P, int K = C::PS>
struct Lay {
    static constexpr int LO = C::lo(P), HI = C::hi(P);
    static_assert(C::T % (1 << LO) == 0, "threads must cover the pass's low bits");
    static_assert(K >= C::PS, "padding beyond the allocated LDS words");
    static_assert(LO >= K || HI < K, "the register field straddles the padding shift");
    static constexpr int RAW_G = (C::T >> LO) << (HI + 1);
    static_assert(C::E == (1 << C::rsize(P)) || RAW_G % (1 << K) == 0, "group stride not a multiple of 2^K");
    static constexpr int SE = (1 << LO) + (LO >= K ? (1 << (LO - K)) : 0);
    static constexpr int SG = RAW_G + (RAW_G >> K);
};

// W is the 8-byte word of the registers and the LDS: uint64_t (integer
// residues) or double (residues held as exact FP64 integers).
template <class C, int P, int K = C::PS, class W = uint64_t>
__device__ __forceinline__ void to_lds(const W (&x)[C::E], W* lds, int tid) {
    constexpr int M = 1 << C::rsize(P), G = C::E >> C::rsize(P);
    W* base = lds + pad<C, K>(index_of<C, P>(0, 0, tid));
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int e = 0; e < M; ++e) base[g * Lay<C, P, K>::SG + e * Lay<C, P, K>::SE] = x[g * M + e];
}

template <class C, int P, int K = C::PS, class W = uint64_t>
__device__ __forceinline__ void from_lds(W (&x)[C::E], const W* lds, int tid) {
    constexpr int M = 1 << C::rsize(P), G = C::E >> C::rsize(P);
    const W* base = lds + pad<C, K>(index_of<C, P>(0, 0, tid));
#pragma unroll
    for (int g = 0; g < G; ++g)
#pragma unroll
        for (int e = 0; e < M; ++e) x[g * M + e] = base[g * Lay<C, P, K>::SG + e * Lay<C, P, K>::SE];
}

// True when the exchange between the layouts of passes P - 1 and P stays
// inside every 64-lane wavefront: each slot a wave reads in pass P's layout
// was written by that same wave in pass P - 1's layout.  (Each wave writes
// and reads 64 E slots, so the inclusion makes the two sets equal, and the
// relation is the same in the inverse direction.)
template <class C, int P>
constexpr bool wave_local() {
    constexpr int LOa = C::lo(P - 1), HIa = C::hi(P - 1);
    constexpr int LOb = C::lo(P), HIb = C::hi(P), RPb = C::rsize(P);
    constexpr int Mb = 1 << RPb, Gb = C::E >> RPb;
    for (int tid = 0; tid < C::T; ++tid)
        for (int g = 0; g < Gb; ++g)
            for (int e = 0; e < Mb; ++e) {
                const int f = g * C::T + tid;
                const int j = ((f >> LOb) << (HIb + 1)) | (e << LOb) | (f & ((1 << LOb) - 1));
                const int fa = ((j >> (HIa + 1)) << LOa) | (j & ((1 << LOa) - 1));
                if ((fa % C::T) / 64 != tid / 64) return false;
            }
    return true;
}

// LDS ordering for an exchange: a wave-local one needs this wave's writes
// complete and visible to itself only; any other one needs the workgroup.
template <bool WAVE>
__device__ __forceinline__ void exchange_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    } else {
        __syncthreads();
    }
}

// An exchange that writes into the buffer the previous exchange used needs a
// workgroup barrier first when that previous exchange was wave-local (it
// ends with a wave-level fence only) and the two pad differently: each wave
// then writes slots that other waves may not have read yet.  With the same
// padding a wave rewrites exactly the slots it read itself.  Exchange XP
// (layouts XP - 1 and XP) was the previous one, exchange XN is next.
template <class C, int XP, int XN>
constexpr bool relaid_after_wave_local() {
    if constexpr (XP < 1 || XP >= C::NP) return false;
    else return wave_local<C, XP>() && xshift<C>(XP) != xshift<C>(XN);
}

// Forward passes 1..NP-1 (pass 0 already applied), exchanging through LDS.
// The first exchange starts with a workgroup barrier: a preceding transform
// in the same kernel may end with a wave-local exchange, and pass 0's slots
// of one wave are other waves' slots there.
template <class C, int P>
__device__ __forceinline__ void fwd_rest(uint64_t (&x)[C::E], uint64_t* lds,
                                         const ulonglong2* __restrict__ tw, uint64_t q,
                                         uint64_t two_q, int tid) {
    if constexpr (P < C::NP) {
        constexpr bool W = wave_local<C, P>();
        const FwdFirst<C, P> pre = fwd_first<C, P>(tw, tid);
        if constexpr (P == 1) __syncthreads();
        else if constexpr (relaid_after_wave_local<C, P - 1, P>()) __syncthreads();
        to_lds<C, P - 1, xshift<C>(P)>(x, lds, tid);
        exchange_sync<W>();
        from_lds<C, P, xshift<C>(P)>(x, lds, tid);
        exchange_sync<W>();
        fwd_pass_t<C, P, true>(x, tw, q, two_q, tid, pre);
        fwd_rest<C, P + 1>(x, lds, tw, q, two_q, tid);
    }
}

// Full forward transform: x holds pass-0 layout on entry (x[e] = a[e*T+tid],
// values below 2q), the last pass's layout on exit, values below
// (2 + 4 log2 N) q -- reduce with reduce64.
template <class C>
__device__ __forceinline__ void fwd_ntt(uint64_t (&x)[C::E], uint64_t* lds,
                                        const ulonglong2* __restrict__ tw, uint64_t q,
                                        uint64_t two_q, int tid) {
    opaque(tw, tid);
    fwd_pass<C, 0>(x, tw, q, two_q, tid);
    fwd_rest<C, 1>(x, lds, tw, q, two_q, tid);
}

// Forward passes with two alternating exchange buffers (pass P uses buffer
// P & 1), for kernels that own 2 LDS_WORDS of LDS.  An exchange then needs
// only its write -> read ordering: the next write into a buffer comes two
// exchanges later, and every wave has finished reading it by the time it
// crosses the workgroup barrier of the exchange in between.  Where that
// exchange is wave-local, a workgroup barrier is put before the write.
template <class C, int P>
__device__ __forceinline__ void fwd_rest_db(uint64_t (&x)[C::E], uint64_t* lds,
                                            const ulonglong2* __restrict__ tw, uint64_t q,
                                            uint64_t two_q, int tid) {
    if constexpr (P < C::NP) {
        constexpr bool W = wave_local<C, P>();
        if constexpr (P == 1) __syncthreads();
        else if constexpr (P >= 3 && wave_local<C, P - 1>()) __syncthreads();
        uint64_t* buf = lds + ((P & 1) ? C::LDS_WORDS : 0);
        const FwdFirst<C, P> pre = fwd_first<C, P>(tw, tid);
        to_lds<C, P - 1, xshift<C>(P)>(x, buf, tid);
        exchange_sync<W>();
        from_lds<C, P, xshift<C>(P)>(x, buf, tid);
        fwd_pass_t<C, P, true>(x, tw, q, two_q, tid, pre);
        fwd_rest_db<C, P + 1>(x, lds, tw, q, two_q, tid);
    }
}

// fwd_ntt through two alternating exchange buffers (2 LDS_WORDS of LDS):
// same entry and exit conditions, but on exit other waves may still be
// reading the LDS, so the caller puts a workgroup barrier before any other
// LDS write (a following fwd_ntt_db starts with one).
template <class C>
__device__ __forceinline__ void fwd_ntt_db(uint64_t (&x)[C::E], uint64_t* lds,
                                           const ulonglong2* __restrict__ tw, uint64_t q,
                                           uint64_t two_q, int tid) {
    static_assert(2 * C::LDS_BYTES <= 160 * 1024, "two exchange buffers exceed the LDS");
    opaque(tw, tid);
    fwd_pass<C, 0>(x, tw, q, two_q, tid);
    fwd_rest_db<C, 1>(x, lds, tw, q, two_q, tid);
}

template <class C, int P>
__device__ __forceinline__ void inv_rest(uint64_t (&x)[C::E], uint64_t* lds,
                                         const ulonglong2* __restrict__ itw, const PrimeConst& pc,
                                         int tid) {
    if constexpr (P >= 0) {
        constexpr bool W = wave_local<C, P + 1>();
        const InvFirst<C, P> pre = inv_first<C, P>(itw, tid);
        if constexpr (relaid_after_wave_local<C, P + 2, P + 1>()) __syncthreads();
        to_lds<C, P + 1, xshift<C>(P + 1)>(x, lds, tid);
        exchange_sync<W>();
        from_lds<C, P, xshift<C>(P + 1)>(x, lds, tid);
        exchange_sync<W>();
        inv_pass<C, P, true>(x, itw, pc, tid, pre);
        inv_rest<C, P - 1>(x, lds, itw, pc, tid);
    }
}

// Inverse passes with two alternating exchange buffers (the exchange into
// pass P's layout uses buffer P & 1); the same ordering argument as
// fwd_rest_db, read in the other direction.  The first exchange starts with
// a workgroup barrier, since what used the LDS before is not known here.
template <class C, int P>
__device__ __forceinline__ void inv_rest_db(uint64_t (&x)[C::E], uint64_t* lds,
                                            const ulonglong2* __restrict__ itw, const PrimeConst& pc,
                                            int tid) {
    if constexpr (P >= 0) {
        constexpr bool W = wave_local<C, P + 1>();
        if constexpr (P == C::NP - 2) __syncthreads();
        else if constexpr (P <= C::NP - 4 && wave_local<C, P + 2>()) __syncthreads();
        uint64_t* buf = lds + ((P & 1) ? C::LDS_WORDS : 0);
        const InvFirst<C, P> pre = inv_first<C, P>(itw, tid);
        to_lds<C, P + 1, xshift<C>(P + 1)>(x, buf, tid);
        exchange_sync<W>();
        from_lds<C, P, xshift<C>(P + 1)>(x, buf, tid);
        inv_pass<C, P, true>(x, itw, pc, tid, pre);
        inv_rest_db<C, P - 1>(x, lds, itw, pc, tid);
    }
}

// inv_ntt through two alternating exchange buffers (2 LDS_WORDS of LDS):
// same entry and exit conditions; on exit other waves may still be reading
// the LDS, so the caller puts a workgroup barrier before any further LDS
// write.
template <class C>
__device__ __forceinline__ void inv_ntt_db(uint64_t (&x)[C::E], uint64_t* lds,
                                           const ulonglong2* __restrict__ itw, const PrimeConst& pc,
                                           int tid) {
    static_assert(2 * C::LDS_BYTES <= 160 * 1024, "two exchange buffers exceed the LDS");
    opaque(itw, tid);
    inv_pass<C, C::NP - 1>(x, itw, pc, tid);
    inv_rest_db<C, C::NP - 2>(x, lds, itw, pc, tid);
}

// inv_ntt_db with the first pass's first-stage twiddles already loaded
// (inv_first<C, NP-1>): a caller that stores results just before the
// transform loads them first, since a wait for a load also waits for every
// memory operation issued before it.
template <class C>
__device__ __forceinline__ void inv_ntt_db(uint64_t (&x)[C::E], uint64_t* lds,
                                           const ulonglong2* __restrict__ itw, const PrimeConst& pc, int tid,
                                           const InvFirst<C, C::NP - 1>& pre) {
    static_assert(2 * C::LDS_BYTES <= 160 * 1024, "two exchange buffers exceed the LDS");
    opaque(itw, tid);
    inv_pass<C, C::NP - 1, true>(x, itw, pc, tid, pre);
    inv_rest_db<C, C::NP - 2>(x, lds, itw, pc, tid);
}

// Full inverse transform: x holds the last forward pass's layout on entry
// (values below 2q), pass-0 (coalesced) layout on exit, values in [0, 2q).
template <class C>
__device__ __forceinline__ void inv_ntt(uint64_t (&x)[C::E], uint64_t* lds,
                                        const ulonglong2* __restrict__ itw, const PrimeConst& pc,
                                        int tid) {
    opaque(itw, tid);
    inv_pass<C, C::NP - 1>(x, itw, pc, tid);
    inv_rest<C, C::NP - 2>(x, lds, itw, pc, tid);
}

// ---------------------------------------------------------------------------
// Two transforms of the same prime, interleaved.  Every twiddle load serves
// the corresponding butterflies of both, the two register sets give the
// scheduler independent work, and both polynomials cross each exchange
// between one pair of barriers.  lds holds two exchange buffers of
// LDS_WORDS each (PAIR_LDS_BYTES).  Results are those of two single
// transforms, value for value.
template <class C>
constexpr bool pair_fits() { return 2 * C::LDS_BYTES <= 160 * 1024; }

// Where kernels pair their transforms: the two exchange buffers fit the LDS
// and each thread holds at most 8 coefficients of each polynomial.  With 16
// or 32 the two register sets crowd out the rest of a kernel's state and
// the compiler spills.
template <class C>
constexpr bool pair_pays() { return pair_fits<C>() && C::E <= 8; }

template <class C, int P, bool PRE = false>
__device__ __forceinline__ void fwd_pass2(uint64_t (&x)[C::E], uint64_t (&y)[C::E],
                                          const ulonglong2* __restrict__ tw, uint64_t q, uint64_t two_q,
                                          int tid, const FwdFirst<C, P>& pre = FwdFirst<C, P>{}) {
    const uint64_t nq = (uint64_t)0 - q, four_q = two_q << 1;
    constexpr int RP = C::rsize(P), LO = C::lo(P), HI = C::hi(P);
    constexpr int M = 1 << RP, G = C::E >> RP;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int fh = group_of<C, P>(g, tid);
#pragma unroll
        for (int b = HI; b >= LO; --b) {
            const int h = 1 << (b - LO);
            const int base = 1 << (C::LOGN - 1 - b);
#pragma unroll
            for (int e = 0; e < M; ++e) {
                if (e & h) continue;
                const int i = (fh << (HI - b)) | (e >> (b - LO + 1));
                const ulonglong2 w = (PRE && b == HI) ? pre.w[g] : tw[base + i];
                ct_bfly_lazy(x[g * M + e], x[g * M + e + h], w.x, w.y, nq, four_q);
                ct_bfly_lazy(y[g * M + e], y[g * M + e + h], w.x, w.y, nq, four_q);
            }
        }
    }
}

template <class C, int P, bool PRE = false>
__device__ __forceinline__ void inv_pass2(uint64_t (&x)[C::E], uint64_t (&y)[C::E],
                                          const ulonglong2* __restrict__ itw, const PrimeConst& pc, int tid,
                                          const InvFirst<C, P>& pre = InvFirst<C, P>{}) {
    constexpr int RP = C::rsize(P), LO = C::lo(P), HI = C::hi(P);
    constexpr int M = 1 << RP, G = C::E >> RP;
    const uint64_t q = pc.q, nq = (uint64_t)0 - q;
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const int fh = group_of<C, P>(g, tid);
#pragma unroll
        for (int b = LO; b <= HI; ++b) {
            const int h = 1 << (b - LO);
            const int base = 1 << (C::LOGN - 1 - b);
            const uint64_t bound = q << (b + 1);
#pragma unroll
            for (int e = 0; e < M; ++e) {
                if (e & h) continue;
                uint64_t& X = x[g * M + e];
                uint64_t& Y = x[g * M + e + h];
                uint64_t& U = y[g * M + e];
                uint64_t& V = y[g * M + e + h];
                if (b == C::LOGN - 1) {
                    const uint64_t x0 = X, x1 = Y, u0 = U, u1 = V;
                    X = mul_shoup_nq(x0 + x1, pc.ninv, pc.ninv_s, nq);
                    Y = mul_shoup_nq(x0 + bound - x1, pc.ninv_w1, pc.ninv_w1_s, nq);
                    U = mul_shoup_nq(u0 + u1, pc.ninv, pc.ninv_s, nq);
                    V = mul_shoup_nq(u0 + bound - u1, pc.ninv_w1, pc.ninv_w1_s, nq);
                } else {
                    const int i = (fh << (HI - b)) | (e >> (b - LO + 1));
                    const ulonglong2 w = (PRE && b == LO) ? pre.w[g * (M / 2) + e / 2] : itw[base + i];
                    gs_bfly_lazy(X, Y, w.x, w.y, nq, bound);
                    gs_bfly_lazy(U, V, w.x, w.y, nq, bound);
                }
            }
        }
    }
}

// One exchange of both register sets from pass A's layout to pass B's.
template <class C, int A, int B, bool W>
__device__ __forceinline__ void exchange2(uint64_t (&x)[C::E], uint64_t (&y)[C::E], uint64_t* lds, int tid) {
    constexpr int K = xshift<C>(A > B ? A : B);
    to_lds<C, A, K>(x, lds, tid);
    to_lds<C, A, K>(y, lds + C::LDS_WORDS, tid);
    exchange_sync<W>();
    from_lds<C, B, K>(x, lds, tid);
    from_lds<C, B, K>(y, lds + C::LDS_WORDS, tid);
    exchange_sync<W>();
}

template <class C, int P>
__device__ __forceinline__ void fwd_rest2(uint64_t (&x)[C::E], uint64_t (&y)[C::E], uint64_t* lds,
                                          const ulonglong2* __restrict__ tw, uint64_t q, uint64_t two_q,
                                          int tid) {
    if constexpr (P < C::NP) {
        const FwdFirst<C, P> pre = fwd_first<C, P>(tw, tid);
        if constexpr (P == 1) __syncthreads();
        else if constexpr (relaid_after_wave_local<C, P - 1, P>()) __syncthreads();
        exchange2<C, P - 1, P, wave_local<C, P>()>(x, y, lds, tid);
        fwd_pass2<C, P, true>(x, y, tw, q, two_q, tid, pre);
        fwd_rest2<C, P + 1>(x, y, lds, tw, q, two_q, tid);
    }
}

// fwd_ntt on x and y at once (same entry and exit conditions).
template <class C>
__device__ __forceinline__ void fwd_ntt2(uint64_t (&x)[C::E], uint64_t (&y)[C::E], uint64_t* lds,
                                         const ulonglong2* __restrict__ tw, uint64_t q, uint64_t two_q,
                                         int tid) {
    static_assert(pair_fits<C>(), "two exchange buffers exceed the LDS");
    opaque(tw, tid);
    fwd_pass2<C, 0>(x, y, tw, q, two_q, tid);
    fwd_rest2<C, 1>(x, y, lds, tw, q, two_q, tid);
}

template <class C, int P>
__device__ __forceinline__ void inv_rest2(uint64_t (&x)[C::E], uint64_t (&y)[C::E], uint64_t* lds,
                                          const ulonglong2* __restrict__ itw, const PrimeConst& pc, int tid) {
    if constexpr (P >= 0) {
        const InvFirst<C, P> pre = inv_first<C, P>(itw, tid);
        if constexpr (relaid_after_wave_local<C, P + 2, P + 1>()) __syncthreads();
        exchange2<C, P + 1, P, wave_local<C, P + 1>()>(x, y, lds, tid);
        inv_pass2<C, P, true>(x, y, itw, pc, tid, pre);
        inv_rest2<C, P - 1>(x, y, lds, itw, pc, tid);
    }
}

// inv_ntt on x and y at once (same entry and exit conditions).
template <class C>
__device__ __forceinline__ void inv_ntt2(uint64_t (&x)[C::E], uint64_t (&y)[C::E], uint64_t* lds,
                                         const ulonglong2* __restrict__ itw, const PrimeConst& pc, int tid) {
    static_assert(pair_fits<C>(), "two exchange buffers exceed the LDS");
    opaque(itw, tid);
    inv_pass2<C, C::NP - 1>(x, y, itw, pc, tid);
    inv_rest2<C, C::NP - 2>(x, y, lds, itw, pc, tid);
}

// HBM <-> registers in pass-0 (coalesced) layout.
template <class C>
__device__ __forceinline__ void load_coalesced(uint64_t (&x)[C::E], const uint64_t* __restrict__ src,
                                               int tid) {
#pragma unroll
    for (int e = 0; e < C::E; ++e) x[e] = src[e * C::T + tid];
}

template <class C>
__device__ __forceinline__ void store_coalesced(const uint64_t (&x)[C::E], uint64_t* __restrict__ dst,
                                                int tid) {
#pragma unroll
    for (int e = 0; e < C::E; ++e) dst[e * C::T + tid] = x[e];
}

// Registers in the last forward pass's layout <-> HBM in bit-reversed
// evaluation order, through LDS so that the HBM side is coalesced.
template <class C>
__device__ __forceinline__ void store_eval(const uint64_t (&x)[C::E], uint64_t* lds,
                                           uint64_t* __restrict__ dst, int tid) {
    to_lds<C, C::NP - 1>(x, lds, tid);
    __syncthreads();
#pragma unroll
    for (int e = 0; e < C::E; ++e) dst[e * C::T + tid] = lds[pad<C>(e * C::T + tid)];
    __syncthreads();
}

template <class C>
__device__ __forceinline__ void load_eval(uint64_t (&x)[C::E], uint64_t* lds,
                                          const uint64_t* __restrict__ src, int tid) {
#pragma unroll
    for (int e = 0; e < C::E; ++e) lds[pad<C>(e * C::T + tid)] = src[e * C::T + tid];
    __syncthreads();
    from_lds<C, C::NP - 1>(x, lds, tid);
    __syncthreads();
}

// Registers in pass-0 (coalesced) layout -> last forward pass's layout.
template <class C>
__device__ __forceinline__ void coalesced_to_last(uint64_t (&x)[C::E], uint64_t* lds, int tid) {
#pragma unroll
    for (int e = 0; e < C::E; ++e) lds[pad<C>(e * C::T + tid)] = x[e];
    __syncthreads();
    from_lds<C, C::NP - 1>(x, lds, tid);
    __syncthreads();
}

// Evaluation-order index of register k in the last forward pass's layout.
template <class C>
__device__ __forceinline__ int eval_index(int k, int tid) {
    constexpr int M = 1 << C::rsize(C::NP - 1);
    return index_of<C, C::NP - 1>(k / M, k % M, tid);
}

// The wave-local exchanges of the built configurations (an exhaustive host
// enumeration gives the same table).
static_assert(!wave_local<Cfg<13, 512>, 1>() && wave_local<Cfg<13, 512>, 2>() && wave_local<Cfg<13, 512>, 3>());
static_assert(!wave_local<Cfg<13, 1024>, 3>() && wave_local<Cfg<13, 1024>, 4>());
static_assert(!wave_local<Cfg<13, 256>, 1>() && wave_local<Cfg<13, 256>, 2>());
static_assert(!wave_local<Cfg<14, 1024>, 2>() && wave_local<Cfg<14, 1024>, 3>());
static_assert(!wave_local<Cfg<14, 512>, 1>() && !wave_local<Cfg<14, 512>, 2>());
static_assert(!wave_local<Cfg<12, 128>, 2>() && wave_local<Cfg<12, 512>, 3>());

}  // namespace ntt
