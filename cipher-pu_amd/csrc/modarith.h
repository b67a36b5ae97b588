// Device-side modular arithmetic for primes q < 2^45 (gfx950).
//
// Replaces the RTL multiplier/reducer pair of the reference
// (rtl/m1/mult_43x43_horner4.sv:31-121, rtl/m2/mod45_fastred.sv:61-122,
// rtl/m2/mod45_barrett.sv:49-75).  The RTL reduction needs the special form
// q = 2^44 + c; these routines work for any prime below 2^45.
//
//  * mul_shoup: a * w mod q for a fixed operand w with the companion
//    w' = floor(w 2^63 / q) precomputed (twiddles, key material, constants).
//    Result in [0, 2q) for any a < 2^62.
//  * mul_barrett: a * b mod q for two data operands (a, b < q), result in
//    [0, q).  Barrett with mu = floor(2^(2l) / q), l = bit length of q.
//  * the forward (Cooley-Tukey) and inverse (Gentleman-Sande) butterflies
//    keep values lazily in [0, 4q) / [0, 2q) (Harvey's bounds).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Every routine here is plain integer C++, compiled for the device and the
// host alike: the host build lets the CPU test suite check the exact
// product code against 128-bit arithmetic (tests/modarith_check.cpp).
#define MODARITH_FN __host__ __device__ __forceinline__

// High 64 bits of a 64 x 64-bit product.
MODARITH_FN uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

#define CKKS_MAX_PRIMES 16

// Per-prime constants, resident in device memory (one entry per prime).
struct PrimeConst {
    uint64_t q;
    uint64_t two_q;
    uint64_t mu;        // floor(2^(2l) / q)
    uint64_t mu64;      // floor(2^64 / q) for reducing arbitrary 64-bit values
    uint32_t ell;       // bit length of q
    uint32_t pad;
    uint64_t ninv, ninv_s;        // N^-1 and its Shoup companion
    uint64_t ninv_w1, ninv_w1_s;  // N^-1 * (inverse root of the last GS stage)
};

// Up to CKKS_MAX_PRIMES prime indices, passed by value to kernels: limb i of
// a tensor uses prime idx[i].
struct LimbMap {
    int idx[CKKS_MAX_PRIMES];
};


// floor(a * ws / 2^63) for a < 2^62 and ws < 2^63.  With ws below 2^63 the
// middle partial sum a0*ws1 + (a1*ws0 + hi(a0*ws0)) stays below 2^64, so
// the 64-bit multiply-adds chain without a carry word.
MODARITH_FN uint64_t mulhi63(uint64_t a, uint64_t ws) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint32_t b0 = (uint32_t)ws, b1 = (uint32_t)(ws >> 32);
    const uint64_t f = (uint64_t)a0 * b0;
    const uint64_t m = (uint64_t)a1 * b0 + (f >> 32);
    const uint64_t n = (uint64_t)a0 * b1 + m;
    const uint64_t h = (uint64_t)a1 * b1 + (n >> 32);
    return (h << 1) | ((uint32_t)n >> 31);
}

// Shoup: a w - floor(a ws / 2^63) q.  The quotient estimate is low by at most
// one (a w/q - a ws/2^63 < a / 2^63 < 1), so the result lies in [0, 2q).
MODARITH_FN uint64_t mul_shoup(uint64_t a, uint64_t w, uint64_t ws, uint64_t q) {
    return a * w - mulhi63(a, ws) * q;
}

MODARITH_FN uint64_t csub(uint64_t a, uint64_t m) { return a >= m ? a - m : a; }

// a * b mod q for a special-form prime q = 2^44 + c, 0 < c < 2^21, a, b < q,
// using 2^44 = -c (mod q) -- the reference's fast reduction
// (scripts/_verify_fastred.py:55-99, rtl mod45_fastred.sv) for c < 2^18,
// regrouped so that every partial product fits 64 bits for c < 2^21:
//   x = a b = xh 2^44 + xl,  xh = xh1 2^23 + xh0 (xh0, xh1 < 2^23)
//   x = xl - c xh0 - c xh1 2^23,  and with u = c xh1 < 2^44:
//   u 2^23 = (u >> 21) 2^44 + (u mod 2^21) 2^23 = (u mod 2^21) 2^23 - c (u >> 21)
// so x = [xl + c (u >> 21)] - [c xh0 + (u mod 2^21) 2^23], both brackets
// below 2^45 < 2q; adding 2q leaves a value in (0, 4q).
MODARITH_FN uint64_t mul_special(uint64_t a, uint64_t b, uint64_t q) {
    const uint64_t c = q - (1ull << 44);
    const uint64_t lo = a * b, hi = mulhi64(a, b);
    const uint64_t xl = lo & ((1ull << 44) - 1);
    const uint64_t xh = (hi << 20) | (lo >> 44);
    const uint64_t xh0 = xh & ((1ull << 23) - 1), xh1 = xh >> 23;
    const uint64_t u = c * xh1;
    const uint64_t pos = xl + c * (u >> 21);
    const uint64_t neg = c * xh0 + ((u & ((1ull << 21) - 1)) << 23);
    const uint64_t r = pos + 2 * q - neg;
    return csub(csub(r, 2 * q), q);
}

MODARITH_FN uint64_t add_mod(uint64_t a, uint64_t b, uint64_t q) { return csub(a + b, q); }

MODARITH_FN uint64_t sub_mod(uint64_t a, uint64_t b, uint64_t q) { return a >= b ? a - b : a + q - b; }

// a * b mod q for a, b < q < 2^45 (HAC 14.42 with base 2).
MODARITH_FN uint64_t mul_barrett(uint64_t a, uint64_t b, const PrimeConst& P) {
    const uint64_t lo = a * b;
    const uint64_t hi = mulhi64(a, b);
    const uint32_t l = P.ell;
    const uint64_t q1 = (hi << (65 - l)) | (lo >> (l - 1));   // floor(ab / 2^(l-1)) < 2^(l+1)
    const uint64_t q3 = mulhi64(q1 << (63 - l), P.mu);    // floor(q1 mu / 2^(l+1))
    uint64_t r = lo - q3 * P.q;                               // in [0, 3q)
    r = csub(r, P.two_q);
    return csub(r, P.q);
}

// x mod q for any 64-bit x: floor(x/q) is estimated from mu64 = floor(2^64/q)
// to within 1, so one conditional subtraction finishes.
MODARITH_FN uint64_t reduce64(uint64_t x, const PrimeConst& P) {
    const uint64_t t = mulhi64(x, P.mu64);
    return csub(x - t * P.q, P.q);
}

// Forward CT butterfly (natural -> bit-reversed order), Harvey's bounds:
// inputs/outputs in [0, 4q).  X' = X + W Y,  Y' = X - W Y.
MODARITH_FN void ct_bfly(uint64_t& X, uint64_t& Y, uint64_t w, uint64_t ws,
                                        uint64_t q, uint64_t two_q) {
    const uint64_t x = csub(X, two_q);
    const uint64_t t = mul_shoup(Y, w, ws, q);
    X = x + t;
    Y = x - t + two_q;
}

// Inverse GS butterfly, inputs/outputs in [0, 2q):  X' = X + Y,
// Y' = (X - Y) W.
MODARITH_FN void gs_bfly(uint64_t& X, uint64_t& Y, uint64_t w, uint64_t ws,
                                        uint64_t q, uint64_t two_q) {
    const uint64_t x = X, y = Y;
    X = csub(x + y, two_q);
    Y = mul_shoup(x - y + two_q, w, ws, q);
}

// Shoup product with the subtraction written as a multiply-add:
// a*w + qt*(2^64 - q) = a*w - qt*q (mod 2^64).  nq = 2^64 - q.
MODARITH_FN uint64_t mul_shoup_nq(uint64_t a, uint64_t w, uint64_t ws, uint64_t nq) {
    return a * w + mulhi63(a, ws) * nq;
}

// floor(a ws / 2^63) minus 0, 1 or 2, for a, ws < 2^63: the low-word product
// a0 w0 (< 2^64, i.e. < 2 at the 2^63 scale) and the fraction of the middle
// sum are dropped.  m = a1 w0 + a0 w1 < 2^63 + 2^63 fits 64 bits.
MODARITH_FN uint64_t mulhi63_lazy(uint64_t a, uint64_t ws) {
    const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
    const uint32_t w0 = (uint32_t)ws, w1 = (uint32_t)(ws >> 32);
    const uint64_t m = (uint64_t)a1 * w0 + (uint64_t)a0 * w1;
    return (((uint64_t)a1 * w1) << 1) + (m >> 31);
}

// a w mod q up to 3q, result in [0, 4q), for a < 2^63: Shoup's quotient is
// at most 1 below floor(a w / q) and the lazy high half at most 2 below that.
MODARITH_FN uint64_t mul_shoup4_nq(uint64_t a, uint64_t w, uint64_t ws, uint64_t nq) {
    return a * w + mulhi63_lazy(a, ws) * nq;
}

// Unreduced butterflies for the LDS-resident transforms; no per-stage
// reduction is needed:
//  * forward (product in [0, 4q)):  X, Y < B  ->  X' = X + WY < B + 4q,
//    Y' = X + 4q - WY < B + 4q.  From inputs below 2q, after s stages every
//    value is below (2 + 4s) q (< 2^51 for log2 N <= 14 and q < 2^45).
//  * inverse:  X, Y < B (B a multiple of q)  ->  X' = X + Y < 2B,
//    Y' = (X - Y + B) W < 4q <= 2B.  From inputs below 2q the bound at stage
//    s is 2^(s+1) q (< 2^60 for log2 N <= 14); the last stage, which folds
//    N^-1, uses the exact product so the transform ends in [0, 2q).
MODARITH_FN void ct_bfly_lazy(uint64_t& X, uint64_t& Y, uint64_t w, uint64_t ws, uint64_t nq,
                                             uint64_t four_q) {
    const uint64_t t = mul_shoup4_nq(Y, w, ws, nq);
    const uint64_t x = X;
    X = x + t;
    Y = x + four_q - t;
}

MODARITH_FN void gs_bfly_lazy(uint64_t& X, uint64_t& Y, uint64_t w, uint64_t ws, uint64_t nq,
                                             uint64_t bound) {
    const uint64_t x = X, y = Y;
    X = x + y;
    Y = mul_shoup4_nq(x + bound - y, w, ws, nq);
}
