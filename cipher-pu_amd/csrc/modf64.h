// Residue arithmetic on the FP64 pipeline for primes q < 2^45 (gfx950).
//
// A residue modulo q is held as a double whose value is an integer
// congruent to it.  Magnitudes stay below 2^52, so sums and differences are
// exact, and every product below is an exact integer computation assembled
// from fused multiply-adds: no rounding error reaches a result.  On gfx950 a
// wave64 FP64 FMA issues at the full VALU rate, and a product modulo q costs
// six FP64 instructions against about twenty 32-bit integer instructions for
// the 64-bit Shoup product of modarith.h.
//
// Replaces, like modarith.h, the RTL multiplier/reducer pair of the reference
// (rtl/m1/mult_43x43_horner4.sv:31-121, rtl/m2/mod45_fastred.sv:61-122); any
// prime below 2^45 works.
//
// Bounds (q < 2^45 throughout; "|x| <= B q" is written "bound B"):
//  * mulw(a, w):  |a| < 2^51                  -> |r| <= 5q/8
//  * mulk(a, w):  |a| < 2^51                  -> |r| <= 7q/8
//  * muld(a, b):  |a b| <= 2^50 q             -> |r| <= 3q/4
//  * rmod(x):     |x| < 2^52                  -> |r| <= (q + 1)/2
//  * canon(x):    |x| <= 16 q                 -> r in [0, q), r = x mod q
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define MODF_FN __host__ __device__ __forceinline__

// Per-prime constants of the FP64 pipeline, resident in device memory.
struct PrimeF {
    double q, nq, qinv;          // q, -q, RN(1/q)
    double ninv, ninv_q;         // N^-1 mod q and RN(ninv / q)
    double ninv_w1, ninv_w1_q;   // N^-1 times the inverse root of the last inverse stage
    double pad;
};

// 1.5 * 2^52: for |t| < 2^51 the sum t + RNDC lies in [2^52, 2^53), where
// doubles are the integers, so fma(a, b, RNDC) - RNDC is a b rounded once to
// the nearest integer.
constexpr double RNDC = 6755399441055744.0;
constexpr double TWO52 = 4503599627370496.0;

// a w mod q for a fixed operand w in [0, q) with wq = RN(w / q) (< 1), and
// any integer a with |a| < 2^51.
//   k = round(a wq): |a wq - a w / q| <= |a| 2^-54 <= 1/8, so |a w / q - k| <= 5/8.
//   h + l = a w exactly (h = RN(a w), l = fma(a, w, -h)); |l| <= 2^43.
//   h - k q = (a w - k q) - l is an integer of magnitude below 2^46: the
//   FMA returns it exactly, and adding l gives r = a w - k q exactly.
MODF_FN double mulw(double a, double w, double wq, double nq) {
#pragma clang fp contract(off)
    const double h = a * w;
    const double l = __builtin_fma(a, w, -h);
    const double k = __builtin_fma(a, wq, RNDC) - RNDC;
    return __builtin_fma(k, nq, h) + l;
}

// a w mod q for a fixed operand w in [0, q) given without its quotient,
// for |a| < 2^51: wq = RN(w qinv) is within (w / q) 2^-53 + 2^-54 < 1.5 2^-53
// of w / q (qinv = RN(1/q) has relative error 2^-53; w qinv < 1 rounds
// within 2^-54), so k = round(a wq) is within 1/2 + 3/8 of a w / q and
// |r| <= 7q/8; r = a w - k q exactly, as in mulw.  Key rows stored as their
// words alone halve the bytes a key product reads, for one more multiply.
MODF_FN double mulk(double a, double w, double qinv, double nq) {
#pragma clang fp contract(off)
    const double wq = w * qinv;
    return mulw(a, w, wq, nq);
}

// a b mod q for data operands with |a b| <= 2^50 q (e.g. both reduced by
// rmod).  k = round(h qinv) with h qinv within |a b / q| 2^-52 <= 1/4 of
// a b / q; the rest as in mulw.
MODF_FN double muld(double a, double b, double qinv, double nq) {
#pragma clang fp contract(off)
    const double h = a * b;
    const double l = __builtin_fma(a, b, -h);
    const double k = __builtin_fma(h, qinv, RNDC) - RNDC;
    return __builtin_fma(k, nq, h) + l;
}

// x mod q into |r| <= (q + 1) / 2 for any integer |x| < 2^52: k = round(x
// qinv) is within 1/2 + 1/(2q) of x / q, and x - k q is exact.
MODF_FN double rmod(double x, double qinv, double nq) {
#pragma clang fp contract(off)
    const double k = __builtin_fma(x, qinv, RNDC) - RNDC;
    return __builtin_fma(k, nq, x);
}

// The canonical residue x mod q in [0, q), for an integer |x| <= 16 q:
// s = RN(x qinv + 2^-47) lies within 2^-48 of x / q + 2^-47 (both the
// product's relative error 2^-53 on |x / q| <= 16 and the rounding of s
// stay below 2^-49), so floor(s) = floor(x / q) -- the fraction of x / q is
// 0 or at least 1/q > 2^-45 -- and x - floor(s) q is exact.
MODF_FN double canon(double x, double qinv, double nq) {
#pragma clang fp contract(off)
    const double k = __builtin_floor(__builtin_fma(x, qinv, 0x1p-47));
    return __builtin_fma(k, nq, x);
}

// uint64 word below 2^52 <-> the double of the same integer value.
MODF_FN double u2d(uint64_t x) {
    return __builtin_bit_cast(double, x | 0x4330000000000000ull) - TWO52;
}

// x: an integer in [0, 2^52).
MODF_FN uint64_t d2u(double x) {
    return __builtin_bit_cast(uint64_t, x + TWO52) & 0x000FFFFFFFFFFFFFull;
}

// Forward (Cooley-Tukey) butterfly on signed residues: X' = X + W Y,
// Y' = X - W Y.  Inputs bounded by B, outputs by B + 7/8 with wq = RN(w
// qinv) (the mulk bound; B + 5/8 with wq = RN(w / q)).
MODF_FN void ct_bfly_f(double& X, double& Y, double w, double wq, double nq) {
#pragma clang fp contract(off)
    const double t = mulw(Y, w, wq, nq);
    const double x = X;
    X = x + t;
    Y = x - t;
}

// Inverse (Gentleman-Sande) butterfly: X' = X + Y, Y' = (X - Y) W.  Inputs
// bounded by B (2 B q < 2^51), X' by 2 B, Y' by 7/8 (wq as above).
MODF_FN void gs_bfly_f(double& X, double& Y, double w, double wq, double nq) {
#pragma clang fp contract(off)
    const double x = X, y = Y;
    X = x + y;
    Y = mulw(x - y, w, wq, nq);
}
