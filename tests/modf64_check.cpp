// Host build of the FP64 residue arithmetic (cipher-pu_amd/csrc/modf64.h)
// checked against exact 128-bit integer arithmetic, on each routine's stated
// input domain, for random and extreme operands: every result must be an
// integer, congruent to the exact value modulo q and inside the documented
// bound.  Moduli: the chain primes 2^44 + 32768 j + 1, the special primes,
// the single-prime product's prime just below 2^45, the largest odd number
// below 2^45 and small moduli.  Prints one line per failure (at most a few)
// and a summary; exit status 0 when everything holds.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "modf64.h"

typedef __int128 i128;

static uint64_t rng_state = 0x2545F4914F6CDD1Dull;
static uint64_t next64() {   // splitmix64
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t below(uint64_t n) { return n ? (uint64_t)(((unsigned __int128)next64() * n) >> 64) : 0; }

static long failures = 0, checks = 0;
#define CHECK(cond, ...)                                        \
    do {                                                        \
        ++checks;                                               \
        if (!(cond)) {                                          \
            if (failures++ < 8) {                               \
                std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
                std::printf(__VA_ARGS__);                       \
                std::printf("\n");                              \
            }                                                   \
        }                                                       \
    } while (0)

// a signed integer with |x| <= lim: random, or an edge of the range
static int64_t signed_operand(int64_t lim, int k) {
    switch (k % 8) {
        case 0: return lim;
        case 1: return -lim;
        case 2: return 0;
        case 3: return lim - (int64_t)below(4);
        case 4: return -lim + (int64_t)below(4);
        default: {
            const int64_t m = (int64_t)below((uint64_t)lim + 1);
            return (next64() & 1) ? m : -m;
        }
    }
}

static i128 mod(i128 a, int64_t q) {
    i128 r = a % q;
    return r < 0 ? r + q : r;
}

static bool is_int(double x) { return std::floor(x) == x && std::fabs(x) < 0x1p53; }

int main(int argc, char** argv) {
    const int iters = argc > 1 ? std::atoi(argv[1]) : 100000;
    const int64_t moduli[] = {17592186634241, 17592186372097, 17592186273793, 17592186175489,   // chain primes
                              17592186765313, 17592187617281,                                    // special primes
                              35184371884033,                                                     // the 45-bit prime
                              35184372088831,                                                     // 2^45 - 1
                              113, 12289, 1073750017};
    for (int64_t q : moduli) {
        const double dq = (double)q, nq = -dq, qinv = 1.0 / dq;
        for (int it = 0; it < iters; ++it) {
            // mulw: |a| < 2^51, w in [0, q)
            {
                const int64_t a = signed_operand((1ll << 51) - 1, it);
                const int64_t w = (it % 5 == 0) ? q - 1 : (it % 5 == 1 ? 0 : (int64_t)below(q));
                const double wq = (double)w / dq;
                const double r = mulw((double)a, (double)w, wq, nq);
                CHECK(is_int(r), "mulw not an integer q=%lld a=%lld w=%lld", (long long)q, (long long)a, (long long)w);
                const i128 ri = (i128)r;
                CHECK(mod(ri - (i128)a * w, q) == 0, "mulw residue q=%lld a=%lld w=%lld r=%.0f", (long long)q,
                      (long long)a, (long long)w, r);
                CHECK(8 * (ri < 0 ? -ri : ri) <= 5 * (i128)q, "mulw bound q=%lld a=%lld w=%lld r=%.0f", (long long)q,
                      (long long)a, (long long)w, r);
            }
            // mulk: |a| < 2^51, w in [0, q), quotient formed from qinv
            {
                const int64_t a = signed_operand((1ll << 51) - 1, it + 1);
                const int64_t w = (it % 5 == 0) ? q - 1 : (it % 5 == 1 ? 1 : (int64_t)below(q));
                const double r = mulk((double)a, (double)w, qinv, nq);
                CHECK(is_int(r), "mulk not an integer q=%lld a=%lld w=%lld", (long long)q, (long long)a, (long long)w);
                const i128 ri = (i128)r;
                CHECK(mod(ri - (i128)a * w, q) == 0, "mulk residue q=%lld a=%lld w=%lld r=%.0f", (long long)q,
                      (long long)a, (long long)w, r);
                CHECK(8 * (ri < 0 ? -ri : ri) <= 7 * (i128)q, "mulk bound q=%lld a=%lld w=%lld r=%.0f", (long long)q,
                      (long long)a, (long long)w, r);
            }
            // muld: |a b| <= 2^50 q (operands up to the bound's square root and reduced ones)
            {
                int64_t a, b;
                if (it & 1) {
                    a = signed_operand((q + 1) / 2, it);
                    b = signed_operand((q + 1) / 2, it / 2 + 3);
                } else {
                    const double lim = std::sqrt(0x1p50 * dq);
                    a = signed_operand((int64_t)lim, it);
                    b = signed_operand((int64_t)lim, it / 2 + 5);
                }
                const double r = muld((double)a, (double)b, qinv, nq);
                CHECK(is_int(r), "muld not an integer q=%lld a=%lld b=%lld", (long long)q, (long long)a, (long long)b);
                const i128 ri = (i128)r;
                CHECK(mod(ri - (i128)a * b, q) == 0, "muld residue q=%lld a=%lld b=%lld", (long long)q, (long long)a,
                      (long long)b);
                CHECK(4 * (ri < 0 ? -ri : ri) <= 3 * (i128)q, "muld bound q=%lld a=%lld b=%lld r=%.0f", (long long)q,
                      (long long)a, (long long)b, r);
            }
            // rmod: |x| < 2^52
            {
                const int64_t x = signed_operand((1ll << 52) - 1, it);
                const double r = rmod((double)x, qinv, nq);
                CHECK(is_int(r), "rmod not an integer");
                const i128 ri = (i128)r;
                CHECK(mod(ri - x, q) == 0, "rmod residue q=%lld x=%lld", (long long)q, (long long)x);
                CHECK(2 * (ri < 0 ? -ri : ri) <= (i128)q + 1, "rmod bound q=%lld x=%lld r=%.0f", (long long)q,
                      (long long)x, r);
            }
            // canon: |x| <= 16 q, exact canonical residue; multiples of q and
            // their neighbours included
            {
                int64_t x;
                if (it % 4 == 0) x = (int64_t)(below(33)) * q - 16 * q + (int64_t)below(3) - 1;
                else x = signed_operand(16 * q, it);
                if (x > 16 * q) x = 16 * q;
                if (x < -16 * q) x = -16 * q;
                const double r = canon((double)x, qinv, nq);
                CHECK(r == (double)(int64_t)mod(x, q), "canon q=%lld x=%lld r=%.0f", (long long)q, (long long)x, r);
            }
            // u2d / d2u round trip below 2^52
            {
                const uint64_t u = (it % 3 == 0) ? (1ull << 52) - 1 - below(2) : below(1ull << 52);
                const double d = u2d(u);
                CHECK(d == (double)u && d2u(d) == u, "u2d/d2u %llu", (unsigned long long)u);
            }
            // butterflies on their bounds: forward from bound B (B q + 5q/8 stays
            // exact), inverse from 2 B q < 2^51
            {
                const int64_t w = (int64_t)below(q);
                const double wq = (double)w / dq;
                const int64_t lim = std::min<int64_t>(60 * q, (1ll << 51) - 1);
                const int64_t X = signed_operand(lim, it), Y = signed_operand(lim, it + 1);
                double x = (double)X, y = (double)Y;
                ct_bfly_f(x, y, (double)w, wq, nq);
                CHECK(is_int(x) && is_int(y) && mod((i128)x - X - (i128)w * Y, q) == 0 &&
                          mod((i128)y - X + (i128)w * Y, q) == 0,
                      "ct_bfly_f q=%lld", (long long)q);
                const int64_t lim2 = std::min<int64_t>(31 * q, (1ll << 50) - 1);
                const int64_t U = signed_operand(lim2, it + 2), V = signed_operand(lim2, it + 3);
                double u = (double)U, v = (double)V;
                gs_bfly_f(u, v, (double)w, wq, nq);
                CHECK(is_int(u) && is_int(v) && (i128)u == (i128)U + V && mod((i128)v - (i128)(U - V) * w, q) == 0,
                      "gs_bfly_f q=%lld", (long long)q);
            }
        }
    }
    std::printf("%ld checks, %ld failures\n", checks, failures);
    return failures ? 1 : 0;
}
