// Host build of the device modular arithmetic (cipher-pu_amd/csrc/modarith.h)
// checked against exact 128-bit arithmetic, on each routine's stated input
// domain, for random and extreme operands and every prime class the library
// meets: the chain primes 2^44 + 32768 j + 1, the special primes, and the
// single-prime product's prime just below 2^45.  Prints one line per failure (at most a
// few) and a summary; exit status 0 when everything holds.
#include <cstdio>
#include <cstdint>
#include <cstdlib>

#include "modarith.h"

typedef unsigned __int128 u128;

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next64() {   // splitmix64
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint64_t below(uint64_t n) { return n ? (uint64_t)(((u128)next64() * n) >> 64) : 0; }

static long failures = 0, checks = 0;
#define CHECK(cond, ...)                                      \
    do {                                                      \
        ++checks;                                             \
        if (!(cond)) {                                        \
            if (failures++ < 8) {                             \
                std::printf("FAIL %s:%d ", __FILE__, __LINE__); \
                std::printf(__VA_ARGS__);                     \
                std::printf("\n");                            \
            }                                                 \
        }                                                     \
    } while (0)

static int bitlen(uint64_t x) {
    int n = 0;
    while (x) {
        ++n;
        x >>= 1;
    }
    return n;
}

static PrimeConst make_pc(uint64_t q) {
    PrimeConst pc{};
    pc.q = q;
    pc.two_q = 2 * q;
    pc.ell = bitlen(q);
    pc.mu = (uint64_t)(((u128)1 << (2 * pc.ell)) / q);
    pc.mu64 = (uint64_t)(~0ull / q);
    return pc;
}

static uint64_t shoup63(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 63) / q); }
static uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }

// an operand below `limit`: random, or one of the edges of that range
static uint64_t operand(uint64_t limit, uint64_t q, int k) {
    switch (k % 16) {
        case 0: return 0;
        case 1: return limit - 1;
        case 2: return q - 1 < limit ? q - 1 : limit - 1;
        case 3: return q < limit ? q : limit - 1;
        case 4: return (limit - 1) / q * q;          // a multiple of q
        default: return below(limit);
    }
}

static void check_prime(uint64_t q, int n) {
    const PrimeConst pc = make_pc(q);
    const uint64_t nq = (uint64_t)0 - q;
    const uint64_t two63 = 1ull << 63, two62 = 1ull << 62;
    for (int k = 0; k < n; ++k) {
        const uint64_t w = (k % 7 == 0) ? q - 1 : (k % 7 == 1 ? 1 : below(q));
        const uint64_t ws = shoup63(w, q);

        // mulhi63: exact high half for a < 2^62
        uint64_t a = operand(two62, q, k);
        CHECK(mulhi63(a, ws) == (uint64_t)(((u128)a * ws) >> 63), "mulhi63 q=%llu a=%llu w=%llu",
              (unsigned long long)q, (unsigned long long)a, (unsigned long long)w);

        // Shoup product, exact quotient: [0, 2q) for a < 2^62
        uint64_t r = mul_shoup_nq(a, w, ws, nq);
        CHECK(r < 2 * q && r % q == mulmod(a % q, w, q), "mul_shoup_nq q=%llu a=%llu w=%llu r=%llu",
              (unsigned long long)q, (unsigned long long)a, (unsigned long long)w, (unsigned long long)r);
        CHECK(mul_shoup(a, w, ws, q) == r, "mul_shoup != mul_shoup_nq");

        // lazy high half and the lazy product: [0, 4q) for a < 2^63
        a = operand(two63, q, k);
        const uint64_t hi = (uint64_t)(((u128)a * ws) >> 63), lz = mulhi63_lazy(a, ws);
        CHECK(lz <= hi && hi - lz <= 2, "mulhi63_lazy q=%llu a=%llu w=%llu", (unsigned long long)q,
              (unsigned long long)a, (unsigned long long)w);
        r = mul_shoup4_nq(a, w, ws, nq);
        CHECK(r < 4 * q && r % q == mulmod(a % q, w, q), "mul_shoup4_nq q=%llu a=%llu w=%llu r=%llu",
              (unsigned long long)q, (unsigned long long)a, (unsigned long long)w, (unsigned long long)r);

        // Barrett, data x data, a, b < q
        const uint64_t x = operand(q, q, k), y = operand(q, q, k + 3);
        CHECK(mul_barrett(x, y, pc) == mulmod(x, y, q), "mul_barrett q=%llu a=%llu b=%llu", (unsigned long long)q,
              (unsigned long long)x, (unsigned long long)y);

        // any 64-bit value
        const uint64_t z = (k % 5 == 0) ? ~0ull - (uint64_t)k : next64();
        CHECK(reduce64(z, pc) == z % q, "reduce64 q=%llu x=%llu", (unsigned long long)q, (unsigned long long)z);

        // lazy butterflies on their documented bounds (forward: inputs below
        // (2 + 4 s) q after s stages, s <= 13; inverse: below 2^(s+1) q)
        const int s = k % 14;
        const uint64_t B = (2 + 4 * (uint64_t)s) * q;
        uint64_t X = below(B), Y = below(B);
        const uint64_t X0 = X, Y0 = Y;
        ct_bfly_lazy(X, Y, w, ws, nq, 4 * q);
        const uint64_t wy = mulmod(Y0 % q, w, q);
        CHECK(X < B + 4 * q && Y < B + 4 * q && X % q == (X0 % q + wy) % q &&
                  Y % q == (X0 % q + q - wy) % q,
              "ct_bfly_lazy q=%llu s=%d", (unsigned long long)q, s);
        const uint64_t bound = q << (s + 1);
        X = below(bound);
        Y = below(bound);
        const uint64_t U0 = X, V0 = Y;
        gs_bfly_lazy(X, Y, w, ws, nq, bound);
        CHECK(X < 2 * bound && Y < 4 * q && X % q == (U0 + V0) % q &&
                  Y % q == mulmod((U0 % q + q - V0 % q) % q, w, q),
              "gs_bfly_lazy q=%llu s=%d", (unsigned long long)q, s);
    }
}

static void check_special(uint64_t q, int n) {
    for (int k = 0; k < n; ++k) {
        const uint64_t a = operand(q, q, k), b = operand(q, q, k + 5);
        const uint64_t r = mul_special(a, b, q);
        CHECK(r == mulmod(a, b, q), "mul_special q=%llu a=%llu b=%llu r=%llu", (unsigned long long)q,
              (unsigned long long)a, (unsigned long long)b, (unsigned long long)r);
    }
}

int main(int argc, char** argv) {
    const int n = argc > 1 ? std::atoi(argv[1]) : 200000;
    const uint64_t primes[] = {
        17592186634241ull, 17592186372097ull, 17592186273793ull, 17592186175489ull,   // chain4
        17592186765313ull, 17592187617281ull,                                         // special, chain5
        35184371884033ull};                                                           // largest 45-bit prime = 1 mod 8192
    for (uint64_t q : primes) check_prime(q, n);
    // special form q = 2^44 + c: chain primes and the largest c the routine allows
    for (uint64_t q : {17592186634241ull, 17592186175489ull, 17592187617281ull, (1ull << 44) + (1ull << 21) - 1})
        check_special(q, n);
    std::printf("%ld checks, %ld failures\n", checks, failures);
    return failures ? 1 : 0;
}
