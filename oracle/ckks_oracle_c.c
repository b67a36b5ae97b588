/* CPU oracle (C restatement) -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain C restatement of the reference's golden arithmetic, for sizes where
 * the big-int Python restatement (oracle/ckks_oracle.py) is too slow.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library, and only as the checker; the product never links it.
 *
 * Exact arithmetic: unsigned __int128 products, every value fully reduced.
 * Each function names the reference file:line it follows; the tests pin it
 * against the Python restatement, which is pinned against the reference's
 * known-answer tests and fixtures.
 *
 * Layouts (all uint64, row-major):
 *   polynomial        [N]
 *   RNS polynomial    [l][N]          limb i reduced mod primes[i]
 *   ciphertext        [2][l][N]       (c0, c1), decrypt = c0 + c1*s
 *   hybrid key        [l][2][l+1][N]  digit j: (b_j, a_j) over primes + p,
 *                                     coefficient form
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) {
    return (uint64_t)(((u128)a * b) % q);
}
static inline uint64_t addmod(uint64_t a, uint64_t b, uint64_t q) {
    uint64_t s = a + b;
    return s >= q ? s - q : s;
}
static inline uint64_t submod(uint64_t a, uint64_t b, uint64_t q) {
    return a >= b ? a - b : a + q - b;
}
static uint64_t powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = mulmod(r, b, q);
        b = mulmod(b, b, q);
        e >>= 1;
    }
    return r;
}
static uint64_t invmod(uint64_t a, uint64_t q) { return powmod(a, q - 2, q); } /* q prime */

static int ilog2(uint64_t n) { int l = 0; while ((1ull << l) < n) l++; return l; }

static uint64_t bitrev(uint64_t i, int bits) {
    uint64_t r = 0;
    for (int b = 0; b < bits; b++) { r = (r << 1) | (i & 1); i >>= 1; }
    return r;
}

/* Cyclic DIT transform, natural order out
 * (tb/ref/gen_poly_mult_vectors_n12.py:54-82): bit-reverse the input, then
 * stages of half-size h = 1, 2, ... with root omega^(N/2h) stepped by
 * repeated multiplication. */
void oracle_ntt_cyclic(uint64_t *a, uint64_t N, uint64_t q, uint64_t omega) {
    int bits = ilog2(N);
    for (uint64_t i = 0; i < N; i++) {
        uint64_t j = bitrev(i, bits);
        if (i < j) { uint64_t t = a[i]; a[i] = a[j]; a[j] = t; }
    }
    for (uint64_t h = 1; h < N; h <<= 1) {
        uint64_t step = powmod(omega, N / (2 * h), q);
        for (uint64_t base = 0; base < N; base += 2 * h) {
            uint64_t w = 1;
            for (uint64_t k = 0; k < h; k++) {
                uint64_t u = a[base + k];
                uint64_t v = mulmod(a[base + k + h], w, q);
                a[base + k] = addmod(u, v, q);
                a[base + k + h] = submod(u, v, q);
                w = mulmod(w, step, q);
            }
        }
    }
}

/* gen_poly_mult_vectors_n12.py:85-100 */
void oracle_intt_cyclic(uint64_t *a, uint64_t N, uint64_t q, uint64_t omega_inv) {
    oracle_ntt_cyclic(a, N, q, omega_inv);
    uint64_t ninv = invmod(N % q, q);
    for (uint64_t i = 0; i < N; i++) a[i] = mulmod(a[i], ninv, q);
}

/* Xneg[k] = sum_j a_j psi^((2k+1)j), natural order. */
void oracle_ntt_negacyclic(uint64_t *a, uint64_t N, uint64_t q, uint64_t psi) {
    uint64_t t = 1;
    for (uint64_t j = 0; j < N; j++) { a[j] = mulmod(a[j], t, q); t = mulmod(t, psi, q); }
    oracle_ntt_cyclic(a, N, q, mulmod(psi, psi, q));
}

void oracle_intt_negacyclic(uint64_t *a, uint64_t N, uint64_t q, uint64_t psi) {
    uint64_t pinv = invmod(psi, q);
    oracle_intt_cyclic(a, N, q, mulmod(pinv, pinv, q));
    uint64_t t = 1;
    for (uint64_t j = 0; j < N; j++) { a[j] = mulmod(a[j], t, q); t = mulmod(t, pinv, q); }
}

/* Schoolbook negacyclic product (scripts/verify_keyswitch.py:48-59). */
void oracle_polymul_schoolbook(const uint64_t *a, const uint64_t *b, uint64_t *out,
                               uint64_t N, uint64_t q) {
    for (uint64_t k = 0; k < N; k++) out[k] = 0;
    for (uint64_t i = 0; i < N; i++) {
        if (!a[i]) continue;
        for (uint64_t j = 0; j < N; j++) {
            uint64_t p = mulmod(a[i], b[j], q);
            uint64_t k = i + j;
            if (k < N) out[k] = addmod(out[k], p, q);
            else out[k - N] = submod(out[k - N], p, q);
        }
    }
}

/* Negacyclic product through the transform. */
void oracle_polymul_ntt(const uint64_t *a, const uint64_t *b, uint64_t *out,
                        uint64_t N, uint64_t q, uint64_t psi) {
    uint64_t *ta = malloc(N * 8), *tb = malloc(N * 8);
    memcpy(ta, a, N * 8);
    memcpy(tb, b, N * 8);
    oracle_ntt_negacyclic(ta, N, q, psi);
    oracle_ntt_negacyclic(tb, N, q, psi);
    for (uint64_t k = 0; k < N; k++) out[k] = mulmod(ta[k], tb[k], q);
    oracle_intt_negacyclic(out, N, q, psi);
    free(ta);
    free(tb);
}

/* RNS rescale by the last prime, per-limb form of rescale_coeff on the CRT
 * value (scripts/verify_rescale.py:38-53):
 *   out_i = ((c_i + h) - [(c_L + h) mod q_L]) * q_L^-1 mod q_i, h = (q_L-1)/2 */
void oracle_rescale(const uint64_t *in, uint64_t *out, uint64_t N, int l,
                    const uint64_t *primes) {
    uint64_t qL = primes[l - 1], h = qL / 2;
    const uint64_t *last = in + (uint64_t)(l - 1) * N;
    for (int i = 0; i < l - 1; i++) {
        uint64_t qi = primes[i], inv = invmod(qL % qi, qi), hi = h % qi;
        for (uint64_t k = 0; k < N; k++) {
            uint64_t t = (last[k] + h) % qL;
            uint64_t v = submod(addmod(in[(uint64_t)i * N + k], hi, qi), t % qi, qi);
            out[(uint64_t)i * N + k] = mulmod(v, inv, qi);
        }
    }
}

/* ModDown by the special prime p (round-half-up of X/p), per limb. The
 * input has l+1 limbs; limb l is the p limb. */
void oracle_mod_down(const uint64_t *in, uint64_t *out, uint64_t N, int l,
                     const uint64_t *primes, uint64_t p) {
    uint64_t h = p / 2;
    const uint64_t *last = in + (uint64_t)l * N;
    for (int i = 0; i < l; i++) {
        uint64_t qi = primes[i], inv = invmod(p % qi, qi), hi = h % qi;
        for (uint64_t k = 0; k < N; k++) {
            uint64_t t = (last[k] + h) % p;
            uint64_t v = submod(addmod(in[(uint64_t)i * N + k], hi, qi), t % qi, qi);
            out[(uint64_t)i * N + k] = mulmod(v, inv, qi);
        }
    }
}

/* Hybrid key switch of d2 (coefficient form, l limbs) with one digit per
 * prime.  out0/out1: [l][N] = ModDown(sum_j D_j b_j), ModDown(-sum_j D_j a_j).
 * psis[i] for i < l are the primes' roots, psis[l] is p's root. */
void oracle_keyswitch(const uint64_t *d2, const uint64_t *key, uint64_t *out0,
                      uint64_t *out1, uint64_t N, int l, int key_limbs,
                      const uint64_t *primes, uint64_t p, const uint64_t *psis,
                      const uint64_t *psi_p) {
    int nb = l + 1;
    uint64_t *acc0 = calloc((size_t)nb * N, 8), *acc1 = calloc((size_t)nb * N, 8);
    uint64_t *dig = malloc(N * 8), *tmp = malloc(N * 8);
    for (int j = 0; j < l; j++) {
        /* key layout [digit][2][key_limbs][N]; the p limb is the last one */
        const uint64_t *bj = key + ((uint64_t)j * 2 + 0) * key_limbs * N;
        const uint64_t *aj = key + ((uint64_t)j * 2 + 1) * key_limbs * N;
        for (int i = 0; i < nb; i++) {
            uint64_t qi = i < l ? primes[i] : p;
            uint64_t ps = i < l ? psis[i] : *psi_p;
            int ki = i < l ? i : key_limbs - 1;
            for (uint64_t k = 0; k < N; k++) dig[k] = d2[(uint64_t)j * N + k] % qi;
            oracle_polymul_ntt(dig, bj + (uint64_t)ki * N, tmp, N, qi, ps);
            for (uint64_t k = 0; k < N; k++)
                acc0[(uint64_t)i * N + k] = addmod(acc0[(uint64_t)i * N + k], tmp[k], qi);
            oracle_polymul_ntt(dig, aj + (uint64_t)ki * N, tmp, N, qi, ps);
            for (uint64_t k = 0; k < N; k++)
                acc1[(uint64_t)i * N + k] = submod(acc1[(uint64_t)i * N + k], tmp[k], qi);
        }
    }
    oracle_mod_down(acc0, out0, N, l, primes, p);
    oracle_mod_down(acc1, out1, N, l, primes, p);
    free(acc0); free(acc1); free(dig); free(tmp);
}

/* The same key switch with the digits given as signed integers
 * (digits[j][k] in (-q_j, q_j)), each reduced into every target prime. */
static void keyswitch_signed(const int64_t *digits, const uint64_t *key, uint64_t *out0,
                             uint64_t *out1, uint64_t N, int l, int key_limbs,
                             const uint64_t *primes, uint64_t p, const uint64_t *psis,
                             const uint64_t *psi_p) {
    int nb = l + 1;
    uint64_t *acc0 = calloc((size_t)nb * N, 8), *acc1 = calloc((size_t)nb * N, 8);
    uint64_t *dig = malloc(N * 8), *tmp = malloc(N * 8);
    for (int j = 0; j < l; j++) {
        const uint64_t *bj = key + ((uint64_t)j * 2 + 0) * key_limbs * N;
        const uint64_t *aj = key + ((uint64_t)j * 2 + 1) * key_limbs * N;
        for (int i = 0; i < nb; i++) {
            uint64_t qi = i < l ? primes[i] : p;
            uint64_t ps = i < l ? psis[i] : *psi_p;
            int ki = i < l ? i : key_limbs - 1;
            for (uint64_t k = 0; k < N; k++) {
                int64_t d = digits[(uint64_t)j * N + k];
                uint64_t m = (uint64_t)(d < 0 ? -d : d) % qi;
                dig[k] = (d < 0 && m) ? qi - m : m;
            }
            oracle_polymul_ntt(dig, bj + (uint64_t)ki * N, tmp, N, qi, ps);
            for (uint64_t k = 0; k < N; k++)
                acc0[(uint64_t)i * N + k] = addmod(acc0[(uint64_t)i * N + k], tmp[k], qi);
            oracle_polymul_ntt(dig, aj + (uint64_t)ki * N, tmp, N, qi, ps);
            for (uint64_t k = 0; k < N; k++)
                acc1[(uint64_t)i * N + k] = submod(acc1[(uint64_t)i * N + k], tmp[k], qi);
        }
    }
    oracle_mod_down(acc0, out0, N, l, primes, p);
    oracle_mod_down(acc1, out1, N, l, primes, p);
    free(acc0); free(acc1); free(dig); free(tmp);
}

/* HMult + relinearise + rescale on one ciphertext pair, coefficient form
 * (scripts/verify_e2e_mult.py:78-102 with the exact rescale of
 * verify_rescale.py).  out: [2][l-1][N]. */
void oracle_hmult(const uint64_t *cta, const uint64_t *ctb, const uint64_t *key,
                  uint64_t *out, uint64_t N, int l, int key_limbs,
                  const uint64_t *primes, uint64_t p, const uint64_t *psis,
                  const uint64_t *psi_p) {
    uint64_t LN = (uint64_t)l * N;
    uint64_t *d0 = calloc(LN, 8), *d1 = calloc(LN, 8), *d2 = calloc(LN, 8);
    uint64_t *t = malloc(N * 8);
    for (int i = 0; i < l; i++) {
        uint64_t q = primes[i], ps = psis[i], o = (uint64_t)i * N;
        const uint64_t *a0 = cta + o, *a1 = cta + LN + o, *b0 = ctb + o, *b1 = ctb + LN + o;
        oracle_polymul_ntt(a0, b0, d0 + o, N, q, ps);
        oracle_polymul_ntt(a0, b1, d1 + o, N, q, ps);
        oracle_polymul_ntt(a1, b0, t, N, q, ps);
        for (uint64_t k = 0; k < N; k++) d1[o + k] = addmod(d1[o + k], t[k], q);
        oracle_polymul_ntt(a1, b1, d2 + o, N, q, ps);
    }
    uint64_t *k0 = malloc(LN * 8), *k1 = malloc(LN * 8);
    oracle_keyswitch(d2, key, k0, k1, N, l, key_limbs, primes, p, psis, psi_p);
    for (int i = 0; i < l; i++)
        for (uint64_t k = 0; k < N; k++) {
            uint64_t o = (uint64_t)i * N + k;
            d0[o] = addmod(d0[o], k0[o], primes[i]);
            d1[o] = addmod(d1[o], k1[o], primes[i]);
        }
    oracle_rescale(d0, out, N, l, primes);
    oracle_rescale(d1, out + (uint64_t)(l - 1) * N, N, l, primes);
    free(d0); free(d1); free(d2); free(t); free(k0); free(k1);
}

/* f(X) -> f(X^g) mod X^N + 1 for odd g (exp_e24_rotation_bootstrap.py:143-162). */
void oracle_automorphism(const uint64_t *in, uint64_t *out, uint64_t N, uint64_t q,
                         uint64_t g) {
    for (uint64_t k = 0; k < N; k++) out[k] = 0;
    for (uint64_t k = 0; k < N; k++) {
        uint64_t t = (g * k) % (2 * N);
        uint64_t j = t % N;
        uint64_t c = in[k] % q;
        out[j] = t < N ? addmod(out[j], c, q) : submod(out[j], c, q);
    }
}

/* Rotation by the Galois element g: (sigma(c0), sigma(c1)) key-switched
 * from sigma(s) to s.  Key layout as in oracle_keyswitch.  out: [2][l][N]. */
void oracle_rotate(const uint64_t *ct, const uint64_t *key, uint64_t *out, uint64_t N,
                   int l, int key_limbs, const uint64_t *primes, uint64_t p,
                   const uint64_t *psis, const uint64_t *psi_p, uint64_t g) {
    uint64_t LN = (uint64_t)l * N;
    uint64_t *s0 = calloc(LN, 8), *s1 = calloc(LN, 8), *k1 = calloc(LN, 8);
    for (int i = 0; i < l; i++) {
        oracle_automorphism(ct + (uint64_t)i * N, s0 + (uint64_t)i * N, N, primes[i], g);
        oracle_automorphism(ct + LN + (uint64_t)i * N, s1 + (uint64_t)i * N, N, primes[i], g);
    }
    oracle_keyswitch(s1, key, out, k1, N, l, key_limbs, primes, p, psis, psi_p);
    for (int i = 0; i < l; i++)
        for (uint64_t k = 0; k < N; k++) {
            uint64_t o = (uint64_t)i * N + k;
            out[o] = addmod(out[o], s0[o], primes[i]);
            out[LN + o] = k1[o];
        }
    free(s0); free(s1); free(k1);
}

/* Hoisted rotation (Halevi-Shoup): the digits are the automorphism over Z
 * of the canonical lifts of c1's limbs -- coefficient k of [c1]_(q_j) goes
 * to (g k) mod N, negated when (g k) mod 2N >= N -- key-switched with the
 * Galois key and added to sigma(c0).  out: [2][l][N]. */
void oracle_rotate_hoisted(const uint64_t *ct, const uint64_t *key, uint64_t *out, uint64_t N,
                           int l, int key_limbs, const uint64_t *primes, uint64_t p,
                           const uint64_t *psis, const uint64_t *psi_p, uint64_t g) {
    uint64_t LN = (uint64_t)l * N;
    int64_t *dig = calloc(LN, 8);
    uint64_t *s0 = calloc(LN, 8), *k1 = calloc(LN, 8);
    for (int j = 0; j < l; j++) {
        const uint64_t *c1 = ct + LN + (uint64_t)j * N;
        for (uint64_t k = 0; k < N; k++) {
            uint64_t t = (g * k) % (2 * N);
            dig[(uint64_t)j * N + t % N] = t < N ? (int64_t)c1[k] : -(int64_t)c1[k];
        }
        oracle_automorphism(ct + (uint64_t)j * N, s0 + (uint64_t)j * N, N, primes[j], g);
    }
    keyswitch_signed(dig, key, out, k1, N, l, key_limbs, primes, p, psis, psi_p);
    for (int i = 0; i < l; i++)
        for (uint64_t k = 0; k < N; k++) {
            uint64_t o = (uint64_t)i * N + k;
            out[o] = addmod(out[o], s0[o], primes[i]);
            out[LN + o] = k1[o];
        }
    free(dig); free(s0); free(k1);
}

