// Host-side checks of the C ABI (cipher-pu_amd/csrc/capi.cpp), built with
// AddressSanitizer on the host code by tests/test_capi_asan.py and run on a
// machine without a GPU: argument checking of every entry point that rejects
// a call before touching the device, the context's host tables (built before
// the device is looked for), and the ModDown / CRT constants against 128-bit
// arithmetic.  The translation unit includes capi.cpp itself, so the
// constants' builders are reachable.  Exit status 0 = all checks passed.
#include "../cipher-pu_amd/csrc/capi.cpp"

#include <stdio.h>

static int failures = 0;
#define CHECK(cond)                                                   \
    do {                                                              \
        if (!(cond)) {                                                \
            fprintf(stderr, "FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                               \
        }                                                             \
    } while (0)

static const uint64_t CHAIN4[4] = {17592186634241ull, 17592186372097ull, 17592186273793ull, 17592186175489ull};
static const uint64_t SPECIAL = 17592186765313ull;

static uint64_t mulmod128(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

static void check_create_arguments() {
    ckks_ctx* c = nullptr;
    const uint64_t q = CHAIN4[0];
    CHECK(ckks_ctx_create(nullptr, 8192, 1, &q, nullptr, 0, 0) == CKKS_ERR_NULL);
    CHECK(ckks_ctx_create(&c, 8192, 1, nullptr, nullptr, 0, 0) == CKKS_ERR_NULL);
    for (uint64_t N : {0ull, 1ull, 3ull, 12ull, 32768ull}) CHECK(ckks_ctx_create(&c, N, 1, &q, nullptr, 0, 0) == CKKS_ERR_BAD_N);
    CHECK(ckks_ctx_create(&c, 8192, 0, &q, nullptr, 0, 0) == CKKS_ERR_BAD_ARG);
    CHECK(ckks_ctx_create(&c, 8192, CKKS_MAX_PRIMES + 1, &q, nullptr, 0, 0) == CKKS_ERR_BAD_ARG);
    const uint64_t bad[] = {2, 15, 17592186634243ull /* composite */, (1ull << 45) + 1, 7681 /* 2N does not divide q-1 */};
    for (uint64_t b : bad) CHECK(ckks_ctx_create(&c, 8192, 1, &b, nullptr, 0, 0) == CKKS_ERR_BAD_MODULUS);
    const uint64_t not_root = 3;
    CHECK(ckks_ctx_create(&c, 8192, 1, &q, &not_root, 0, 0) == CKKS_ERR_BAD_MODULUS);
    CHECK(c == nullptr);
    CHECK(ckks_ctx_create(&c, 8192, 1, &q, nullptr, 0, -1) == CKKS_ERR_NO_DEVICE);
}

// Valid contexts of every ring degree: the host tables are built in full
// before the device is looked for, so on a machine without a GPU this walks
// all of the table construction and returns CKKS_ERR_NO_DEVICE.
static void check_host_tables() {
    uint64_t primes[5] = {CHAIN4[0], CHAIN4[1], CHAIN4[2], CHAIN4[3], SPECIAL};
    for (uint64_t N = 2; N <= 16384; N <<= 1)
        for (int flags : {0, CKKS_CYCLIC}) {
            ckks_ctx* c = nullptr;
            const int st = ckks_ctx_create(&c, N, 5, primes, nullptr, flags, 0);
            CHECK(st == CKKS_OK || st == CKKS_ERR_NO_DEVICE);
            if (st == CKKS_OK) {
                uint64_t r = 0;
                CHECK(ckks_ctx_root(c, 4, &r) == CKKS_OK && r != 0);
                CHECK(ckks_ctx_root(c, 5, &r) == CKKS_ERR_BAD_SHAPE);
                CHECK(ckks_ctx_destroy(c) == CKKS_OK);
            } else {
                CHECK(c == nullptr);
            }
        }
    // explicit roots: psi of order 2N for the negacyclic ring
    const uint64_t q = 35184371884033ull, psi = 7388705119ull;
    ckks_ctx* c = nullptr;
    const int st = ckks_ctx_create(&c, 4096, 1, &q, &psi, 0, 0);
    CHECK(st == CKKS_OK || st == CKKS_ERR_NO_DEVICE);
    if (c) ckks_ctx_destroy(c);
}

// ModDown (by p) and rescale (by q_(l-1)) constants, and the decode CRT
// constants, on a context struct filled on the host.
static void check_constants() {
    ckks_ctx c;
    c.N = 8192;
    c.logN = 13;
    c.K = 5;
    for (int i = 0; i < 4; ++i) c.primes[i] = CHAIN4[i];
    c.primes[4] = SPECIAL;
    for (int l = 1; l <= 4; ++l) {
        const ModDownConst m = moddown_const(&c, l, 4);
        const uint64_t qL = c.primes[l - 1];
        for (int i = 0; i < l; ++i) {
            const uint64_t qi = c.primes[i];
            CHECK(m.hp[i] == (SPECIAL >> 1) % qi);
            CHECK(mulmod128(m.pinv[i], SPECIAL % qi, qi) == 1);
            CHECK(m.pinv_s[i] == (uint64_t)(((u128)m.pinv[i] << 63) / qi));
            CHECK(m.p_mod[i] == SPECIAL % qi);
            if (i < l - 1) {
                CHECK(m.hq[i] == (qL >> 1) % qi);
                CHECK(mulmod128(m.qinv[i], qL % qi, qi) == 1);
                CHECK(m.qinv_s[i] == (uint64_t)(((u128)m.qinv[i] << 63) / qi));
            }
        }
    }
    for (int l = 1; l <= 5; ++l) {
        CrtConst cc;
        CHECK(crt_const(&c, 0, l, &cc) == CKKS_OK);
        // Q = prod q_i, and qhat_i q_i = Q word for word
        uint64_t Q[CRT_MAX_WORDS] = {1};
        for (int i = 0; i < l; ++i) mw_mul_small(Q, cc.words, c.primes[i]);
        for (int k = 0; k < cc.words; ++k) CHECK(Q[k] == cc.Q[k]);
        for (int i = 0; i < l; ++i) {
            uint64_t t[CRT_MAX_WORDS];
            for (int k = 0; k < CRT_MAX_WORDS; ++k) t[k] = cc.qhat[i][k];
            mw_mul_small(t, cc.words, c.primes[i]);
            for (int k = 0; k < cc.words; ++k) CHECK(t[k] == cc.Q[k]);
            const uint64_t qi = c.primes[i];
            CHECK(mulmod128(cc.qhat_inv[i], mw_mod_small(cc.qhat[i], cc.words, qi), qi) == 1);
        }
        // half_Q = floor(Q / 2): doubling it gives Q or Q - 1
        uint64_t h2[CRT_MAX_WORDS] = {0};
        for (int k = 0; k < cc.words; ++k) h2[k] = cc.half_Q[k];
        mw_mul_small(h2, cc.words, 2);
        CHECK(h2[0] + (Q[0] & 1) == Q[0]);
        for (int k = 1; k < cc.words; ++k) CHECK(h2[k] == Q[k]);
    }
    CrtConst cc;
    CHECK(crt_const(&c, 0, CRT_MAX_LIMBS + 1, &cc) == CKKS_ERR_BAD_SHAPE);
}

static void check_gaussian_table() {
    uint64_t cdf[GAUSS_TABLE];
    CHECK(ckks_gaussian_cdf(3.2, cdf, GAUSS_TABLE) == CKKS_OK);
    for (int k = 1; k < GAUSS_TABLE; ++k) CHECK(cdf[k] >= cdf[k - 1]);
    CHECK(cdf[GAUSS_TABLE - 1] == ~0ull);
    CHECK(ckks_gaussian_cdf(3.2, nullptr, 4) == CKKS_ERR_NULL);
    CHECK(ckks_gaussian_cdf(3.2, cdf, GAUSS_TABLE + 1) != CKKS_OK);
    CHECK(ckks_gaussian_cdf(-1.0, cdf, 4) != CKKS_OK);
}

// Every compute entry point rejects a NULL context or bad buffers before any
// device call.
static void check_entry_arguments() {
    alignas(16) uint64_t buf[64] = {0};
    uint64_t* mis = (uint64_t*)((char*)buf + 4);
    CHECK(ckks_ntt_fwd(nullptr, buf, 1, 1, 0, 0, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_ntt_inv(nullptr, buf, 1, 1, 0, 0, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_poly_mul(nullptr, buf, buf, buf, 1, 1, 0, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_pointwise(nullptr, buf, buf, buf, 1, 1, 0, 0, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_rescale(nullptr, buf, buf, 1, 2, 0, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_relin(nullptr, buf, buf, buf, 5, buf, 1, 4, 4, buf, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_hmult(nullptr, buf, buf, buf, 5, buf, 1, 4, 4, buf, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_hmult_noalloc(nullptr, buf, buf, buf, 5, buf, 1, 4, 4, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_rotate(nullptr, buf, buf, 5, buf, 1, 4, 4, 5, buf, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_automorphism(nullptr, buf, buf, 1, 1, 0, 5, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_noise_stats(nullptr, buf, buf, 1, 1, 0, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_lincomb(nullptr, buf, buf, buf, 1, 1, 1, 1, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_encode(nullptr, (double*)buf, buf, 1, 1, 0, 1.0, (double*)buf, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_timing_start(nullptr, 4) == CKKS_ERR_NULL);
    CHECK(ckks_workspace_bytes(nullptr, 4, 4) == 0);
    CHECK(ckks_ctx_destroy(nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_version() != nullptr);
    // context-free entry points: shapes, pointers, alignment, moduli
    CHECK(ckks_mod_switch(buf, buf, 0, 113, 7, nullptr) == CKKS_OK);
    CHECK(ckks_mod_switch(nullptr, buf, 4, 113, 7, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_mod_switch(mis, buf, 4, 113, 7, nullptr) == CKKS_ERR_BAD_ALIGNMENT);
    CHECK(ckks_mod_switch(buf, buf, 4, 7, 113, nullptr) == CKKS_ERR_BAD_MODULUS);
    CHECK(ckks_mod_switch_scaled(buf, buf, 0, 113, 113, 30, nullptr) == CKKS_OK);
    CHECK(ckks_mod_switch_scaled(buf, nullptr, 4, 113, 113, 30, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_mod_switch_scaled(buf, mis, 4, 113, 113, 30, nullptr) == CKKS_ERR_BAD_ALIGNMENT);
    CHECK(ckks_mod_switch_scaled(buf, buf, 4, 1, 113, 30, nullptr) == CKKS_ERR_BAD_MODULUS);
    CHECK(ckks_mod_switch_scaled(buf, buf, 4, 113, 1ull << 45, 30, nullptr) == CKKS_ERR_BAD_MODULUS);
    CHECK(ckks_mod_switch_scaled(buf, buf, 4, 113, 113, 128, nullptr) == CKKS_ERR_BAD_ARG);
    CHECK(ckks_noise_stats_q(buf, buf, 0, 32, 113, nullptr) == CKKS_OK);
    CHECK(ckks_noise_stats_q(buf, buf, -1, 32, 113, nullptr) == CKKS_ERR_BAD_SHAPE);
    CHECK(ckks_noise_stats_q(nullptr, buf, 1, 32, 113, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_noise_stats_q(buf, buf, 1, 32, 1, nullptr) == CKKS_ERR_BAD_MODULUS);
    const uint64_t q45 = 17592186765313ull;
    CHECK(ckks_probe_dft64(nullptr, buf, buf, q45, 1, 64, 0, 1, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_probe_dft64(buf, buf, nullptr, q45, 1, 64, 0, 1, nullptr) == CKKS_ERR_NULL);
    for (size_t n : {(size_t)0, (size_t)32, (size_t)96}) CHECK(ckks_probe_dft64(buf, buf, buf, q45, 1, n, 0, 1, nullptr) == CKKS_ERR_BAD_SHAPE);
    for (int m : {-1, 4}) CHECK(ckks_probe_dft64(buf, buf, buf, q45, 1, 64, m, 1, nullptr) == CKKS_ERR_BAD_ARG);
    for (int r : {0, 65}) CHECK(ckks_probe_dft64(buf, buf, buf, q45, 1, 64, 1, r, nullptr) == CKKS_ERR_BAD_ARG);
    for (uint64_t bq : {(uint64_t)2, q45 + 1, (uint64_t)1 << 45}) CHECK(ckks_probe_dft64(buf, buf, buf, bq, 1, 64, 1, 1, nullptr) == CKKS_ERR_BAD_MODULUS);
    CHECK(ckks_probe_dft64(buf, buf, (char*)buf + 8, q45, 1, 64, 1, 1, nullptr) == CKKS_ERR_BAD_ALIGNMENT);
    CHECK(ckks_probe_dft64(mis, buf, buf, q45, 1, 64, 1, 1, nullptr) == CKKS_ERR_BAD_ALIGNMENT);
    CHECK(ckks_noise_stats_q(buf, mis, 1, 32, 113, nullptr) == CKKS_ERR_BAD_ALIGNMENT);
    CHECK(ckks_poly_mul_modq(nullptr, buf, buf, buf, 1, 113, buf, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_bitrev(buf, 1, 12, nullptr) == CKKS_ERR_BAD_N);
    CHECK(ckks_bitrev(buf, -1, 16, nullptr) == CKKS_ERR_BAD_SHAPE);
    CHECK(ckks_bitrev(nullptr, 1, 16, nullptr) == CKKS_ERR_NULL);
    CHECK(ckks_bitrev(mis, 1, 16, nullptr) == CKKS_ERR_BAD_ALIGNMENT);
}

int main() {
    check_create_arguments();
    check_host_tables();
    check_constants();
    check_gaussian_table();
    check_entry_arguments();
    if (failures) {
        fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    printf("capi host checks passed\n");
    return 0;
}
