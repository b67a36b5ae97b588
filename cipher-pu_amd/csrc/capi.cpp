// C ABI of the CKKS evaluation core: contexts, argument checking and the
// composition of kernels into the reference's operations.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "ckks_core.h"
#include "launch.h"

// Scratch of ckks_hmult_noalloc.  A device variable: every GPU the library
// runs on holds its own copy, at its own address, until the process ends.
// In host code this name is only a stand-in for the symbol; the device
// address comes from hipGetSymbolAddress (scratch_on_device), and only that
// address is ever handed to a kernel.
static __device__ __align__(256) uint64_t scratch[CKKS_NOALLOC_SCRATCH_BYTES / 8];

namespace {

typedef unsigned __int128 u128;

uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)(((u128)a * b) % q); }

uint64_t powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = mulmod(r, b, q);
        b = mulmod(b, b, q);
        e >>= 1;
    }
    return r;
}

uint64_t invmod(uint64_t a, uint64_t q) { return powmod(a % q, q - 2, q); }

bool is_prime(uint64_t n) {
    if (n < 2) return false;
    static const uint64_t bases[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t p : bases)
        if (n % p == 0) return n == p;
    uint64_t d = n - 1;
    int s = 0;
    while (!(d & 1)) { d >>= 1; ++s; }
    for (uint64_t a : bases) {
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int r = 1; r < s; ++r) {
            x = mulmod(x, x, n);
            if (x == n - 1) { comp = false; break; }
        }
        if (comp) return false;
    }
    return true;
}

// smallest generator of Z_q^* (tb/ref/gen_ntt_reference.py:38-61)
uint64_t primitive_root(uint64_t q) {
    std::vector<uint64_t> fs;
    uint64_t n = q - 1;
    for (uint64_t d = 2; d * d <= n; d += (d == 2 ? 1 : 2)) {
        if (n % d == 0) {
            fs.push_back(d);
            while (n % d == 0) n /= d;
        }
    }
    if (n > 1) fs.push_back(n);
    for (uint64_t g = 2;; ++g) {
        bool ok = true;
        for (uint64_t f : fs)
            if (powmod(g, (q - 1) / f, q) == 1) { ok = false; break; }
        if (ok) return g;
    }
}

// Shoup companion floor(w 2^63 / q) (see mul_shoup in modarith.h)
uint64_t shoup(uint64_t w, uint64_t q) { return (uint64_t)(((u128)w << 63) / q); }

uint32_t bitrev(uint32_t i, int bits) {
    uint32_t r = 0;
    for (int b = 0; b < bits; ++b) { r = (r << 1) | (i & 1); i >>= 1; }
    return r;
}

int bitlen(uint64_t x) { int l = 0; while (x) { ++l; x >>= 1; } return l; }

}  // namespace

struct ckks_ctx {
    int device = 0, logN = 0, K = 0, flags = 0;
    uint64_t N = 0;
    uint64_t primes[CKKS_MAX_PRIMES] = {}, roots[CKKS_MAX_PRIMES] = {};
    PrimeConst hpc[CKKS_MAX_PRIMES] = {};
    PrimeConst* dpc = nullptr;
    ulonglong2* dtw = nullptr;
    PrimeF* dpf = nullptr;        // FP64 pipeline: per-prime constants
    double* dtwf = nullptr;       // FP64 pipeline: the twiddle words as doubles, order of dtw
    int fp64 = 1;                 // transform arithmetic (ckks_ctx_set_arith)
    double2* fft_tw = nullptr;    // exp(2 pi i k / n), k < n/2, n = N/2
    double2* fft_zeta = nullptr;  // exp(i pi j / N), j < n
    int* slot_k = nullptr;        // (5^j mod 2N - 1)/4
    int ntt_threads = 0, ks_threads = 0, tensor_threads = 0;
    int front_fused = 0;   // HMult front: one fused kernel (1) or 2 transforms + tensor (0)
    HoistOrder hoist = {0, 0};   // hoisted key-switch placement; btile 0 = default for the output form
    int ks_pair_major = 1;       // HMult key-switch placement (ckks_ctx_set_ks_order)
    // Optional per-launch timing (ckks_timing_start/stop): an event pair
    // around every kernel the entry points launch, tagged by kernel kind.
    struct Timer {
        bool on = false;
        int used = 0;
        std::vector<hipEvent_t> ev;   // 2 per slot
        std::vector<int> kind;
    };
    mutable Timer timer;
    int narrow = 0;        // every prime below twice every other one
    void* scratch = nullptr;   // this device's copy of the no-alloc scratch array
    int n_cu = 256;        // compute units of the device
    DevTables tables() const { return DevTables{dpc, dtw, (int)N, logN, narrow, dpf, dtwf, fp64, n_cu}; }
};

namespace {

#define HIPCHK(x)                                  \
    do {                                           \
        if ((x) != hipSuccess) return CKKS_ERR_HIP; \
    } while (0)

inline bool misaligned(const void* p) { return ((uintptr_t)p & 7) != 0; }
// (value, Shoup companion) pair buffers -- prepared keys and plaintexts --
// are read as 16-byte pairs
inline bool misaligned_pairs(const void* p) { return ((uintptr_t)p & 15) != 0; }

// CKKS_OK, an error, or NOTHING_TO_DO for a valid call with zero rows
const int NOTHING_TO_DO = 1;

int check_rows(const ckks_ctx* ctx, int n_polys, int n_limbs, int prime0) {
    if (!ctx) return CKKS_ERR_NULL;
    if (n_polys < 0 || n_limbs < 1 || prime0 < 0 || prime0 + n_limbs > ctx->K) return CKKS_ERR_BAD_SHAPE;
    return n_polys == 0 ? NOTHING_TO_DO : CKKS_OK;
}

LimbMap limb_map(int prime0, int n) {
    LimbMap m;
    for (int i = 0; i < CKKS_MAX_PRIMES; ++i) m.idx[i] = i < n ? prime0 + i : 0;
    return m;
}

hipStream_t S(void* s) { return (hipStream_t)s; }

// Timing slots: -1 when timing is off or every slot is taken.
int timer_begin(const ckks_ctx* c, int kind, hipStream_t s) {
    auto& t = c->timer;
    if (!t.on || 2 * (t.used + 1) > (int)t.ev.size()) return -1;
    const int slot = t.used++;
    t.kind[slot] = kind;
    hipEventRecord(t.ev[2 * slot], s);
    return slot;
}

void timer_end(const ckks_ctx* c, int slot, hipStream_t s) {
    if (slot >= 0) hipEventRecord(c->timer.ev[2 * slot + 1], s);
}

// Launch a kernel through its launcher, timed when timing is on.
#define TIMED(KIND, CALL)                                 \
    do {                                                  \
        const int slot_ = timer_begin(c, KIND, S(stream)); \
        const hipError_t e_ = (CALL);                     \
        timer_end(c, slot_, S(stream));                   \
        if (e_ != hipSuccess) return CKKS_ERR_HIP;        \
    } while (0)

int set_device(const ckks_ctx* ctx) { return hipSetDevice(ctx->device) == hipSuccess ? CKKS_OK : CKKS_ERR_HIP; }

// Device address of the scratch array on the current device, looked up once
// per device (with that device current) and kept.
const int SCRATCH_DEVICES = 64;
std::mutex scratch_mu;
void* scratch_addr[SCRATCH_DEVICES] = {};

void* scratch_on_device(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES) return nullptr;
    std::lock_guard<std::mutex> g(scratch_mu);
    if (!scratch_addr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
        scratch_addr[device] = p;
    }
    return scratch_addr[device];
}

// constants for ModDown by the special prime sp and rescale by limb l-1
ModDownConst moddown_const(const ckks_ctx* c, int l, int sp) {
    ModDownConst m;
    memset(&m, 0, sizeof(m));
    const uint64_t p = sp >= 0 ? c->primes[sp] : 1;
    const uint64_t qL = c->primes[l - 1];
    for (int i = 0; i < l && i < CKKS_MAX_PRIMES; ++i) {
        const uint64_t qi = c->primes[i];
        if (sp >= 0) {
            m.hp[i] = (p >> 1) % qi;
            m.pinv[i] = invmod(p, qi);
            m.pinv_s[i] = shoup(m.pinv[i], qi);
            m.p_mod[i] = p % qi;
            m.p_mod_s[i] = shoup(m.p_mod[i], qi);
            m.pmf[i] = (double)m.p_mod[i];
            m.pmf_q[i] = (double)m.p_mod[i] / (double)qi;
            m.hpf[i] = (double)m.hp[i];
            m.pinvf[i] = (double)m.pinv[i];
            m.pinvf_q[i] = (double)m.pinv[i] / (double)qi;
        }
        if (i < l - 1) {
            m.hq[i] = (qL >> 1) % qi;
            m.qinv[i] = invmod(qL, qi);
            m.qinv_s[i] = shoup(m.qinv[i], qi);
            m.hqf[i] = (double)m.hq[i];
            m.qinvf[i] = (double)m.qinv[i];
            m.qinvf_q[i] = (double)m.qinv[i] / (double)qi;
        }
    }
    return m;
}

// little-endian multiword helpers for the CRT constants
void mw_mul_small(uint64_t* a, int w, uint64_t m) {
    u128 carry = 0;
    for (int k = 0; k < w; ++k) {
        u128 t = (u128)a[k] * m + carry;
        a[k] = (uint64_t)t;
        carry = t >> 64;
    }
}

uint64_t mw_mod_small(const uint64_t* a, int w, uint64_t m) {
    u128 r = 0;
    for (int k = w - 1; k >= 0; --k) r = ((r << 64) | a[k]) % m;
    return (uint64_t)r;
}

int crt_const(const ckks_ctx* c, int prime0, int l, CrtConst* cc) {
    if (l > CRT_MAX_LIMBS) return CKKS_ERR_BAD_SHAPE;
    memset(cc, 0, sizeof(*cc));
    cc->l = l;
    int bits = 2;
    for (int i = 0; i < l; ++i) bits += bitlen(c->primes[prime0 + i]);
    cc->words = (bits + 63) / 64;
    if (cc->words > CRT_MAX_WORDS) return CKKS_ERR_BAD_SHAPE;
    cc->Q[0] = 1;
    for (int i = 0; i < l; ++i) mw_mul_small(cc->Q, cc->words, c->primes[prime0 + i]);
    for (int i = 0; i < l; ++i) {
        const uint64_t qi = c->primes[prime0 + i];
        cc->qhat[i][0] = 1;
        for (int j = 0; j < l; ++j)
            if (j != i) mw_mul_small(cc->qhat[i], cc->words, c->primes[prime0 + j]);
        cc->qhat_inv[i] = invmod(mw_mod_small(cc->qhat[i], cc->words, qi), qi);
        cc->qhat_inv_s[i] = shoup(cc->qhat_inv[i], qi);
    }
    for (int k = 0; k < cc->words; ++k)
        cc->half_Q[k] = (cc->Q[k] >> 1) | (k + 1 < cc->words ? cc->Q[k + 1] << 63 : 0);
    return CKKS_OK;
}

}  // namespace

extern "C" {

const char* ckks_version(void) { return "cipherpu-ckks 0.1 gfx950"; }

int ckks_ctx_create(ckks_ctx** out, uint64_t N, int n_primes, const uint64_t* primes, const uint64_t* roots,
                    int flags, int device) {
    if (!out || !primes) return CKKS_ERR_NULL;
    *out = nullptr;
    if (N < 2 || N > 16384 || (N & (N - 1))) return CKKS_ERR_BAD_N;
    if (n_primes < 1 || n_primes > CKKS_MAX_PRIMES) return CKKS_ERR_BAD_ARG;
    const bool cyclic = flags & CKKS_CYCLIC;
    const uint64_t order = cyclic ? N : 2 * N;
    for (int i = 0; i < n_primes; ++i) {
        const uint64_t q = primes[i];
        if (q < 3 || q >= (1ull << 45) || !is_prime(q) || (q - 1) % order) return CKKS_ERR_BAD_MODULUS;
        if (roots) {
            const uint64_t r = roots[i] % q;
            if (powmod(r, order, q) != 1 || powmod(r, order / 2, q) != q - 1) return CKKS_ERR_BAD_MODULUS;
        }
    }
    // host tables first, then the device: the table construction is plain
    // host code (checked under AddressSanitizer on machines without a GPU)
    ckks_ctx* c = new ckks_ctx();
    c->device = device;
    c->N = N;
    c->logN = bitlen(N) - 1;
    c->K = n_primes;
    c->flags = flags;
    {
        uint64_t lo = primes[0], hi = primes[0];
        for (int i = 1; i < n_primes; ++i) {
            lo = primes[i] < lo ? primes[i] : lo;
            hi = primes[i] > hi ? primes[i] : hi;
        }
        c->narrow = hi < 2 * lo;
    }
    const int logN = c->logN;
    std::vector<ulonglong2> tw((size_t)n_primes * 2 * N);
    std::vector<double> twf((size_t)n_primes * 2 * N);
    std::vector<PrimeF> pf(n_primes);
    for (int p = 0; p < n_primes; ++p) {
        const uint64_t q = primes[p];
        const uint64_t r = roots ? roots[p] % q : powmod(primitive_root(q), (q - 1) / order, q);
        c->primes[p] = q;
        c->roots[p] = r;
        ulonglong2* f = &tw[(size_t)p * 2 * N];
        ulonglong2* iv = f + N;
        f[0] = make_ulonglong2(1, shoup(1, q));
        iv[0] = f[0];
        for (uint64_t k = 1; k < N; ++k) {
            int s = bitlen(k) - 1;
            uint64_t i = k - (1ull << s);
            uint64_t e;
            if (cyclic)
                e = (uint64_t)bitrev((uint32_t)i, s) * (N >> (s + 1));   // omega^(br_s(i) N / 2^(s+1))
            else
                e = bitrev((uint32_t)k, logN);                          // psi^(br_logN(k))
            const uint64_t w = powmod(r, e, q);
            const uint64_t wi = invmod(w, q);
            f[k] = make_ulonglong2(w, shoup(w, q));
            iv[k] = make_ulonglong2(wi, shoup(wi, q));
        }
        PrimeConst& pc = c->hpc[p];
        pc.q = q;
        pc.two_q = 2 * q;
        pc.ell = bitlen(q);
        pc.mu = (uint64_t)(((u128)1 << (2 * pc.ell)) / q);
        pc.mu64 = (uint64_t)(~0ull / q);
        pc.ninv = invmod(N % q, q);
        pc.ninv_s = shoup(pc.ninv, q);
        pc.ninv_w1 = mulmod(pc.ninv, iv[1].x, q);
        pc.ninv_w1_s = shoup(pc.ninv_w1, q);
        // the FP64 pipeline's constants and twiddles: integers below 2^45 are
        // exact doubles, and IEEE division rounds 1 / q and the N^-1
        // constants' quotients once (the twiddles' quotients are formed by
        // the butterflies from 1 / q)
        const double dq = (double)q;
        pf[p] = PrimeF{dq, -dq, 1.0 / dq, (double)pc.ninv, (double)pc.ninv / dq, (double)pc.ninv_w1,
                       (double)pc.ninv_w1 / dq, 0.0};
        double* ff = &twf[(size_t)p * 2 * N];
        for (uint64_t k = 0; k < 2 * N; ++k) ff[k] = (double)f[k].x;
    }
    // canonical-embedding tables (negacyclic rings): the length-N/2 FFT's
    // twiddles, the twist zeta^j and the 5^j slot order
    const bool emb = !cyclic && N >= 4;
    const uint64_t n = N / 2;
    std::vector<double2> ft(emb && n / 2 > 0 ? n / 2 : 1), zt(emb ? n : 1);
    std::vector<int> sk(emb ? n : 1);
    if (emb) {
        const long double PI = 3.141592653589793238462643383279502884L;
        for (uint64_t k = 0; k < n / 2; ++k) {
            const long double a = 2.0L * PI * (long double)k / (long double)n;
            ft[k] = make_double2((double)cosl(a), (double)sinl(a));
        }
        for (uint64_t j = 0; j < n; ++j) {
            const long double a = PI * (long double)j / (long double)N;
            zt[j] = make_double2((double)cosl(a), (double)sinl(a));
        }
        uint64_t e = 1;
        for (uint64_t j = 0; j < n; ++j) {
            sk[j] = (int)((e - 1) / 4);
            e = e * 5 % (2 * N);
        }
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= device || device < 0) {
        delete c;
        return CKKS_ERR_NO_DEVICE;
    }
    bool ok = hipSetDevice(device) == hipSuccess;
    ok = ok && hipMalloc(&c->dpc, sizeof(PrimeConst) * n_primes) == hipSuccess;
    ok = ok && hipMalloc(&c->dtw, tw.size() * sizeof(ulonglong2)) == hipSuccess;
    ok = ok && hipMemcpy(c->dpc, c->hpc, sizeof(PrimeConst) * n_primes, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(c->dtw, tw.data(), tw.size() * sizeof(ulonglong2), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMalloc(&c->dpf, pf.size() * sizeof(PrimeF)) == hipSuccess;
    ok = ok && hipMalloc(&c->dtwf, twf.size() * sizeof(double)) == hipSuccess;
    ok = ok && hipMemcpy(c->dpf, pf.data(), pf.size() * sizeof(PrimeF), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(c->dtwf, twf.data(), twf.size() * sizeof(double), hipMemcpyHostToDevice) == hipSuccess;
    if (ok && emb) {
        ok = ok && hipMalloc(&c->fft_tw, ft.size() * sizeof(double2)) == hipSuccess;
        ok = ok && hipMalloc(&c->fft_zeta, zt.size() * sizeof(double2)) == hipSuccess;
        ok = ok && hipMalloc(&c->slot_k, sk.size() * sizeof(int)) == hipSuccess;
        ok = ok && hipMemcpy(c->fft_tw, ft.data(), ft.size() * sizeof(double2), hipMemcpyHostToDevice) == hipSuccess;
        ok = ok && hipMemcpy(c->fft_zeta, zt.data(), zt.size() * sizeof(double2), hipMemcpyHostToDevice) == hipSuccess;
        ok = ok && hipMemcpy(c->slot_k, sk.data(), sk.size() * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    }
    if (!ok) {
        ckks_ctx_destroy(c);
        return CKKS_ERR_HIP;
    }
    // set-up of the no-alloc entry point, with the device current; a failed
    // lookup leaves scratch NULL and only ckks_hmult_noalloc reports it
    c->scratch = scratch_on_device(device);
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && n_cu > 0)
        c->n_cu = n_cu;
    *out = c;
    return CKKS_OK;
}

int ckks_ctx_destroy(ckks_ctx* c) {
    if (!c) return CKKS_ERR_NULL;
    hipSetDevice(c->device);
    for (hipEvent_t e : c->timer.ev) hipEventDestroy(e);
    if (c->dpc) hipFree(c->dpc);
    if (c->dtw) hipFree(c->dtw);
    if (c->dpf) hipFree(c->dpf);
    if (c->dtwf) hipFree(c->dtwf);
    if (c->fft_tw) hipFree(c->fft_tw);
    if (c->fft_zeta) hipFree(c->fft_zeta);
    if (c->slot_k) hipFree(c->slot_k);
    delete c;
    return CKKS_OK;
}

int ckks_ctx_root(const ckks_ctx* c, int prime, uint64_t* root) {
    if (!c || !root) return CKKS_ERR_NULL;
    if (prime < 0 || prime >= c->K) return CKKS_ERR_BAD_SHAPE;
    *root = c->roots[prime];
    return CKKS_OK;
}

int ckks_ctx_set_arith(ckks_ctx* c, int arith) {
    if (!c) return CKKS_ERR_NULL;
    if (arith != CKKS_ARITH_INT64 && arith != CKKS_ARITH_FP64) return CKKS_ERR_BAD_ARG;
    c->fp64 = arith == CKKS_ARITH_FP64;
    return CKKS_OK;
}

int ckks_ctx_set_front(ckks_ctx* c, int fused) {
    if (!c) return CKKS_ERR_NULL;
    c->front_fused = fused ? 1 : 0;
    return CKKS_OK;
}

int ckks_ctx_set_ks_order(ckks_ctx* c, int pair_major) {
    if (!c) return CKKS_ERR_NULL;
    c->ks_pair_major = pair_major ? 1 : 0;
    return CKKS_OK;
}

int ckks_ctx_set_hoist_order(ckks_ctx* c, int btile, int xcd) {
    if (!c) return CKKS_ERR_NULL;
    if (btile < 0) return CKKS_ERR_BAD_ARG;
    c->hoist = HoistOrder{btile, xcd ? 1 : 0};
    return CKKS_OK;
}

// cdf[k] = floor(2^64 P(|round(Z)| <= k)), Z ~ N(0, sigma^2), from the tail
// P(|round(Z)| > k) = erfc((k + 1/2) / (sigma sqrt 2)) in 80-bit arithmetic.
static void gauss_table(double sigma, GaussTable& t) {
    const long double two64 = 18446744073709551616.0L;
    for (int k = 0; k < GAUSS_TABLE; ++k) {
        const long double tail = erfcl((k + 0.5L) / ((long double)sigma * 1.41421356237309504880L)) * two64;
        t.cdf[k] = tail < 1.0L ? UINT64_MAX : (uint64_t)(two64 - tail);
    }
    t.cdf[GAUSS_TABLE - 1] = UINT64_MAX;
}

int ckks_gaussian_cdf(double sigma, uint64_t* cdf, int n) {
    if (!cdf) return CKKS_ERR_NULL;
    if (!(sigma > 0.0) || sigma > 10.0 || n < 1 || n > GAUSS_TABLE) return CKKS_ERR_BAD_ARG;
    GaussTable t;
    gauss_table(sigma, t);
    for (int k = 0; k < n; ++k) cdf[k] = t.cdf[k];
    return CKKS_OK;
}

int ckks_sample_uniform(const ckks_ctx* c, uint64_t* out, int rows, int l, int prime0, uint64_t seed,
                        void* stream) {
    int st = check_rows(c, rows, l, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!out) return CKKS_ERR_NULL;
    if (misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_sample_uniform(c->tables(), out, rows, l, limb_map(prime0, l), seed, S(stream)));
    return CKKS_OK;
}

int ckks_sample_gaussian(const ckks_ctx* c, uint64_t* out, int rows, int l, int prime0, double sigma, uint64_t seed,
                         void* stream) {
    int st = check_rows(c, rows, l, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!out) return CKKS_ERR_NULL;
    if (!(sigma > 0.0) || sigma > 10.0) return CKKS_ERR_BAD_ARG;
    if (misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    GaussTable t;
    gauss_table(sigma, t);
    HIPCHK(launch_sample_gauss(c->tables(), out, rows, l, limb_map(prime0, l), seed, t, S(stream)));
    return CKKS_OK;
}

int ckks_encrypt_sk(const ckks_ctx* c, const uint64_t* pt, const uint64_t* s_eval, uint64_t* out, int batch, int l,
                    double sigma, uint64_t seed_a, uint64_t seed_e, void* stream) {
    int st = check_rows(c, batch, l, 0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!pt || !s_eval || !out) return CKKS_ERR_NULL;
    if (pt == out || s_eval == out) return CKKS_ERR_BAD_ARG;
    if (!(sigma > 0.0) || sigma > 10.0) return CKKS_ERR_BAD_ARG;
    if (!c->fp64) return CKKS_ERR_BAD_ARG;
    if (c->logN < 12 || c->logN > 14 || (c->flags & CKKS_CYCLIC)) return CKKS_ERR_BAD_SHAPE;
    if (misaligned(pt) || misaligned(s_eval) || misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    GaussTable t;
    gauss_table(sigma, t);
    HIPCHK(launch_encrypt_sk(c->tables(), pt, s_eval, out, batch, l, limb_map(0, l), seed_a, seed_e, t, S(stream)));
    return CKKS_OK;
}

int ckks_decrypt_sk(const ckks_ctx* c, const uint64_t* ct, const uint64_t* s_eval, uint64_t* out, int batch, int l,
                    void* stream) {
    int st = check_rows(c, batch, l, 0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!ct || !s_eval || !out) return CKKS_ERR_NULL;
    if (ct == out || s_eval == out) return CKKS_ERR_BAD_ARG;
    if (!c->fp64) return CKKS_ERR_BAD_ARG;
    if (c->logN < 12 || c->logN > 14 || (c->flags & CKKS_CYCLIC)) return CKKS_ERR_BAD_SHAPE;
    if (misaligned(ct) || misaligned(s_eval) || misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_decrypt_sk(c->tables(), ct, s_eval, out, batch, l, limb_map(0, l), S(stream)));
    return CKKS_OK;
}

int ckks_timing_start(ckks_ctx* c, int max_launches) {
    if (!c) return CKKS_ERR_NULL;
    if (max_launches < 1 || max_launches > (1 << 20)) return CKKS_ERR_BAD_SHAPE;
    if (set_device(c)) return CKKS_ERR_HIP;
    auto& t = c->timer;
    while ((int)t.ev.size() < 2 * max_launches) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        t.ev.push_back(e);
    }
    t.kind.assign(t.ev.size() / 2, 0);
    t.used = 0;
    t.on = true;
    return CKKS_OK;
}

int ckks_timing_stop(ckks_ctx* c, double* ms, int* launches, int n_kinds) {
    if (!c || !ms || !launches) return CKKS_ERR_NULL;
    if (n_kinds < 1) return CKKS_ERR_BAD_SHAPE;
    auto& t = c->timer;
    t.on = false;
    for (int k = 0; k < n_kinds; ++k) {
        ms[k] = 0.0;
        launches[k] = 0;
    }
    if (t.used && hipEventSynchronize(t.ev[2 * t.used - 1]) != hipSuccess) return CKKS_ERR_HIP;
    for (int i = 0; i < t.used; ++i) {
        float e = 0.f;
        HIPCHK(hipEventElapsedTime(&e, t.ev[2 * i], t.ev[2 * i + 1]));
        if (t.kind[i] < n_kinds) {
            ms[t.kind[i]] += e;
            ++launches[t.kind[i]];
        }
    }
    const int n = t.used;
    t.used = 0;
    return n;
}

int ckks_ctx_set_threads(ckks_ctx* c, int ntt_threads, int ks_threads, int tensor_threads) {
    if (!c) return CKKS_ERR_NULL;
    for (int t : {ntt_threads, ks_threads, tensor_threads})
        if (t != 0 && t != 128 && t != 256 && t != 512 && t != 1024) return CKKS_ERR_BAD_ARG;
    c->ntt_threads = ntt_threads;
    c->ks_threads = ks_threads;
    c->tensor_threads = tensor_threads;
    return CKKS_OK;
}

static int do_ntt(const ckks_ctx* c, uint64_t* data, int n_polys, int n_limbs, int prime0, int flags, void* stream,
                  bool inv) {
    int st = check_rows(c, n_polys, n_limbs, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!data) return CKKS_ERR_NULL;
    if (misaligned(data)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    const DevTables t = c->tables();
    const int rows = n_polys * n_limbs;
    if (inv && (flags & CKKS_NATURAL_ORDER)) HIPCHK(launch_bitrev(data, rows, c->logN, S(stream)));
    TIMED(CKKS_KIND_NTT, launch_ntt(t, data, data, n_polys, n_limbs, limb_map(prime0, n_limbs), inv, false,
                                    c->ntt_threads, S(stream)));
    if (!inv && (flags & CKKS_NATURAL_ORDER)) HIPCHK(launch_bitrev(data, rows, c->logN, S(stream)));
    return CKKS_OK;
}

int ckks_ntt_fwd(const ckks_ctx* c, uint64_t* data, int n_polys, int n_limbs, int prime0, int flags, void* stream) {
    return do_ntt(c, data, n_polys, n_limbs, prime0, flags, stream, false);
}

int ckks_ntt_inv(const ckks_ctx* c, uint64_t* data, int n_polys, int n_limbs, int prime0, int flags, void* stream) {
    return do_ntt(c, data, n_polys, n_limbs, prime0, flags, stream, true);
}

int ckks_poly_mul(const ckks_ctx* c, const uint64_t* a, const uint64_t* b, uint64_t* out, int n_polys, int n_limbs,
                  int prime0, void* stream) {
    int st = check_rows(c, n_polys, n_limbs, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!a || !b || !out) return CKKS_ERR_NULL;
    if (misaligned(a) || misaligned(b) || misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    TIMED(CKKS_KIND_POLYMUL, launch_polymul(c->tables(), a, b, out, n_polys, n_limbs, limb_map(prime0, n_limbs),
                                            c->ntt_threads, S(stream)));
    return CKKS_OK;
}

int ckks_pointwise(const ckks_ctx* c, const uint64_t* a, const uint64_t* b, uint64_t* out, int n_polys, int n_limbs,
                   int prime0, int op, void* stream) {
    int st = check_rows(c, n_polys, n_limbs, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (op < 0 || op > 4) return CKKS_ERR_BAD_ARG;
    if (op == 4)   // special form: every prime of the window 2^44 + c, 0 < c < 2^21
        for (int i = 0; i < n_limbs; ++i) {
            const uint64_t q = c->primes[prime0 + i];
            if (q <= (1ull << 44) || q - (1ull << 44) >= (1ull << 21)) return CKKS_ERR_BAD_MODULUS;
        }
    if (!a || !out || (op != 3 && !b)) return CKKS_ERR_NULL;
    if (misaligned(a) || misaligned(out) || (b && misaligned(b))) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_pointwise(c->tables(), a, b, out, n_polys, n_limbs, limb_map(prime0, n_limbs), op, S(stream)));
    return CKKS_OK;
}

int ckks_rescale(const ckks_ctx* c, const uint64_t* in, uint64_t* out, int n_polys, int l, int prime0, void* stream) {
    int st = check_rows(c, n_polys, l, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (l < 2) return CKKS_ERR_BAD_SHAPE;
    if (!in || !out) return CKKS_ERR_NULL;
    if (misaligned(in) || misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    // constants are indexed by limb; shift the prime window to prime0
    ckks_ctx shifted = *c;
    for (int i = 0; i < l; ++i) shifted.primes[i] = c->primes[prime0 + i];
    const ModDownConst mc = moddown_const(&shifted, l, -1);
    HIPCHK(launch_rescale(c->tables(), in, out, n_polys, l, limb_map(prime0, l), mc, S(stream)));
    return CKKS_OK;
}

uint64_t ckks_workspace_bytes(const ckks_ctx* c, int batch, int l) {
    if (!c || batch < 0 || l < 1) return 0;
    return (uint64_t)batch * c->N * 8 * (uint64_t)(10 * l + 2);
}

int ckks_key_prepare(const ckks_ctx* c, const uint64_t* key_coeff, uint64_t* key_out, int dnum, int key_limbs,
                     int special, void* stream) {
    if (!c || !key_coeff || !key_out) return CKKS_ERR_NULL;
    if (dnum < 1 || key_limbs < 2 || key_limbs - 1 > c->K || special < 0 || special >= c->K)
        return CKKS_ERR_BAD_SHAPE;
    if (misaligned(key_coeff) || misaligned_pairs(key_out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    LimbMap m = limb_map(0, key_limbs);
    m.idx[key_limbs - 1] = special;
    const DevTables t = c->tables();
    uint64_t* tmp = const_cast<uint64_t*>(key_coeff);   // transformed in place (documented)
    HIPCHK(launch_ntt(t, tmp, tmp, dnum * 2, key_limbs, m, false, true, c->ntt_threads, S(stream)));
    HIPCHK(launch_prepare_key(t, tmp, key_out, dnum * 2, key_limbs, m, S(stream)));
    return CKKS_OK;
}

static int ks_common_check(const ckks_ctx* c, const void* key, int key_limbs, const void* out, int batch, int l,
                           int special, const void* ws) {
    if (!c) return CKKS_ERR_NULL;
    if (batch < 0 || l < 1 || l + 1 > CKKS_MAX_PRIMES || l > key_limbs - 1 || key_limbs - 1 > c->K)
        return CKKS_ERR_BAD_SHAPE;
    if (batch == 0) return NOTHING_TO_DO;
    if (!key || !out || !ws) return CKKS_ERR_NULL;
    if (special < 0 || special >= c->K || special < l) return CKKS_ERR_BAD_SHAPE;
    if (c->flags & CKKS_CYCLIC) return CKKS_ERR_BAD_ARG;
    if (c->logN < 12) return CKKS_ERR_BAD_N;
    if (misaligned_pairs(key) || misaligned(out) || misaligned(ws)) return CKKS_ERR_BAD_ALIGNMENT;
    return CKKS_OK;
}

int ckks_relin(const ckks_ctx* c, const uint64_t* d01, const uint64_t* d2, const uint64_t* key, int key_limbs,
               uint64_t* out, int batch, int l, int special, void* workspace, void* stream) {
    int st = ks_common_check(c, key, key_limbs, out, batch, l, special, workspace);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!d2) return CKKS_ERR_NULL;
    if (set_device(c)) return CKKS_ERR_HIP;
    const DevTables t = c->tables();
    const LimbMap m = limb_map(0, l);
    const ModDownConst mc = moddown_const(c, l, special);
    uint64_t* acc = (uint64_t*)workspace;
    const size_t LN = (size_t)l * c->N;
    TIMED(CKKS_KIND_KEYSWITCH, launch_keyswitch(t, d2, LN, nullptr, nullptr, key, key_limbs, acc,
                                                batch, l, m, special, c->ks_threads, S(stream)));
    TIMED(CKKS_KIND_MODDOWN, launch_moddown_rescale(t, acc, d01, 3, out, batch, l, m, special, mc, false, S(stream)));
    return CKKS_OK;
}

int ckks_hmult(const ckks_ctx* c, const uint64_t* a, const uint64_t* b, const uint64_t* key, int key_limbs,
               uint64_t* out, int batch, int l, int special, void* workspace, void* stream) {
    int st = ks_common_check(c, key, key_limbs, out, batch, l, special, workspace);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (l < 2) return CKKS_ERR_BAD_SHAPE;
    if (!a || !b) return CKKS_ERR_NULL;
    if (misaligned(a) || misaligned(b)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    const DevTables t = c->tables();
    const LimbMap m = limb_map(0, l);
    const ModDownConst mc = moddown_const(c, l, special);
    const size_t LN = (size_t)l * c->N, B = (size_t)batch;
    uint64_t* ea = (uint64_t*)workspace;
    uint64_t* eb = ea + B * 2 * LN;
    uint64_t* d01 = eb + B * 2 * LN;
    uint64_t* d2e = d01 + B * 2 * LN;
    uint64_t* d2c = d2e + B * LN;
    uint64_t* acc = d2c + B * LN;
    if (c->front_fused) {
        TIMED(CKKS_KIND_FRONT, launch_front(t, a, b, d01, d2e, d2c, batch, l, m, mc, c->tensor_threads, S(stream)));
    } else {
        // a squaring (b == a) transforms its operand once
        TIMED(CKKS_KIND_NTT, launch_ntt(t, ea, a, batch * 2, l, m, false, false, c->ntt_threads, S(stream)));
        if (b != a) TIMED(CKKS_KIND_NTT, launch_ntt(t, eb, b, batch * 2, l, m, false, false, c->ntt_threads, S(stream)));
        TIMED(CKKS_KIND_TENSOR, launch_tensor_intt(t, ea, b != a ? eb : ea, d01, d2e, d2c, batch, l, m, mc,
                                                   c->tensor_threads, S(stream)));
    }
    // the key switch of limb l-1 and of the special prime first, then the
    // other chain limbs with ModDown and the rescale fused in
    KsTargets tail, rest;
    tail.first = l - 1;
    tail.count = 2;
    tail.pair_major = c->ks_pair_major;
    rest.first = 0;
    rest.count = l - 1;
    rest.out = out;
    rest.mc = &mc;
    rest.pair_major = c->ks_pair_major;
    TIMED(CKKS_KIND_KEYSWITCH, launch_keyswitch(t, d2c, LN, d2e, d01, key, key_limbs, acc, batch,
                                                l, m, special, c->ks_threads, S(stream), nullptr, tail));
    TIMED(CKKS_KIND_KEYSWITCH_FUSED, launch_keyswitch(t, d2c, LN, d2e, d01, key, key_limbs, acc,
                                                      batch, l, m, special, c->ks_threads, S(stream), nullptr, rest));
    return CKKS_OK;
}

int ckks_hmult_noalloc(const ckks_ctx* c, const uint64_t* a, const uint64_t* b, const uint64_t* key, int key_limbs,
                       uint64_t* out, int batch, int l, int special, void* stream) {
    if (!c) return CKKS_ERR_NULL;
    if (batch < 0 || l < 1) return CKKS_ERR_BAD_SHAPE;
    if (ckks_workspace_bytes(c, batch, l) > CKKS_NOALLOC_SCRATCH_BYTES) return CKKS_ERR_BAD_SHAPE;
    if (!c->scratch) return CKKS_ERR_HIP;
    // the same kernels as ckks_hmult, with the device's scratch array as the
    // workspace; nothing in that path allocates, and no part of the
    // workspace needs to start cleared (every word is written before it is read)
    return ckks_hmult(c, a, b, key, key_limbs, out, batch, l, special, c->scratch, stream);
}

int ckks_automorphism(const ckks_ctx* c, const uint64_t* in, uint64_t* out, int n_polys, int n_limbs, int prime0,
                      uint64_t galois, void* stream) {
    int st = check_rows(c, n_polys, n_limbs, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!in || !out) return CKKS_ERR_NULL;
    if (in == out) return CKKS_ERR_BAD_ARG;
    if (!(galois & 1) || galois >= 2 * c->N) return CKKS_ERR_BAD_ARG;
    if (c->flags & CKKS_CYCLIC) return CKKS_ERR_BAD_ARG;
    if (misaligned(in) || misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_automorphism(c->tables(), in, out, n_polys, n_limbs, limb_map(prime0, n_limbs), galois, S(stream)));
    return CKKS_OK;
}

namespace {
// ckks_rotate, and with add = true ckks_rotate_add (the input added in the
// ModDown pass: out = ct + rot(ct)).
int rotate_impl(const ckks_ctx* c, const uint64_t* ct, const uint64_t* gkey, int key_limbs, uint64_t* out,
                int batch, int l, int special, uint64_t galois, void* workspace, void* stream, bool add) {
    int st = ks_common_check(c, gkey, key_limbs, out, batch, l, special, workspace);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!ct) return CKKS_ERR_NULL;
    if (ct == out) return CKKS_ERR_BAD_ARG;
    if (!(galois & 1) || galois >= 2 * c->N) return CKKS_ERR_BAD_ARG;
    if (misaligned(ct)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    const DevTables t = c->tables();
    const LimbMap m = limb_map(0, l);
    const ModDownConst mc = moddown_const(c, l, special);
    const size_t LN = (size_t)l * c->N, B = (size_t)batch;
    uint64_t* sct = (uint64_t*)workspace;   // sigma(c0), sigma(c1)
    uint64_t* acc = sct + B * 2 * LN;
    TIMED(CKKS_KIND_AUTOMORPHISM, launch_automorphism(t, ct, sct, batch * 2, l, m, galois, S(stream)));
    TIMED(CKKS_KIND_KEYSWITCH, launch_keyswitch(t, sct + LN, 2 * LN, nullptr, nullptr, gkey,
                                                key_limbs, acc, batch, l, m, special, c->ks_threads, S(stream)));
    TIMED(CKKS_KIND_MODDOWN, launch_moddown_rescale(t, acc, sct, 1, out, batch, l, m, special, mc, false, S(stream),
                                                    add ? ct : nullptr));
    return CKKS_OK;
}
}  // namespace

int ckks_rotate(const ckks_ctx* c, const uint64_t* ct, const uint64_t* gkey, int key_limbs, uint64_t* out, int batch,
                int l, int special, uint64_t galois, void* workspace, void* stream) {
    return rotate_impl(c, ct, gkey, key_limbs, out, batch, l, special, galois, workspace, stream, false);
}

int ckks_rotate_add(const ckks_ctx* c, const uint64_t* ct, const uint64_t* gkey, int key_limbs, uint64_t* out,
                    int batch, int l, int special, uint64_t galois, void* workspace, void* stream) {
    return rotate_impl(c, ct, gkey, key_limbs, out, batch, l, special, galois, workspace, stream, true);
}

int ckks_rotate_many(const ckks_ctx* c, const uint64_t* ct, const uint64_t* key_table, int key_limbs,
                     const uint64_t* galois_table, uint64_t* out, int batch, int l, int special, void* workspace,
                     void* stream) {
    int st = ks_common_check(c, key_table, key_limbs, out, batch, l, special, workspace);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!ct || !galois_table) return CKKS_ERR_NULL;
    if (ct == out) return CKKS_ERR_BAD_ARG;
    if (misaligned(ct) || misaligned(galois_table)) return CKKS_ERR_BAD_ALIGNMENT;
    if (c->N < 128 || c->N > 16384) return CKKS_ERR_BAD_N;   // the per-ciphertext map runs in LDS
    if (set_device(c)) return CKKS_ERR_HIP;
    const DevTables t = c->tables();
    const LimbMap m = limb_map(0, l);
    const ModDownConst mc = moddown_const(c, l, special);
    const size_t LN = (size_t)l * c->N, B = (size_t)batch;
    uint64_t* sct = (uint64_t*)workspace;   // sigma_b(c0), sigma_b(c1) per ciphertext b
    uint64_t* acc = sct + B * 2 * LN;
    TIMED(CKKS_KIND_AUTOMORPHISM, launch_automorphism_groups(t, ct, sct, batch * 2, l, (size_t)l * c->N, m, 1,
                                                             S(stream), galois_table, 2));
    TIMED(CKKS_KIND_KEYSWITCH, launch_keyswitch(t, sct + LN, 2 * LN, nullptr, nullptr, nullptr, key_limbs, acc, batch,
                                                l, m, special, c->ks_threads, S(stream), key_table));
    TIMED(CKKS_KIND_MODDOWN, launch_moddown_rescale(t, acc, sct, 1, out, batch, l, m, special, mc, false, S(stream)));
    return CKKS_OK;
}

uint64_t ckks_workspace_hoisted_bytes(const ckks_ctx* c, int batch, int l, int R) {
    if (!c || batch < 0 || l < 1 || R < 1) return 0;
    return (uint64_t)batch * c->N * 8 * (uint64_t)(l * (l + 1) + R * (4 * l + 2));
}

int ckks_key_permute(const ckks_ctx* c, const uint64_t* key_in, uint64_t* key_out, int rows, uint64_t galois,
                     void* stream) {
    if (!c || !key_in || !key_out) return CKKS_ERR_NULL;
    if (rows < 0) return CKKS_ERR_BAD_SHAPE;
    if (!(galois & 1) || galois >= 2 * c->N || (c->flags & CKKS_CYCLIC)) return CKKS_ERR_BAD_ARG;
    if (key_in == key_out) return CKKS_ERR_BAD_ARG;
    if (misaligned_pairs(key_in) || misaligned_pairs(key_out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_key_permute(c->tables(), key_in, key_out, rows, (uint32_t)galois, S(stream)));
    return CKKS_OK;
}

int ckks_rotate_hoisted(const ckks_ctx* c, const uint64_t* ct, const uint64_t* keys, int key_limbs,
                        const uint64_t* galois, int R, uint64_t* out, int batch, int l, int special, void* workspace,
                        void* stream) {
    int st = ks_common_check(c, keys, key_limbs, out, batch, l, special, workspace);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!ct || !galois) return CKKS_ERR_NULL;
    if (R < 1 || R > HOIST_MAX) return CKKS_ERR_BAD_SHAPE;
    if (misaligned(ct)) return CKKS_ERR_BAD_ALIGNMENT;
    GaloisList gl = {};
    for (int r = 0; r < R; ++r) {
        if (!(galois[r] & 1) || galois[r] >= 2 * c->N) return CKKS_ERR_BAD_ARG;
        gl.g[r] = (uint32_t)galois[r];
    }
    if (set_device(c)) return CKKS_ERR_HIP;
    const DevTables t = c->tables();
    const LimbMap m = limb_map(0, l);
    const ModDownConst mc = moddown_const(c, l, special);
    const size_t LN = (size_t)l * c->N, B = (size_t)batch;
    uint64_t* D = (uint64_t*)workspace;                       // [B][l][l+1][N]
    uint64_t* sct = D + B * l * (l + 1) * c->N;               // [R][B][2][l][N]: sigma_g(ct)
    uint64_t* acc = sct + (size_t)R * B * 2 * LN;             // [R][B][2][l+1][N]
    TIMED(CKKS_KIND_KEYSWITCH, launch_modup(t, ct, D, batch, l, m, special, c->ks_threads, S(stream)));
    // only sigma_g(c0) is read (ModDown adds poly 0 alone); sigma_g(c1)'s part of sct stays unwritten
    for (int r = 0; r < R; ++r)
        TIMED(CKKS_KIND_AUTOMORPHISM, launch_automorphism_groups(t, ct, sct + (size_t)r * B * 2 * LN, batch, l,
                                                                 2 * LN, m, galois[r], S(stream)));
    // one prepared key per rotation: [dnum = key_limbs - 1][2][key_limbs][N] pairs
    const size_t key_stride = (size_t)(key_limbs - 1) * 2 * key_limbs * c->N;
    TIMED(CKKS_KIND_KEYSWITCH, launch_ks_hoisted(t, D, keys, key_stride, key_limbs, gl, acc, batch,
                                                 l, m, special, R, 0, c->hoist, c->ks_threads, S(stream)));
    TIMED(CKKS_KIND_MODDOWN, launch_moddown_rescale(t, acc, sct, 1, out, batch * R, l, m, special, mc, false,
                                                    S(stream)));
    return CKKS_OK;
}

int ckks_rotate_hoisted_eval(const ckks_ctx* c, const uint64_t* ct, const uint64_t* keys, int key_limbs,
                             const uint64_t* galois, int R, uint64_t* out, int batch, int l, int special,
                             void* workspace, void* stream) {
    int st = ks_common_check(c, keys, key_limbs, out, batch, l, special, workspace);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!ct || !galois) return CKKS_ERR_NULL;
    if (R < 1 || R > HOIST_MAX) return CKKS_ERR_BAD_SHAPE;
    if (misaligned(ct)) return CKKS_ERR_BAD_ALIGNMENT;
    GaloisList gl = {};
    for (int r = 0; r < R; ++r) {
        if (!(galois[r] & 1) || galois[r] >= 2 * c->N) return CKKS_ERR_BAD_ARG;
        gl.g[r] = (uint32_t)galois[r];
    }
    if (set_device(c)) return CKKS_ERR_HIP;
    const DevTables t = c->tables();
    const LimbMap m = limb_map(0, l);
    const ModDownConst mc = moddown_const(c, l, special);
    const size_t LN = (size_t)l * c->N, B = (size_t)batch;
    uint64_t* D = (uint64_t*)workspace;                       // [B][l][l+1][N]
    uint64_t* cn = D + B * l * (l + 1) * c->N;                // [B][2][l][N]: ct in evaluation form
    uint64_t* acc = cn + B * 2 * LN;                          // [R][B][2][l+1][N]
    TIMED(CKKS_KIND_KEYSWITCH, launch_modup(t, ct, D, batch, l, m, special, c->ks_threads, S(stream)));
    TIMED(CKKS_KIND_NTT, launch_ntt(t, cn, ct, batch * 2, l, m, false, false, c->ntt_threads, S(stream)));
    const size_t key_stride = (size_t)(key_limbs - 1) * 2 * key_limbs * c->N;
    TIMED(CKKS_KIND_KEYSWITCH, launch_ks_hoisted(t, D, keys, key_stride, key_limbs, gl, acc, batch,
                                                 l, m, special, R, 1, c->hoist, c->ks_threads, S(stream)));
    TIMED(CKKS_KIND_MODDOWN, launch_moddown_eval(t, acc, cn, out, batch, l, m, special, mc, gl, R, c->ks_threads,
                                                 S(stream)));
    return CKKS_OK;
}

int ckks_encode(const ckks_ctx* c, const double* slots, uint64_t* pt, int n_vec, int l, int prime0, double scale,
                double* work, void* stream) {
    int st = check_rows(c, n_vec, l, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!slots || !pt || !work) return CKKS_ERR_NULL;
    if (!c->fft_tw) return CKKS_ERR_BAD_ARG;
    if (!(scale > 0) || !(scale < 9.0e18)) return CKKS_ERR_BAD_ARG;
    if (misaligned(slots) || misaligned(pt) || misaligned(work)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_encode_fft(c->fft_tw, c->fft_zeta, c->slot_k, (const double2*)slots, work, n_vec, c->logN,
                             S(stream)));
    HIPCHK(launch_encode_round(c->tables(), work, pt, n_vec, l, limb_map(prime0, l), scale, S(stream)));
    return CKKS_OK;
}

int ckks_decode(const ckks_ctx* c, const uint64_t* pt, double* slots, int n_vec, int l, int prime0, double scale,
                double* work, void* stream) {
    int st = check_rows(c, n_vec, l, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!slots || !pt || !work) return CKKS_ERR_NULL;
    if (!c->fft_tw) return CKKS_ERR_BAD_ARG;
    if (!(scale > 0)) return CKKS_ERR_BAD_ARG;
    if (misaligned(slots) || misaligned(pt) || misaligned(work)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    CrtConst cc;
    st = crt_const(c, prime0, l, &cc);
    if (st) return st;
    HIPCHK(launch_decode_crt(c->tables(), pt, work, n_vec, limb_map(prime0, l), cc, 1.0 / scale, S(stream)));
    HIPCHK(launch_decode_fft(c->fft_tw, c->fft_zeta, c->slot_k, work, (double2*)slots, n_vec, c->logN, S(stream)));
    return CKKS_OK;
}

int ckks_bitrev(uint64_t* data, int n_rows, uint64_t N, void* stream) {
    if (N < 1 || N > (1u << 20) || (N & (N - 1))) return CKKS_ERR_BAD_N;
    if (n_rows < 0) return CKKS_ERR_BAD_SHAPE;
    if (n_rows == 0) return CKKS_OK;
    if (!data) return CKKS_ERR_NULL;
    if (misaligned(data)) return CKKS_ERR_BAD_ALIGNMENT;
    HIPCHK(launch_bitrev(data, n_rows, bitlen(N) - 1, S(stream)));
    return CKKS_OK;
}

int ckks_mod_switch(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t q_in, uint64_t q_out, void* stream) {
    if (n == 0) return CKKS_OK;
    if (!in || !out) return CKKS_ERR_NULL;
    if (misaligned(in) || misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (q_in < 3 || q_out < 2 || q_in >= (1ull << 45) || bitlen(q_out) > bitlen(q_in)) return CKKS_ERR_BAD_MODULUS;
    const int ell = bitlen(q_in);
    const uint64_t mu = (uint64_t)(((u128)1 << (2 * ell)) / q_in);
    HIPCHK(launch_mod_switch(in, out, n, q_in, q_out, mu, ell, S(stream)));
    return CKKS_OK;
}

int ckks_mod_switch_scaled(const uint64_t* in, uint64_t* out, uint64_t n, uint64_t q_in, uint64_t q_out,
                           int scale_bits, void* stream) {
    if (n == 0) return CKKS_OK;
    if (!in || !out) return CKKS_ERR_NULL;
    if (misaligned(in) || misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (q_in < 2 || q_out < 2 || q_in >= (1ull << 45) || q_out >= (1ull << 45)) return CKKS_ERR_BAD_MODULUS;
    if (scale_bits < 0 || scale_bits > 127) return CKKS_ERR_BAD_ARG;
    HIPCHK(launch_mod_switch_toy(in, out, n, q_in, q_out, scale_bits, S(stream)));
    return CKKS_OK;
}

int ckks_poly_mul_modq(const ckks_ctx* c, const uint64_t* a, const uint64_t* b, uint64_t* out, int n_polys,
                       uint64_t q, void* workspace, void* stream) {
    int st = check_rows(c, n_polys, 3, 0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (q < 2 || q >= (1ull << 45)) return CKKS_ERR_BAD_MODULUS;
    // every product coefficient lies in (-N q^2, N q^2); Garner's sign test
    // needs it inside (-P/4, P/4), P = p0 p1 p2
    const int pbits = bitlen(c->primes[0]) + bitlen(c->primes[1]) + bitlen(c->primes[2]) - 3;
    if (pbits < 2 * bitlen(q - 1) + c->logN + 3) return CKKS_ERR_BAD_MODULUS;
    if (!a || !b || !out || !workspace) return CKKS_ERR_NULL;
    if (misaligned(a) || misaligned(b) || misaligned(out) || misaligned(workspace)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    const uint64_t p0 = c->primes[0], p1 = c->primes[1], p2 = c->primes[2];
    Crt3Const cc;
    memset(&cc, 0, sizeof(cc));
    cc.p0inv_p1 = invmod(p0 % p1, p1);
    cc.p0inv_p1_s = shoup(cc.p0inv_p1, p1);
    cc.p0_p2 = p0 % p2;
    cc.p0_p2_s = shoup(cc.p0_p2, p2);
    cc.p01inv_p2 = invmod(mulmod(p0 % p2, p1 % p2, p2), p2);
    cc.p01inv_p2_s = shoup(cc.p01inv_p2, p2);
    cc.p0_q = p0 % q;
    cc.p01_q = mulmod(cc.p0_q, p1 % q, q);
    cc.P_q = mulmod(cc.p01_q, p2 % q, q);
    PrimeConst& Q = cc.Pq;
    Q.q = q;
    Q.two_q = 2 * q;
    Q.ell = bitlen(q);
    Q.mu = (uint64_t)(((u128)1 << (2 * Q.ell)) / q);
    Q.mu64 = (uint64_t)(~0ull / q);
    const DevTables t = c->tables();
    const size_t rows3 = (size_t)n_polys * 3 * c->N;
    uint64_t* A3 = (uint64_t*)workspace;
    uint64_t* B3 = A3 + rows3;
    uint64_t* C3 = B3 + rows3;
    HIPCHK(launch_lift3(t, a, A3, n_polys, Q, S(stream)));
    HIPCHK(launch_lift3(t, b, B3, n_polys, Q, S(stream)));
    TIMED(CKKS_KIND_POLYMUL, launch_polymul(t, A3, B3, C3, n_polys, 3, limb_map(0, 3), c->ntt_threads, S(stream)));
    HIPCHK(launch_crt3_modq(t, C3, out, n_polys, cc, S(stream)));
    return CKKS_OK;
}

int ckks_probe_dft64(const uint64_t* x, uint64_t* y, const void* tables, uint64_t q, uint64_t ws40, size_t count,
                     int method, int reps, void* stream) {
    if (!x || !y || !tables) return CKKS_ERR_NULL;
    if (count == 0 || count % 64 || count > (1ull << 26)) return CKKS_ERR_BAD_SHAPE;
    if (method < 0 || method > 3 || reps < 1 || reps > 64) return CKKS_ERR_BAD_ARG;
    // q < 2^45 keeps the radix-2 path's lazy values below 2^62 for 64 reps
    // (q + 24 reps q) and the matrix path's below 2^48
    if (q < 3 || !(q & 1) || q >= (1ull << 45)) return CKKS_ERR_BAD_MODULUS;
    if (misaligned(x) || misaligned(y) || misaligned_pairs(tables)) return CKKS_ERR_BAD_ALIGNMENT;
    HIPCHK(launch_dft64_probe(x, y, tables, q, ws40, count, method, reps, S(stream)));
    return CKKS_OK;
}

int ckks_noise_stats_q(const uint64_t* data, uint64_t* stats, int rows, uint64_t n, uint64_t q, void* stream) {
    if (rows < 0 || n > (1ull << 30)) return CKKS_ERR_BAD_SHAPE;
    if (rows == 0 || n == 0) return CKKS_OK;
    if (q < 2 || q >= (1ull << 45)) return CKKS_ERR_BAD_MODULUS;
    if (!data || !stats) return CKKS_ERR_NULL;
    if (misaligned(data) || misaligned(stats)) return CKKS_ERR_BAD_ALIGNMENT;
    HIPCHK(launch_noise_stats_q(data, stats, rows, n, q, S(stream)));
    return CKKS_OK;
}

int ckks_embed(const ckks_ctx* c, const double* coeffs, double* slots, int n_vec, int order, void* stream) {
    if (!c || !coeffs || !slots) return CKKS_ERR_NULL;
    if (!c->fft_tw) return CKKS_ERR_BAD_ARG;
    if (n_vec < 0 || order < 0 || order > 1) return CKKS_ERR_BAD_ARG;
    if (misaligned(coeffs) || misaligned(slots)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_decode_fft(c->fft_tw, c->fft_zeta, order ? nullptr : c->slot_k, coeffs, (double2*)slots, n_vec,
                             c->logN, S(stream)));
    return CKKS_OK;
}

int ckks_embed_inv(const ckks_ctx* c, const double* slots, double* coeffs, int n_vec, int order, void* stream) {
    if (!c || !coeffs || !slots) return CKKS_ERR_NULL;
    if (!c->fft_tw) return CKKS_ERR_BAD_ARG;
    if (n_vec < 0 || order < 0 || order > 1) return CKKS_ERR_BAD_ARG;
    if (misaligned(coeffs) || misaligned(slots)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_encode_fft(c->fft_tw, c->fft_zeta, order ? nullptr : c->slot_k, (const double2*)slots, coeffs,
                             n_vec, c->logN, S(stream)));
    return CKKS_OK;
}

int ckks_pt_prepare(const ckks_ctx* c, uint64_t* pt, uint64_t* out, int n_polys, int l, int prime0, void* stream) {
    int st = check_rows(c, n_polys, l, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!pt || !out) return CKKS_ERR_NULL;
    if (misaligned(pt) || misaligned_pairs(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (c->flags & CKKS_CYCLIC) return CKKS_ERR_BAD_ARG;
    if (set_device(c)) return CKKS_ERR_HIP;
    const DevTables t = c->tables();
    const LimbMap m = limb_map(prime0, l);
    HIPCHK(launch_ntt(t, pt, pt, n_polys, l, m, false, true, c->ntt_threads, S(stream)));
    HIPCHK(launch_prepare_pairs(t, pt, out, n_polys, l, m, S(stream)));
    return CKKS_OK;
}

int ckks_lincomb(const ckks_ctx* c, const uint64_t* E, const uint64_t* D, uint64_t* out, int nb, int ng, int batch,
                 int l, void* stream) {
    if (!c) return CKKS_ERR_NULL;
    if (nb < 1 || nb > 32 || ng < 0 || batch < 0 || l < 1 || l > c->K) return CKKS_ERR_BAD_SHAPE;
    if (!batch || !ng) return CKKS_OK;
    if (!E || !D || !out) return CKKS_ERR_NULL;
    if (misaligned(E) || misaligned_pairs(D) || misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_lincomb(c->tables(), E, D, out, nb, ng, batch, l, limb_map(0, l), 0, S(stream)));
    return CKKS_OK;
}

int ckks_lincomb_scalar(const ckks_ctx* c, const uint64_t* E, const uint64_t* D, uint64_t* out, int nb, int ng,
                        int batch, int l, void* stream) {
    if (!c) return CKKS_ERR_NULL;
    if (nb < 1 || nb > 32 || ng < 0 || batch < 0 || l < 1 || l > c->K) return CKKS_ERR_BAD_SHAPE;
    if (!batch || !ng) return CKKS_OK;
    if (!E || !D || !out) return CKKS_ERR_NULL;
    if (misaligned(E) || misaligned_pairs(D) || misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_lincomb(c->tables(), E, D, out, nb, ng, batch, l, limb_map(0, l), 1, S(stream)));
    return CKKS_OK;
}

int ckks_lift_small(const ckks_ctx* c, const uint64_t* in, uint64_t* out, int rows, int l_in, int l_out,
                    void* stream) {
    int st = check_rows(c, rows, l_out, 0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (l_in < 1 || l_in > c->K) return CKKS_ERR_BAD_SHAPE;
    if (!in || !out) return CKKS_ERR_NULL;
    if (misaligned(in) || misaligned(out)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_lift_small(c->tables(), in, out, rows, l_in, l_out, limb_map(0, l_out), S(stream)));
    return CKKS_OK;
}

int ckks_noise_stats(const ckks_ctx* c, const uint64_t* data, uint64_t* stats, int n_polys, int n_limbs, int prime0,
                     void* stream) {
    int st = check_rows(c, n_polys, n_limbs, prime0);
    if (st) return st == NOTHING_TO_DO ? CKKS_OK : st;
    if (!data || !stats) return CKKS_ERR_NULL;
    if (misaligned(data) || misaligned(stats)) return CKKS_ERR_BAD_ALIGNMENT;
    if (set_device(c)) return CKKS_ERR_HIP;
    HIPCHK(launch_noise_stats(c->tables(), data, stats, n_polys, n_limbs, limb_map(prime0, n_limbs), S(stream)));
    return CKKS_OK;
}

}  // extern "C"
