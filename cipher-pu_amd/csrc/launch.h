// Internal launcher interface shared by the kernel translation units and
// the C ABI (capi.cpp).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "modarith.h"
#include "modf64.h"

// A value the compiler cannot prove wave-uniform (a runtime division of the
// block index, say), marked uniform: it then lives in an SGPR, and the
// table entries it selects come through scalar loads instead of a chain of
// vector loads at the start of every workgroup.
__device__ __forceinline__ int wave_uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

struct DevTables {
    const PrimeConst* pc;   // [K]
    const ulonglong2* tw;   // [K][2][N]: forward table, inverse table
    int N;
    int logN;
    // 1 when every prime is below twice every other one: a residue mod one
    // prime is then below 2q for any other q, which the lazy forward
    // transform accepts as input without a reduction.
    int narrow = 0;
    // The FP64 pipeline's tables (modf64.h): per-prime constants and the
    // twiddle words as doubles, in the order of tw.
    const PrimeF* pf = nullptr;   // [K]
    const double* twf = nullptr;  // [K][2][N]
    // Arithmetic of the transform kernels that have both forms: 0 = 64-bit
    // integer words, 1 = FP64 (results identical).
    int fp64 = 0;
    // Compute units of the device: the grid of kernels that loop over their
    // work items is sized from it.
    int n_cu = 256;
};

// Cumulative table of |e| for the rounded Gaussian: cdf[k] = floor(2^64
// P(|round(Z)| <= k)), saturated at 2^64 - 1.
#define GAUSS_TABLE 48
struct GaussTable {
    uint64_t cdf[GAUSS_TABLE];
};

// Thread-count choice for the LDS-resident kernels, per log2 N.
struct NttPlan {
    int threads;   // 0 = default
};

// ntt kernels (k_ntt.cu)
hipError_t launch_ntt(const DevTables& t, uint64_t* dst, const uint64_t* src, int n_polys, int n_limbs, const LimbMap& map,
                      bool inverse, bool reduce_input, int threads, hipStream_t s);
hipError_t launch_polymul(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* out,
                          int n_polys, int n_limbs, const LimbMap& map, int threads, hipStream_t s);
// Prepared pairs of reduced words, in the context's arithmetic: (value,
// RN(value / q)) doubles for FP64, (value, floor(value 2^63 / q)) for INT64;
// the form of keys and plaintexts the key switches and combinations read.
hipError_t launch_prepare_pairs(const DevTables& t, const uint64_t* in, void* out, int n_polys, int n_limbs,
                                const LimbMap& map, hipStream_t s);
// Prepared key rows: the words as doubles for FP64 (8 bytes a word), the
// prepared pairs above for INT64.
hipError_t launch_prepare_key(const DevTables& t, const uint64_t* in, void* out, int n_polys, int n_limbs,
                              const LimbMap& map, hipStream_t s);
hipError_t launch_bitrev(uint64_t* data, int n_rows, int logN, hipStream_t s);
hipError_t launch_pointwise(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* out,
                            int n_polys, int n_limbs, const LimbMap& map, int op, hipStream_t s);

// ckks kernels (k_ckks.cu)
struct ModDownConst {
    // ModDown by the special prime (index pidx_p): for limb i < l
    uint64_t hp[CKKS_MAX_PRIMES];      // (p-1)/2 mod q_i
    uint64_t pinv[CKKS_MAX_PRIMES];    // p^-1 mod q_i
    uint64_t pinv_s[CKKS_MAX_PRIMES];
    // rescale by the last limb q_(l-1): for limb i < l-1
    uint64_t hq[CKKS_MAX_PRIMES];
    uint64_t qinv[CKKS_MAX_PRIMES];
    uint64_t qinv_s[CKKS_MAX_PRIMES];
    uint64_t p_mod[CKKS_MAX_PRIMES];   // p mod q_i, and Shoup companion
    uint64_t p_mod_s[CKKS_MAX_PRIMES];
    // The same constants for the FP64 pipeline (modf64.h): values as doubles,
    // multipliers with their RN(w / q_i) companions.
    double pmf[CKKS_MAX_PRIMES], pmf_q[CKKS_MAX_PRIMES];         // p mod q_i
    double hpf[CKKS_MAX_PRIMES], hqf[CKKS_MAX_PRIMES];           // hp, hq
    double pinvf[CKKS_MAX_PRIMES], pinvf_q[CKKS_MAX_PRIMES];     // p^-1 mod q_i
    double qinvf[CKKS_MAX_PRIMES], qinvf_q[CKKS_MAX_PRIMES];     // q_(l-1)^-1 mod q_i
};

hipError_t launch_tensor_intt(const DevTables& t, const uint64_t* ea, const uint64_t* eb, uint64_t* d01,
                              uint64_t* d2e, uint64_t* d2c, int batch, int l, const LimbMap& map,
                              const ModDownConst& mc, int threads, hipStream_t s);
hipError_t launch_front(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* d01, uint64_t* d2e,
                        uint64_t* d2c, int batch, int l, const LimbMap& map, const ModDownConst& mc, int threads,
                        hipStream_t s);
// Which target limbs one key-switch launch covers: first .. first + count - 1
// (count 0: all l + 1).  With out set (HMult only; first + count <= l - 1),
// the workgroups finish ModDown and the rescale by the last chain prime
// themselves, from the accumulators of limb l - 1 and of the special prime
// that an earlier launch left in acc, and write [batch][2][l-1][N] to out.
// pair_major: workgroup placement -- 0 = target-major (all ciphertexts of
// one target limb in flight together: that limb's key slice stays in L2);
// 1 = ciphertext-major with each XCD given one contiguous range of
// ciphertexts, so a ciphertext's targets run back to back on one XCD and
// its digits (and the fused ModDown's inputs) are read from that XCD's L2.
struct KsTargets {
    int first = 0, count = 0;
    uint64_t* out = nullptr;
    const ModDownConst* mc = nullptr;
    int pair_major = 0;
};

// key_table (optional): ciphertext b uses the key at device address
// key_table[b] instead of key.
// key_pairs: prepared key rows in the context's arithmetic (doubles w for
// FP64, ulonglong2 (w, Shoup companion) pairs for INT64).
hipError_t launch_keyswitch(const DevTables& t, const uint64_t* d2c, size_t d2c_bstride, const uint64_t* d2e,
                            const uint64_t* d01, const void* key_pairs, int key_limbs, uint64_t* acc,
                            int batch, int l, const LimbMap& map, int pidx_special, int threads,
                            hipStream_t s, const uint64_t* key_table = nullptr, const KsTargets& tg = KsTargets());
// The integer-word versions (k_ckks_int.hip) the launchers above dispatch to
// when the context's arithmetic is INT64.
hipError_t launch_tensor_intt_int(const DevTables& t, const uint64_t* ea, const uint64_t* eb, uint64_t* d01,
                                  uint64_t* d2e, uint64_t* d2c, int batch, int l, const LimbMap& map,
                                  const ModDownConst& mc, int threads, hipStream_t s);
hipError_t launch_front_int(const DevTables& t, const uint64_t* a, const uint64_t* b, uint64_t* d01, uint64_t* d2e,
                            uint64_t* d2c, int batch, int l, const LimbMap& map, const ModDownConst& mc, int threads,
                            hipStream_t s);
hipError_t launch_keyswitch_int(const DevTables& t, const uint64_t* d2c, size_t d2c_bstride, const uint64_t* d2e,
                                const uint64_t* d01, const ulonglong2* key, int key_limbs, uint64_t* acc, int batch,
                                int l, const LimbMap& map, int pidx_p, int threads, hipStream_t s,
                                const uint64_t* key_table, const KsTargets& tg);
hipError_t launch_lincomb_int(const DevTables& t, const uint64_t* E, const ulonglong2* D, uint64_t* out, int nb,
                              int ng, int batch, int l, const LimbMap& map, int scalar, hipStream_t s);
// dcoef (rows selected by dmask bit poly) and dadd (both polys), optional, are
// added to the result; both are [batch][2][l][N].
hipError_t launch_moddown_rescale(const DevTables& t, const uint64_t* acc, const uint64_t* dcoef, int dmask,
                                  uint64_t* out, int batch, int l, const LimbMap& map, int pidx_special,
                                  const ModDownConst& mc, bool rescale, hipStream_t s,
                                  const uint64_t* dadd = nullptr);
hipError_t launch_rescale(const DevTables& t, const uint64_t* in, uint64_t* out, int n_polys, int l,
                          const LimbMap& map, const ModDownConst& mc, hipStream_t s);
hipError_t launch_automorphism(const DevTables& t, const uint64_t* in, uint64_t* out, int n_polys,
                               int n_limbs, const LimbMap& map, uint64_t galois, hipStream_t s);
// The automorphism over n_groups groups of n_limbs rows, group g at g group_stride.
// gal_table (optional, device): group g uses gal_table[g / gal_div] instead
// of galois; needs N in [128, 16384].
hipError_t launch_automorphism_groups(const DevTables& t, const uint64_t* in, uint64_t* out, int n_groups,
                                      int n_limbs, size_t group_stride, const LimbMap& map, uint64_t galois,
                                      hipStream_t s, const uint64_t* gal_table = nullptr, int gal_div = 1);
hipError_t launch_mod_switch(const uint64_t* in, uint64_t* out, size_t n, uint64_t q_in, uint64_t q_out, uint64_t mu,
                             int ell, hipStream_t s);
hipError_t launch_lincomb(const DevTables& t, const uint64_t* E, const void* D_pairs, uint64_t* out, int nb, int ng,
                          int batch, int l, const LimbMap& map, int scalar, hipStream_t s);
hipError_t launch_lift_small(const DevTables& t, const uint64_t* in, uint64_t* out, int rows, int l_in, int l_out,
                             const LimbMap& map, hipStream_t s);
hipError_t launch_noise_stats(const DevTables& t, const uint64_t* in, uint64_t* stats, int n_polys,
                              int n_limbs, const LimbMap& map, hipStream_t s);
// The monitor for rows of any length n under one modulus q (no context).
hipError_t launch_noise_stats_q(const uint64_t* in, uint64_t* stats, int rows, size_t n, uint64_t q, hipStream_t s);
// The reference's toy rescale floor((c q_out + floor(q_in/2)) / (q_in 2^sb)) mod q_out.
hipError_t launch_mod_switch_toy(const uint64_t* in, uint64_t* out, size_t n, uint64_t q_in, uint64_t q_out, int sb,
                                 hipStream_t s);
// Exact product modulo any q < 2^45 through the context's primes 0..2.
struct Crt3Const {
    uint64_t p0inv_p1, p0inv_p1_s;     // p0^-1 mod p1, Shoup companion
    uint64_t p0_p2, p0_p2_s;           // p0 mod p2
    uint64_t p01inv_p2, p01inv_p2_s;   // (p0 p1)^-1 mod p2
    uint64_t p0_q, p01_q, P_q;         // p0, p0 p1, p0 p1 p2 modulo q
    PrimeConst Pq;                     // q, mu, mu64, ell (Barrett / 64-bit reduction mod q)
};
hipError_t launch_lift3(const DevTables& t, const uint64_t* in, uint64_t* out, int rows, const PrimeConst& Pq,
                        hipStream_t s);
hipError_t launch_crt3_modq(const DevTables& t, const uint64_t* in, uint64_t* out, int rows, const Crt3Const& cc,
                            hipStream_t s);

// fft kernels (k_fft.hip)
#define CRT_MAX_WORDS 8
#define CRT_MAX_LIMBS 10
struct CrtConst {
    int l, words;
    uint64_t qhat[CRT_MAX_LIMBS][CRT_MAX_WORDS];   // Q / q_i, little-endian words
    uint64_t qhat_inv[CRT_MAX_LIMBS];              // (Q/q_i)^-1 mod q_i
    uint64_t qhat_inv_s[CRT_MAX_LIMBS];
    uint64_t Q[CRT_MAX_WORDS];
    uint64_t half_Q[CRT_MAX_WORDS];                // floor(Q / 2)
};

hipError_t launch_decode_fft(const double2* tw, const double2* zeta, const int* slot_k, const double* coeffs,
                             double2* slots, int n_vec, int logN, hipStream_t s);
hipError_t launch_encode_fft(const double2* tw, const double2* zeta, const int* slot_k, const double2* slots,
                             double* coeffs, int n_vec, int logN, hipStream_t s);
hipError_t launch_encode_round(const DevTables& t, const double* coeffs, uint64_t* pt, int n_vec, int l,
                               const LimbMap& map, double scale, hipStream_t s);
hipError_t launch_decode_crt(const DevTables& t, const uint64_t* pt, double* coeffs, int n_vec, const LimbMap& map,
                             const CrtConst& cc, double inv_scale, hipStream_t s);

hipError_t launch_sample_uniform(const DevTables& t, uint64_t* out, int rows, int l, const LimbMap& map,
                                 uint64_t seed, hipStream_t s);
// Fused symmetric encryption on the FP64 pipeline (N = 2^12..2^14; other
// rings and the integer arithmetic: hipErrorInvalidValue): pt [batch][l][N],
// s_eval [l][N] evaluation order, out [batch][2][l][N].
hipError_t launch_encrypt_sk(const DevTables& t, const uint64_t* pt, const uint64_t* s_eval, uint64_t* out, int batch,
                             int l, const LimbMap& map, uint64_t seed_a, uint64_t seed_e, const GaussTable& tab,
                             hipStream_t s);
// Fused decryption m = c0 + c1 s, same rings and arithmetic: ct
// [batch][2][l][N], s_eval [l][N], out [batch][l][N].
hipError_t launch_decrypt_sk(const DevTables& t, const uint64_t* ct, const uint64_t* s_eval, uint64_t* out, int batch,
                             int l, const LimbMap& map, hipStream_t s);
hipError_t launch_sample_gauss(const DevTables& t, uint64_t* out, int rows, int l, const LimbMap& map,
                               uint64_t seed, const GaussTable& tab, hipStream_t s);

// Galois elements of one hoisted-rotation call, passed by value.
#define HOIST_MAX 32
struct GaloisList {
    uint32_t g[HOIST_MAX];
};

// Workgroup placement of the hoisted key switch, one workgroup per (target
// t, rotation r, ciphertext b).  Each (r, t) key slice (1 MB at N = 8192,
// l = 4) is read by every ciphertext and each (b, t) digit slice (l rows) by
// every rotation, so the order decides what L2 keeps.  Logical order: t
// outermost; inside one t, tiles of btile ciphertexts, r-major and b-minor
// inside a tile.  btile = batch walks one key slice across all ciphertexts;
// btile = 1 walks one ciphertext's digits across all rotations.  With xcd
// set and the grid G a multiple of 8, the logical order is dealt out in 8
// contiguous ranges, one per XCD: the hardware hands workgroup i to XCD
// i mod 8, so workgroup i takes logical index (i mod 8) G/8 + i/8 and each
// XCD's L2 sees one contiguous stretch.  For 1 <= btile <= batch every
// placement is a bijection from [0, G) onto [0, l+1) x [0, R) x [0, batch)
// (tests/hoist_order_check.cpp).
struct HoistOrder {
    int btile;
    int xcd;
};
struct TRB {
    int t, r, b;
};
MODARITH_FN TRB hoist_slot(int i, int G, int R, int batch, HoistOrder o) {
    const int L = (o.xcd && (G & 7) == 0) ? (i & 7) * (G >> 3) + (i >> 3) : i;
    const int RB = R * batch;
    const int t = L / RB;
    const int rem = L - t * RB;
    const int tile = rem / (R * o.btile);
    const int b0 = tile * o.btile;
    const int nb = batch - b0 < o.btile ? batch - b0 : o.btile;
    const int inner = rem - tile * R * o.btile;
    const int r = inner / nb;
    return TRB{t, r, b0 + inner - r * nb};
}

hipError_t launch_modup(const DevTables& t, const uint64_t* ct, uint64_t* D, int batch, int l, const LimbMap& map,
                        int pidx_p, int threads, hipStream_t s);
// eval_out = 1: the chain-limb accumulators are stored in evaluation form
// (pi_g applied, no inverse transform) for launch_moddown_eval; the special
// limb's in coefficient form as always.
// order: workgroup placement (hoist_slot); btile 0 picks the measured
// default for the output form, btile above batch means batch.
hipError_t launch_ks_hoisted(const DevTables& t, const uint64_t* D, const void* keys, size_t key_stride,
                             int key_limbs, const GaloisList& galois, uint64_t* acc, int batch, int l,
                             const LimbMap& map, int pidx_p, int R, int eval_out, HoistOrder order, int threads,
                             hipStream_t s);
hipError_t launch_moddown_eval(const DevTables& t, const uint64_t* acc, const uint64_t* Cn, uint64_t* out, int batch,
                               int l, const LimbMap& map, int pidx_p, const ModDownConst& mc,
                               const GaloisList& galois, int R, int threads, hipStream_t s);
// rows of N prepared key entries (8-byte doubles for FP64, 16-byte pairs for
// INT64) permuted by pi_g^-1.
hipError_t launch_key_permute(const DevTables& t, const void* in, void* out, int rows, uint32_t g, hipStream_t s);


// Matrix-core probe (k_mfma_probe.hip): y = G^reps x for V length-64 vectors
// stored [64][V]; method 0 = radix-2 VALU stages, 1 / 2 = int8 MFMA product
// with accumulator layout candidate 0 / 1, 3 = candidate 0 software-pipelined.  tables: 64 twiddle pairs, row
// constants (128 words: kq, (-BIAS) mod q, BIAS), digit operands
// [2][4][6][64] x 16 bytes.
hipError_t launch_dft64_probe(const uint64_t* X, uint64_t* Y, const void* tables, uint64_t q, uint64_t ws40,
                              size_t V, int method, int reps, hipStream_t s);
