/* ckks_core.h -- C ABI of the MI355X (gfx950) CKKS evaluation core.
 *
 * The reference (YI-HAIBO/cipher-pu) has no FFI: its drop-in surface is the
 * set of Python golden-model functions on lists of integers, driven by
 * scripts/gen_ntt_vectors.py, verify_keyswitch.py, verify_rescale.py and
 * verify_e2e_mult.py.  Each entry point below names the function it
 * replaces.  The Python package cipherpu binds this ABI with ctypes and
 * keeps the reference's names.
 *
 * Conventions
 *  - All data pointers are caller-owned device pointers (hipMalloc or a
 *    torch tensor's data_ptr()), 8-byte aligned; buffers of prepared pairs
 *    -- prepared keys and prepared plaintexts, each word w < q stored as
 *    the two IEEE doubles (w, RN(w / q)) -- 16-byte aligned.  Nothing is
 *    allocated behind the caller's back except inside ckks_ctx_create.
 *  - Polynomials are uint64 rows of N coefficients; RNS tensors are
 *    [poly][limb][N] with limb i reduced modulo prime (prime0 + i) of the
 *    context, or [batch][2][limb][N] for ciphertexts (c0, c1), decrypt =
 *    c0 + c1*s (scripts/exp_e33_real_ciphertext.py:138-141).
 *  - "eval" data is in bit-reversed NTT order as produced by
 *    ckks_ntt_fwd; CKKS_NATURAL_ORDER selects the reference's natural order.
 *  - Every call is asynchronous on the given stream (NULL = default stream)
 *    and returns 0 or a negative ckks_status.  Contexts are read-only after
 *    creation, so compute calls may share one context from several host
 *    threads and streams; calls on one stream run in order.  The settings
 *    calls (ckks_ctx_set_threads, ckks_ctx_set_front) and the timing hook
 *    change the context: make them from one thread while no other call is
 *    using it.  Between ckks_timing_start and ckks_timing_stop the compute
 *    calls record their launches into the context, so a timed region is
 *    driven from one host thread.
 *  - Primes must be odd primes below 2^45 (the reference's 45-bit datapath,
 *    rtl/m1/ntt_unit.sv:38-51); 2N must divide q-1 for the negacyclic ring
 *    (N for the cyclic one).
 */
#ifndef CKKS_CORE_H
#define CKKS_CORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CKKS_OK = 0,
    CKKS_ERR_BAD_N = -1,          /* N not a power of two in [2, 16384] */
    CKKS_ERR_BAD_MODULUS = -2,    /* not an odd prime < 2^45, or no root of the needed order */
    CKKS_ERR_BAD_ALIGNMENT = -3,  /* pointer not 8-byte aligned (pair buffers: 16) */
    CKKS_ERR_BAD_SHAPE = -4,      /* limb range outside the context, batch < 0, ... */
    CKKS_ERR_HIP = -5,            /* a HIP runtime call failed */
    CKKS_ERR_NULL = -6,           /* required pointer is NULL */
    CKKS_ERR_BAD_ARG = -7,        /* other invalid argument */
    CKKS_ERR_NO_DEVICE = -8       /* no usable gfx950 device */
} ckks_status;

/* flags */
#define CKKS_CYCLIC 1          /* ring Z_q[X]/(X^N - 1) (reference ntt/intt); default negacyclic */
#define CKKS_NATURAL_ORDER 2   /* transforms read/write natural evaluation order */

typedef struct ckks_ctx ckks_ctx;

/* Create a context for ring degree N over the given primes (all limbs, then
 * any special primes).  roots[i] is the root of unity to use for prime i:
 * of order 2N (negacyclic) or N (cyclic); NULL derives g^((q-1)/order) from
 * the smallest generator g, the chain's construction
 * (scripts/gen_modulus_chain.py:128-130).  Precomputes the twiddle tables
 * (scripts/gen_omega_tables.py:96-106) with their Shoup companions, and
 * the FP64 pipeline's (w, RN(w / q)) twiddle pairs.
 * Replaces the module constants of tb/ref/gen_poly_mult_vectors_n12.py:36-44
 * and the chain loader verify_rescale.py:34-35. */
int ckks_ctx_create(ckks_ctx **out, uint64_t N, int n_primes, const uint64_t *primes,
                    const uint64_t *roots, int flags, int device);
int ckks_ctx_destroy(ckks_ctx *ctx);
/* Root of unity used for prime i (host value). */
int ckks_ctx_root(const ckks_ctx *ctx, int prime, uint64_t *root);
/* Library version string and the device architecture it was built for. */
const char *ckks_version(void);

/* In-place forward transform of n_polys x n_limbs rows (limb i uses prime
 * prime0 + i).  Replaces ntt(a, omega, N, q)
 * (tb/ref/gen_poly_mult_vectors_n12.py:68-82): with CKKS_NATURAL_ORDER and
 * a CKKS_CYCLIC context the result equals the reference's output. */
int ckks_ntt_fwd(const ckks_ctx *ctx, uint64_t *data, int n_polys, int n_limbs, int prime0, int flags,
                 void *stream);
/* In-place inverse transform; replaces intt(a, omega_inv, N, q)
 * (gen_poly_mult_vectors_n12.py:85-100). */
int ckks_ntt_inv(const ckks_ctx *ctx, uint64_t *data, int n_polys, int n_limbs, int prime0, int flags,
                 void *stream);

/* out = a * b in the ring, coefficient form, fused NTT -> point-wise ->
 * INTT per row.  Replaces poly_mult (scripts/verify_keyswitch.py:48-59)
 * and poly_mult_via_ntt (gen_poly_mult_vectors_n12.py:103-107). */
int ckks_poly_mul(const ckks_ctx *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out, int n_polys,
                  int n_limbs, int prime0, void *stream);

/* Element-wise RNS arithmetic, any domain: op 0 = a*b, 1 = a+b, 2 = a-b,
 * 3 = -a, 4 = a*b by special-form reduction (primes 2^44 + c, c < 2^21:
 * the fast_red of scripts/_verify_fastred.py:55-99 generalised; else
 * CKKS_ERR_BAD_MODULUS).  Replaces poly_add / poly_sub
 * (verify_keyswitch.py:40-45) and the point-wise product of
 * poly_mult_via_ntt. */
int ckks_pointwise(const ckks_ctx *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out, int n_polys,
                   int n_limbs, int prime0, int op, void *stream);

/* Rescale by the last limb: in [n_polys][l][N] -> out [n_polys][l-1][N],
 * coefficient form, out_i = [round-half-up(CRT(c) / q_(l-1))]_(q_i).
 * Replaces rescale_poly / rescale_coeff (scripts/verify_rescale.py:38-57)
 * applied to the CRT value. */
int ckks_rescale(const ckks_ctx *ctx, const uint64_t *in, uint64_t *out, int n_polys, int l, int prime0,
                 void *stream);

/* Bytes of device workspace needed by ckks_keyswitch / ckks_relin /
 * ckks_hmult / ckks_rotate for a batch at level l. */
uint64_t ckks_workspace_bytes(const ckks_ctx *ctx, int batch, int l);

/* Key material for the hybrid key switch, in the layout the kernels read,
 * evaluation order, in the context's arithmetic (ckks_ctx_set_arith):
 * CKKS_ARITH_FP64, [dnum][2][key_limbs][N] doubles (the words w; the key
 * products form RN(w / q) themselves), as many 8-byte words as the input;
 * CKKS_ARITH_INT64, [dnum][2][key_limbs][N] (w, floor(w 2^63 / q)) pairs,
 * twice the input's words.  A buffer of the pair size serves both.  Builds
 * it from coefficient-form b_j, a_j ([dnum][2][key_limbs][N] uint64, limbs
 * = primes 0..key_limbs-2 then the special prime).  Device pointers, key_out
 * 16-byte aligned.  key_coeff is transformed in place (it holds the
 * evaluation form afterwards).  A key is read in the arithmetic it was
 * prepared in. */
int ckks_key_prepare(const ckks_ctx *ctx, const uint64_t *key_coeff, uint64_t *key_out, int dnum,
                     int key_limbs, int special, void *stream);

/* Hybrid key switch (one digit per prime, one special prime) of d2
 * ([batch][l][N], coefficient form).  out: [batch][2][l][N] =
 * (ModDown(sum_j D_j b_j), ModDown(-sum_j D_j a_j)), then plus (d0, d1) if
 * given ([batch][2][l][N], coefficient form).  Generalises relin(e0, e1, e2,
 * a, b, q) (scripts/verify_keyswitch.py:62-68; key convention
 * exp_e33_real_ciphertext.py:153-175) to RNS; with d0/d1 it IS relin.
 * Residues must be reduced (limb i in [0, q_i)), as every entry point
 * writes them. */
int ckks_relin(const ckks_ctx *ctx, const uint64_t *d01, const uint64_t *d2, const uint64_t *key,
               int key_limbs, uint64_t *out, int batch, int l, int special, void *workspace, void *stream);

/* Fused HMult + relinearise + rescale of ciphertext pairs, coefficient form:
 * a, b [batch][2][l][N] -> out [batch][2][l-1][N].  Replaces e2e_mult
 * (scripts/verify_e2e_mult.py:78-102; RTL rtl/m4/e2e_ckks_mult_ctrl_m4.sv
 * S_MUL_0..S_RESCALE_1) with the exact RNS rescale.  With b == a (one
 * pointer: a squaring) the operand is transformed once; the words are those
 * of a product with a copy. */
int ckks_hmult(const ckks_ctx *ctx, const uint64_t *a, const uint64_t *b, const uint64_t *key, int key_limbs,
               uint64_t *out, int batch, int l, int special, void *workspace, void *stream);

/* ckks_hmult without a workspace argument, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * workspace is a library-owned device array of CKKS_NOALLOC_SCRATCH_BYTES
 * per GPU (held until the process ends); the call allocates nothing.
 * batch * ckks_workspace_bytes(ctx, 1, l) must fit in it, else
 * CKKS_ERR_BAD_SHAPE (at N = 8192, l = 4: at most 24 pairs).  Every call on
 * one GPU shares that GPU's array, so calls that may overlap on one GPU
 * (other streams, other host threads) race: issue them on one stream.
 * Results are word for word those of ckks_hmult. */
#define CKKS_NOALLOC_SCRATCH_BYTES (64ull << 20)
int ckks_hmult_noalloc(const ckks_ctx *ctx, const uint64_t *a, const uint64_t *b, const uint64_t *key,
                       int key_limbs, uint64_t *out, int batch, int l, int special, void *stream);

/* Galois automorphism f(X) -> f(X^g), g odd, coefficient form, out of
 * place.  Replaces automorphism_exact (exp_e24_rotation_bootstrap.py:143-162). */
int ckks_automorphism(const ckks_ctx *ctx, const uint64_t *in, uint64_t *out, int n_polys, int n_limbs,
                      int prime0, uint64_t galois, void *stream);

/* Slot rotation: automorphism by g = 5^r mod 2N then key switch from
 * sigma_g(s) back to s with the Galois key.  ct, out [batch][2][l][N].
 * Completes the EvalRotate that verify_multi_slot_simd.py:146-161 leaves
 * open (exp_e24:232-238 only multiplies by X^r). */
int ckks_rotate(const ckks_ctx *ctx, const uint64_t *ct, const uint64_t *gkey, int key_limbs, uint64_t *out,
                int batch, int l, int special, uint64_t galois, void *workspace, void *stream);

/* out = ct + rotate(ct): ckks_rotate with the input added in its ModDown
 * pass, the same words as ckks_rotate followed by ckks_pointwise op 1.  One
 * step of a rotate-and-sum (the slot sums of the encrypted layers, where
 * the reference sums coefficient-encoded products after decryption,
 * exp_e14_cnn_e2e.py:90-127).  Arguments as for ckks_rotate. */
int ckks_rotate_add(const ckks_ctx *ctx, const uint64_t *ct, const uint64_t *gkey, int key_limbs, uint64_t *out,
                    int batch, int l, int special, uint64_t galois, void *workspace, void *stream);

/* ckks_rotate with a step per ciphertext, in one call: ciphertext b is
 * rotated by galois_table[b] with the prepared Galois key at device address
 * key_table[b].  Both tables are device arrays of batch uint64 words; the
 * table pointer and every key it names must be 16-byte aligned (only the
 * table pointer is checked on the host).  Output word for word that of
 * ckks_rotate per ciphertext.  Same workspace as ckks_rotate; N in [128,
 * 16384].  No reference counterpart: it lets a batch of rotations by mixed
 * steps (one key each) fill the GPU as one launch per kernel. */
int ckks_rotate_many(const ckks_ctx *ctx, const uint64_t *ct, const uint64_t *key_table, int key_limbs,
                     const uint64_t *galois_table, uint64_t *out, int batch, int l, int special, void *workspace,
                     void *stream);

/* Hoisted rotations (no counterpart in the reference, which rotates one
 * step at a time: exp_e24_rotation_bootstrap.py:143-176; the method is
 * Halevi-Shoup, CRYPTO 2018).  R <= 32 rotations of the same ciphertexts
 * share one digit decomposition.  ct [batch][2][l][N]; keys: R prepared
 * Galois keys, each permuted for its element by ckks_key_permute, contiguous;
 * galois: host array of the R elements; out [R][batch][2][l][N].  The
 * digits are the automorphism of the canonical lifts, so outputs decrypt to
 * the rotated message like ckks_rotate's but are not bit-identical to them.
 * Workspace: ckks_workspace_hoisted_bytes(ctx, batch, l, R). */
uint64_t ckks_workspace_hoisted_bytes(const ckks_ctx *ctx, int batch, int l, int R);
int ckks_rotate_hoisted(const ckks_ctx *ctx, const uint64_t *ct, const uint64_t *keys, int key_limbs,
                        const uint64_t *galois, int R, uint64_t *out, int batch, int l, int special, void *workspace,
                        void *stream);
/* ckks_rotate_hoisted with the outputs in evaluation (bit-reversed NTT)
 * order, for callers whose next step is a plaintext product (no counterpart
 * in the reference).  ModDown runs in the evaluation domain: only the
 * special-prime limb is transformed back, so each rotated ciphertext costs
 * 2 inverse and 2 l forward transforms instead of 2 (l + 1) inverse ones
 * plus the caller's 2 l forward ones.  The words equal ckks_ntt_fwd of
 * ckks_rotate_hoisted's.  Same arguments and workspace. */
int ckks_rotate_hoisted_eval(const ckks_ctx *ctx, const uint64_t *ct, const uint64_t *keys, int key_limbs,
                             const uint64_t *galois, int R, uint64_t *out, int batch, int l, int special,
                             void *workspace, void *stream);
/* key_out = key_in (prepared key rows of N entries, in the context's
 * arithmetic as ckks_key_prepare wrote them) with every row permuted by the
 * inverse evaluation-domain automorphism of galois, for ckks_rotate_hoisted. */
int ckks_key_permute(const ckks_ctx *ctx, const uint64_t *key_in, uint64_t *key_out, int rows, uint64_t galois,
                     void *stream);

/* Canonical-embedding encode: slots [n_vec][N/2] complex (re, im doubles),
 * scaled by `scale`, rounded, reduced into l limbs -> pt [n_vec][l][N].
 * work: n_vec*N doubles of device scratch.  Replaces encode
 * (scripts/verify_multi_slot_simd.py:75-84) in the 5^j slot order. */
int ckks_encode(const ckks_ctx *ctx, const double *slots, uint64_t *pt, int n_vec, int l, int prime0,
                double scale, double *work, void *stream);
/* Decode: exact CRT of l limbs, centered, / scale, evaluated at the slots.
 * slots out [n_vec][N/2] complex.  Replaces decode
 * (verify_multi_slot_simd.py:87-91). */
int ckks_decode(const ckks_ctx *ctx, const uint64_t *pt, double *slots, int n_vec, int l, int prime0,
                double scale, double *work, void *stream);

/* In-place bit-reversal permutation of n_rows rows of N words
 * (bit_rev_perm, tb/ref/gen_poly_mult_vectors_n12.py:54-65). */
int ckks_bitrev(uint64_t *data, int n_rows, uint64_t N, void *stream);

/* Single-prime modulus switch out = round-half-up(in * q_out / q_in) mod
 * q_out for n values in [0, q_in).  Replaces rescale_coeff / rescale_poly
 * exactly (scripts/verify_rescale.py:38-57).  Needs bitlen(q_out) <=
 * bitlen(q_in) < 46. */
int ckks_mod_switch(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t q_in, uint64_t q_out, void *stream);

/* The reference's toy rescale: out = floor((in * q_out + floor(q_in / 2)) /
 * (q_in * 2^scale_bits)) mod q_out for n 64-bit values, exactly as
 * rescale(s, q_in, q_out, scale_bits) of scripts/verify_e2e_mult.py:70-75
 * (also scripts/gen_tb_e2e_vec.py:59-63, gen_tb_rescale_vec.py:28-33)
 * computes it on Python integers.  2 <= q_in, q_out < 2^45, 0 <= scale_bits
 * <= 127. */
int ckks_mod_switch_scaled(const uint64_t *in, uint64_t *out, uint64_t n, uint64_t q_in, uint64_t q_out,
                           int scale_bits, void *stream);

/* Product modulo ANY q (2 <= q < 2^45, prime or not, no root of unity
 * needed): a, b [n_polys][N] in [0, q) -> out [n_polys][N], in the
 * context's ring (negacyclic, or cyclic for a CKKS_CYCLIC context).  The
 * exact product over Z is formed from its residues modulo the context's
 * primes 0..2 (their product must exceed 8 N q^2, else
 * CKKS_ERR_BAD_MODULUS; the 45-bit chain primes serve any q < 2^45 up to
 * N = 16384) and reduced modulo q.  workspace: 9 n_polys N words.
 * Replaces the schoolbook poly_mult (scripts/verify_keyswitch.py:48-59,
 * verify_e2e_mult.py:56-67) and neg_conv_modq
 * (scripts/exp_e24_rotation_bootstrap.py:80-106), which accept every q,
 * and cyclic_conv_direct (tb/ref/gen_poly_mult_vectors_n12.py:110-117). */
int ckks_poly_mul_modq(const ckks_ctx *ctx, const uint64_t *a, const uint64_t *b, uint64_t *out, int n_polys,
                       uint64_t q, void *workspace, void *stream);

/* The plain noise monitor of ckks_noise_stats for rows of any length n
 * under one modulus q (2 <= q < 2^45), with no context: data [rows][n] in
 * [0, q) -> stats [rows][5].  The monitor needs only q, and the
 * reference's call site uses a toy modulus
 * (noise_monitor_energy(f0p + f1p, TOY_Q=113) on 32 coefficients,
 * scripts/verify_m4_h_multilayer.py:162-174,260). */
int ckks_noise_stats_q(const uint64_t *data, uint64_t *stats, int rows, uint64_t n, uint64_t q, void *stream);

/* Raw canonical embedding of real coefficients (doubles [n_vec][N]) to
 * complex slots [n_vec][N/2]: order 0 = slot j at zeta^(5^j), order 1 =
 * slot k at zeta^(4k+1).  ckks_embed_inv is its inverse (no rounding).
 * The float halves of encode/decode (verify_multi_slot_simd.py:32-52). */
int ckks_embed(const ckks_ctx *ctx, const double *coeffs, double *slots, int n_vec, int order, void *stream);
int ckks_embed_inv(const ckks_ctx *ctx, const double *slots, double *coeffs, int n_vec, int order, void *stream);

/* Plaintext rows for ckks_lincomb: coefficient-form pt [n_polys][l][N]
 * (transformed in place) -> out [n_polys][l][N] prepared pairs in
 * evaluation order, in the context's arithmetic: the doubles w and
 * RN(w / q) (FP64), or w and floor(w 2^63 / q) (INT64). */
int ckks_pt_prepare(const ckks_ctx *ctx, uint64_t *pt, uint64_t *out, int n_polys, int l, int prime0,
                    void *stream);

/* Batched plaintext linear combination in the evaluation domain, the inner
 * product of the diagonal matrix-vector method (encrypted_linear_layer,
 * scripts/exp_e11_ckks_e2e_real.py:142-166, re-packed as diagonals):
 *   out[g] = sum_(b<nb) E[b] * D[g][b]
 * E: [nb][batch][2][l][N] evaluation-form ciphertexts (limbs = primes
 * 0..l-1); D: [ng][nb][l][N] rows from ckks_pt_prepare; out:
 * [ng][batch][2][l][N].  nb <= 32. */
int ckks_lincomb(const ckks_ctx *ctx, const uint64_t *E, const uint64_t *D, uint64_t *out, int nb, int ng,
                 int batch, int l, void *stream);
/* The same with one constant per (g, b, limb): D is [ng][nb][l] prepared
 * pairs (the doubles w and RN(w / q)) of the residues w of an integer weight.  A constant
 * multiplies every coefficient alike, so E may be in either domain.  Used by
 * the encrypted convolution (the reference's encrypted_conv,
 * scripts/exp_e14_cnn_e2e.py:90-127, as rotations and scalar weights). */
int ckks_lincomb_scalar(const ckks_ctx *ctx, const uint64_t *E, const uint64_t *D, uint64_t *out, int nb, int ng,
                        int batch, int l, void *stream);

/* DARE lift (the message-keeping refresh of
 * exp_e24_rotation_bootstrap.py:287-304): rows of a decrypted plaintext
 * in [rows][l_in][N] -> [rows][l_out][N], taking the centred residue of
 * limb 0 as the integer message (exact while |m| < q_0 / 2). */
int ckks_lift_small(const ckks_ctx *ctx, const uint64_t *in, uint64_t *out, int rows, int l_in, int l_out,
                    void *stream);

/* Plain noise monitor (rtl/m4/noise_monitor.sv:44-78,
 * verify_m4_h_multilayer.py:162-174): per row, with c' = min(c, q-c),
 * stats[row][5] = {sum c', sum c'^2 low word, high word, max c',
 * sum of centered values}. */
int ckks_noise_stats(const ckks_ctx *ctx, const uint64_t *data, uint64_t *stats, int n_polys, int n_limbs,
                     int prime0, void *stream);

/* ---- learned gear controller (small dense network, FP32 weights,
 * activations optionally stored as BF16) ----
 * It predicts the output SNR of each sparsity gear from ciphertext
 * statistics; the plain noise monitor stays the fallback and the arbiter.  It
 * has no counterpart in the reference (whose gear choice is the threshold
 * policy of rtl/m4/adaptive_gear_ctrl.sv:42-100).  Activations [B][I] /
 * [B][J] row-major, weights [J][I]; *_bf16 selects BF16 storage. */
int ckks_gear_dense_fwd(const void *in, int in_bf16, const float *W, const float *bias, void *out, int out_bf16,
                        int B, int I, int J, int relu, void *stream);
/* Backward of one dense layer: dpre = dout * relu'(out); dW = dpre^T in;
 * db = column sums of dpre; din = dpre W (skipped if din is NULL). */
int ckks_gear_dense_bwd(const float *dout, const void *out, int out_bf16, const void *in, int in_bf16,
                        const float *W, float *dpre, float *dW, float *db, float *din, int B, int I, int J, int relu,
                        void *stream);
/* One Adam step on n parameters (step >= 1 for the bias corrections). */
int ckks_gear_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2, float eps,
                   int step, void *stream);
/* dy = 2 (y - target) / B and per-row squared error, for the MSE loss. */
int ckks_gear_mse_grad(const float *y, const float *target, float *dy, float *row_loss, int B, int J, void *stream);

/* Threads per workgroup (0 = default, else 128/256/512/1024 where built)
 * for this context's LDS-resident kernels: forward/inverse transforms and
 * products, the key switch, and the tensor-product kernel.  Set by the
 * tuner in cipherpu.tune. */
int ckks_ctx_set_threads(ckks_ctx *ctx, int ntt_threads, int ks_threads, int tensor_threads);
/* HMult front end: 1 = one fused kernel per (pair, limb) doing the four
 * forward transforms, the tensor product and INTT(d2); 0 (default) =
 * separate transform launches then the tensor kernel.  Same results. */
int ckks_ctx_set_front(ckks_ctx *ctx, int fused);
/* Residue arithmetic of every kernel on the path (transforms, products,
 * the HMult front, key switches, hoisted key switch, plaintext
 * combinations): CKKS_ARITH_INT64 (64-bit integer words, Shoup and Barrett
 * products on __umul64hi) or CKKS_ARITH_FP64 (default: residues held as
 * exact FP64 integers, products assembled from fused multiply-adds).
 * Results are identical word for word; words at rest are canonical uint64
 * residues in both.  Prepared keys and plaintexts (ckks_key_prepare,
 * ckks_pt_prepare) are in the form of the arithmetic in force when they
 * were prepared: prepare them after setting it.  No reference counterpart
 * (the RTL multiplier rtl/m1/mult_43x43_horner4.sv is one fixed datapath). */
#define CKKS_ARITH_INT64 0
#define CKKS_ARITH_FP64 1
int ckks_ctx_set_arith(ckks_ctx *ctx, int arith);
/* Workgroup placement of the hoisted key switch (ckks_rotate_hoisted and
 * its evaluation-form variant): inside one target limb, tiles of btile
 * ciphertexts with every rotation of a tile before the next tile (btile =
 * batch: one rotation's key slice across all ciphertexts); xcd = 1 gives
 * each of the 8 XCDs one contiguous stretch of that order.  btile = 0
 * (default) picks per call: tiles of 4 per XCD for coefficient-form
 * output, the whole batch for evaluation form.  Results do not depend on
 * it.  No reference counterpart (a launch-shape setting). */
int ckks_ctx_set_hoist_order(ckks_ctx *ctx, int btile, int xcd);
/* Workgroup placement of ckks_hmult's key switch: 0 target-major (one
 * target limb's key slice shared in L2 by every ciphertext in flight), 1
 * (default) ciphertext-major with one contiguous range of ciphertexts per
 * XCD (a ciphertext's digits and ModDown inputs shared in L2 by its targets;
 * 5% faster at N = 8192, L = 4, batch 256).
 * Results do not depend on it.  No reference counterpart (a launch-shape
 * setting). */
int ckks_ctx_set_ks_order(ckks_ctx *ctx, int pair_major);

/* Device sampling for encryption (the reference samples on the host:
 * uniform a and the rounded Box-Muller dgauss of
 * scripts/exp_e33_real_ciphertext.py:75-85, sigma = 3.2).  Counter-based
 * splitmix64 draws: the output depends only on seed and the element index.
 * Not a cryptographic generator.  uniform: out [rows][l][N] uniform in
 * [0, q_(prime0+i)).  gaussian: one rounded N(0, sigma^2) integer per (row,
 * n), written as residues into all l limbs; 0 < sigma <= 10. */
int ckks_sample_uniform(const ckks_ctx *ctx, uint64_t *out, int rows, int l, int prime0, uint64_t seed,
                        void *stream);
int ckks_sample_gaussian(const ckks_ctx *ctx, uint64_t *out, int rows, int l, int prime0, double sigma,
                         uint64_t seed, void *stream);
/* Symmetric encryption in one pass (encrypt of
 * scripts/exp_e33_real_ciphertext.py:171-175: c1 = -a, c0 = a s + e + m):
 * pt [batch][l][N] coefficient form, s_eval [l][N] the secret key's
 * residues in evaluation order (ckks_ntt_fwd of s), out [batch][2][l][N].
 * The mask is ckks_sample_uniform(rows = batch, l, prime 0, seed_a) and the
 * error ckks_sample_gaussian(batch, l, 0, sigma, seed_e), so the words are
 * those of the two samplers followed by ckks_poly_mul and ckks_pointwise.
 * FP64 arithmetic, negacyclic rings N = 2^12 .. 2^14 (else
 * CKKS_ERR_BAD_ARG / CKKS_ERR_BAD_SHAPE). */
int ckks_encrypt_sk(const ckks_ctx *ctx, const uint64_t *pt, const uint64_t *s_eval, uint64_t *out, int batch,
                    int l, double sigma, uint64_t seed_a, uint64_t seed_e, void *stream);
/* Decryption in one pass (decrypt of scripts/exp_e33_real_ciphertext.py:
 * 178-180: m = c0 + c1 s): ct [batch][2][l][N], s_eval as for
 * ckks_encrypt_sk, out [batch][l][N]; the words of ckks_poly_mul of c1 and
 * s followed by ckks_pointwise op 1.  Same arithmetic and rings. */
int ckks_decrypt_sk(const ckks_ctx *ctx, const uint64_t *ct, const uint64_t *s_eval, uint64_t *out, int batch,
                    int l, void *stream);
/* Host only: the Gaussian sampler's table, cdf[k] = floor(2^64
 * P(|round(Z)| <= k)) for k < n <= 48. */
int ckks_gaussian_cdf(double sigma, uint64_t *cdf, int n);

/* Kernel timing (measurement hook; no counterpart in the reference, whose
 * RTL counts cycles in tb/tb_ntt_cycles.sv:1-13).  Between start and stop
 * every kernel that ckks_ntt_fwd/inv, ckks_poly_mul, ckks_hmult, ckks_relin
 * and ckks_rotate launch is
 * bracketed by a pair of HIP events on the call's stream, up to
 * max_launches launches.  stop waits for the last event, writes per kind
 * (CKKS_KIND_*) the summed milliseconds and the launch count, and returns
 * the number of timed launches (negative status on error). */
enum {
    CKKS_KIND_NTT = 0,
    CKKS_KIND_TENSOR = 1,
    CKKS_KIND_FRONT = 2,
    CKKS_KIND_KEYSWITCH = 3,
    CKKS_KIND_MODDOWN = 4,
    CKKS_KIND_AUTOMORPHISM = 5,
    CKKS_KIND_POLYMUL = 6,
    CKKS_KIND_KEYSWITCH_FUSED = 7,   /* HMult's key switch of limbs 0..l-2 with ModDown and rescale fused */
    CKKS_KIND_COUNT = 8
};
int ckks_timing_start(ckks_ctx *ctx, int max_launches);
int ckks_timing_stop(ckks_ctx *ctx, double *ms, int *launches, int n_kinds);

/* Matrix-core measurement probe (no counterpart in the reference; it
 * settles whether an exact int8 matrix-core product can replace the six
 * radix-2 stages of a 64-point transform step): y = G^reps x mod q for
 * count vectors of 64 words stored element-major, x [64][count] in [0, q),
 * G the 64-point negacyclic forward transform in bit-reversed output order.
 * method 0 runs the radix-2 VALU stages, 1 and 2 the int8 product
 * (v_mfma_i32_16x16x64_i8) under the two candidate accumulator layouts, 3
 * the first layout with the next row tile's products interleaved into the
 * recombination.
 * tables (CKKS_PROBE_DFT64_TABLE_BYTES, device memory) and ws40 (the Shoup
 * companion of 2^40) come from cipherpu.dft64_probe.  count is a positive
 * multiple of 64, 1 <= reps <= 64, q an odd prime below 2^45. */
#define CKKS_PROBE_DFT64_TABLE_BYTES 51200
int ckks_probe_dft64(const uint64_t *x, uint64_t *y, const void *tables, uint64_t q, uint64_t ws40, size_t count,
                     int method, int reps, void *stream);

#ifdef __cplusplus
}
#endif
#endif
