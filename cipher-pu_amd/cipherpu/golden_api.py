"""Drop-in replacements for the reference's golden-model functions.

Same names, argument meaning and return type (lists of ints / numpy arrays)
as the Python functions the reference's scripts call; the work runs in the
HIP library on the GPU.  Contexts are cached per (N, q, root, ring).

Reference surface -> here:
  tb/ref/gen_poly_mult_vectors_n12.py  bit_rev_perm, ntt, intt,
                                       poly_mult_via_ntt, cyclic_conv_direct
  scripts/verify_keyswitch.py          poly_add, poly_sub, poly_mult, relin
  scripts/verify_rescale.py            rescale_coeff, rescale_poly
  scripts/verify_e2e_mult.py           e2e_mult, rescale (the toy rescale with
                                       scale_bits; the exact modulus switch
                                       when scale_bits is None)
  scripts/exp_e24_rotation_bootstrap.py automorphism_exact, automorphism_key_exact
  scripts/verify_multi_slot_simd.py    neg_fft, neg_ifft, encode, decode
  scripts/verify_m4_h_multilayer.py    noise_monitor_energy
  scripts/verify_m4h2_prod_calibration.py  compute_snr, snr_to_gear
  scripts/exp_e11_ckks_e2e_real.py     estimate_r_hat, select_gear
  scripts/verify_project_b_mlp_reference.py  AdaptiveGearCtrl
  rtl/m4/adaptive_gear_ctrl.sv         RtlGearCtrl (register transfer)
  scripts/exp_e24_rotation_bootstrap.py neg_conv_modq, decrypt_exact_int,
                                       rlwe_encrypt_mint, reencrypt (DARE)
  scripts/exp_e11_ckks_e2e_real.py     rlwe_encrypt, rlwe_decrypt
"""
from __future__ import annotations

import ctypes
import math
from functools import lru_cache
from typing import Sequence

import numpy as np
import torch

from ._lib import call
from .params import CHAIN4
from .ring import RingContext, bitrev_, mod_switch, to_numpy


def _dev():
    return torch.device("cuda", torch.cuda.current_device())


@lru_cache(maxsize=64)
def _ring(N: int, q: int, root, cyclic: bool) -> RingContext:
    return RingContext(N, [q], roots=None if root is None else [root], cyclic=cyclic,
                       device=torch.cuda.current_device())


def _t(a, ctx: RingContext):
    return ctx.tensor(np.asarray(a, dtype=np.uint64).reshape(1, -1))


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _list(t) -> list:
    return [int(v) for v in to_numpy(t).reshape(-1)]


def bit_rev_perm(a: Sequence[int], n: int) -> list:
    """gen_poly_mult_vectors_n12.py:54-65"""
    assert len(a) == n and n & (n - 1) == 0
    x = torch.from_numpy(np.asarray(a, dtype=np.uint64).copy()).to(_dev())
    return _list(bitrev_(x.view(1, n)))


def ntt(a: Sequence[int], omega: int, N: int, q: int) -> list:
    """Cyclic transform X[k] = sum_j a_j omega^(jk), natural order
    (gen_poly_mult_vectors_n12.py:68-82)."""
    ctx = _ring(N, q, omega % q, True)
    return _list(ctx.ntt_(_t(a, ctx), natural=True))


def intt(a: Sequence[int], omega_inv: int, N: int, q: int) -> list:
    """gen_poly_mult_vectors_n12.py:85-100"""
    ctx = _ring(N, q, pow(omega_inv, -1, q), True)
    return _list(ctx.intt_(_t(a, ctx), natural=True))


def poly_mult_via_ntt(a, b, N, q, omega, omega_inv) -> list:
    """Cyclic product (gen_poly_mult_vectors_n12.py:103-107)."""
    assert omega * omega_inv % q == 1
    ctx = _ring(N, q, omega % q, True)
    return _list(ctx.poly_mul(_t(a, ctx), _t(b, ctx)))


def cyclic_conv_direct(a, b, N, q) -> list:
    """Cyclic convolution mod (X^N - 1, q) for any modulus q < 2^45
    (gen_poly_mult_vectors_n12.py:110-117)."""
    return _list(_prod(_residues(a, q), _residues(b, q), q, cyclic=True))


def poly_add(a, b, q) -> list:
    """verify_keyswitch.py:40-41"""
    return _list(_ew(a, b, q, 1))


def poly_sub(a, b, q) -> list:
    """verify_keyswitch.py:44-45"""
    return _list(_ew(a, b, q, 2))


def _ntt_ok(n, q, cyclic=False):
    return n >= 2 and n & (n - 1) == 0 and (q - 1) % (n if cyclic else 2 * n) == 0 and _isprime(q)


def _ew(a, b, q, op):
    n = len(a)
    if _ntt_ok(n, q) or _ntt_ok(n, q, True):
        ctx = _ring(n, q, None, not _ntt_ok(n, q))
        return ctx._pw(_t(a, ctx), _t(b, ctx) if b is not None else None, op, 0, None)
    # element-wise add/sub needs no transform: exact on the GPU for any q
    x = torch.from_numpy(_residues(a, q).astype(np.int64)).to(_dev())
    y = torch.from_numpy(_residues(b, q).astype(np.int64)).to(_dev())
    return torch.remainder(x + y if op == 1 else x - y, q).to(torch.uint64)


@lru_cache(maxsize=16)
def _crt_ring(n: int, cyclic: bool) -> RingContext:
    """Three 45-bit chain primes (2^15 | q - 1): the exact product over Z
    of two polynomials mod any q < 2^45, n <= 16384, fits their product."""
    return RingContext(n, list(CHAIN4[:3]), cyclic=cyclic, device=torch.cuda.current_device())


def _prod(x, y, q, cyclic=False) -> torch.Tensor:
    """Ring product of residue vectors (numpy uint64 in [0, q), [n] or
    [rows, n]) mod (X^n -+ 1, q) on the GPU: the single-prime transform
    product where q admits one, else the exact product through three primes
    and CRT (ckks_poly_mul_modq)."""
    x = np.asarray(x, dtype=np.uint64)
    n = x.shape[-1]
    if _ntt_ok(n, q, cyclic):
        ctx = _ring(n, q, None, cyclic)
        X = ctx.tensor(x.reshape(-1, 1, n))
        Y = ctx.tensor(np.asarray(y, np.uint64).reshape(-1, 1, n))
        return ctx.poly_mul(X, Y).reshape(-1, n)
    ctx = _crt_ring(n, cyclic)
    X = ctx.tensor(x.reshape(-1, n))
    Y = ctx.tensor(np.asarray(y, np.uint64).reshape(-1, n))
    rows = X.shape[0]
    out = torch.empty(rows, n, dtype=torch.uint64, device=X.device)
    ws = torch.empty(9 * rows * n, dtype=torch.uint64, device=X.device)
    call("ckks_poly_mul_modq", ctx.handle, _ptr(X), _ptr(Y), _ptr(out), rows, int(q), _ptr(ws),
         ctx.stream())
    return out


def poly_mult(a, b, q) -> list:
    """Negacyclic product mod (X^n + 1, q) for any modulus q < 2^45
    (verify_keyswitch.py:48-59, verify_e2e_mult.py:56-67)."""
    return _list(_prod(_residues(a, q), _residues(b, q), q))


def relin(e0, e1, e2, a, b, q):
    """f0 = e0 + e2*b, f1 = e1 - e2*a (verify_keyswitch.py:62-68), any q."""
    n = len(e0)
    E2 = _residues(e2, q)
    pb = _prod(np.stack([E2, E2]), np.stack([_residues(b, q), _residues(a, q)]), q)
    e01 = torch.from_numpy(np.stack([_residues(e0, q), _residues(e1, q)]).astype(np.int64)).to(pb.device)
    f = torch.remainder(torch.stack([e01[0] + pb[0].to(torch.int64), e01[1] - pb[1].to(torch.int64)]), q)
    f = to_numpy(f.to(torch.uint64))
    return [int(v) for v in f[0].reshape(n)], [int(v) for v in f[1].reshape(n)]


def rescale_coeff(c: int, q_in: int, q_out: int) -> int:
    """round-half-up(c q_out / q_in) mod q_out (verify_rescale.py:38-53)."""
    return rescale_poly([c], q_in, q_out)[0]


def rescale_poly(c, q_in, q_out) -> list:
    """verify_rescale.py:56-57"""
    x = torch.from_numpy(np.asarray(c, dtype=np.uint64)).to(_dev())
    return _list(mod_switch(x, q_in, q_out))


def rescale(s, q_in, q_out, scale_bits):
    """The reference's toy rescale floor((s q_out + floor(q_in / 2)) /
    (q_in 2^scale_bits)) mod q_out (verify_e2e_mult.py:70-75) of one value
    or a list of values in [0, 2^64)."""
    scalar = np.ndim(s) == 0
    x = torch.from_numpy(np.asarray([s] if scalar else s, dtype=np.uint64).reshape(-1).copy()).to(_dev())
    out = _rescale_toy(x, q_in, q_out, scale_bits)
    return out[0] if scalar else out


def _rescale_toy(x: torch.Tensor, q_in, q_out, scale_bits) -> list:
    out = torch.empty_like(x)
    call("ckks_mod_switch_scaled", _ptr(x), _ptr(out), x.numel(), int(q_in), int(q_out), int(scale_bits),
         _stream())
    return _list(out)


def e2e_mult(c0_1, c1_1, c0_2, c1_2, a, b, q_in, q_out, scale_bits=None):
    """Tensor product -> relinearise -> rescale on single-modulus ciphertexts
    (verify_e2e_mult.py:78-102), for any modulus q_in < 2^45.  With
    scale_bits (the reference's positional call, :173,:205,:220) the rescale
    is the reference's toy division by q_in 2^scale_bits (:70-75), word for
    word; with scale_bits=None it is the exact modulus switch q_in -> q_out
    of verify_rescale.py."""
    n = len(c0_1)
    R = lambda v: _residues(v, q_in)
    A0, A1, B0, B1, KA, KB = (R(v) for v in (c0_1, c1_1, c0_2, c1_2, a, b))
    # m0 = A0 B0, m1 = A0 B1 + A1 B0, m2 = A1 B1 in one batched call
    m = _prod(np.stack([A0, A0, A1, A1]), np.stack([B0, B1, B0, B1]), q_in).to(torch.int64)
    m0, m1, m2 = m[0], torch.remainder(m[1] + m[2], q_in), m[3]
    m2n = to_numpy(m2.to(torch.uint64))
    e = _prod(np.stack([m2n, m2n]), np.stack([KB, KA]), q_in).to(torch.int64)   # m2 b, m2 a
    f = torch.remainder(torch.stack([m0 + e[0], m1 - e[1]]), q_in).to(torch.uint64).reshape(-1).contiguous()
    if scale_bits is None:
        g = _list(mod_switch(f, q_in, q_out))
    else:
        g = _rescale_toy(f, q_in, q_out, scale_bits)
    return g[:n], g[n:]


def automorphism_exact(ct, nu, n, q):
    """(c0, c1) -> (c0(X^nu), c1(X^nu)) mod X^n + 1 for odd nu
    (exp_e24_rotation_bootstrap.py:143-162)."""
    if nu % 2 == 0:
        raise ValueError("automorphism needs an odd Galois element")
    ctx = _ring(n, q, None, False)
    x = ctx.tensor(np.asarray([np.asarray(ct[0]) % q, np.asarray(ct[1]) % q], dtype=np.uint64).reshape(2, 1, n))
    y = to_numpy(ctx.automorphism(x, nu % (2 * n)))
    return y[0, 0].astype(np.int64), y[1, 0].astype(np.int64)


def automorphism_key_exact(s, nu, n, q):
    """exp_e24_rotation_bootstrap.py:165-176"""
    return automorphism_exact((s, np.zeros(n, dtype=np.int64)), nu, n, q)[0]


# -- natural-order canonical embedding (verify_multi_slot_simd.py) -----------

@lru_cache(maxsize=16)
def _emb_ring(n: int) -> RingContext:
    # any NTT-friendly prime works: only the floating-point tables are used
    return RingContext(n, [_friendly_prime(n)], device=torch.cuda.current_device())


def _friendly_prime(n):
    q = (1 << 44) // (2 * n) * (2 * n) + 1
    while not _isprime(q):
        q += 2 * n
    return q


def _isprime(q):
    if q < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if q % p == 0:
            return q == p
    d, s = q - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        x = pow(a, d, q)
        if x in (1, q - 1):
            continue
        for _ in range(s - 1):
            x = x * x % q
            if x == q - 1:
                break
        else:
            return False
    return True


def _natural_gather(n: int):
    """Slot k of the natural order sits at Phi^(2k+1).  With e = 2k+1:
    e = 1 mod 4 -> FFT-order slot (e-1)/4; e = 3 mod 4 -> conjugate of
    FFT-order slot (2n - e - 1)/4 (m is real)."""
    e = 2 * np.arange(n) + 1
    direct = (e % 4) == 1
    idx = np.where(direct, (e - 1) // 4, (2 * n - e - 1) // 4)
    return idx, ~direct


def neg_fft(m):
    """Evaluation at Phi^(2k+1), k < n, natural order, for real m
    (verify_multi_slot_simd.py:32-39)."""
    m = np.asarray(m, dtype=np.float64)
    n = m.shape[-1]
    ctx = _emb_ring(n)
    v = ctx.embed(torch.from_numpy(m.reshape(-1, n).copy()), order=1)
    idx, conj = _natural_gather(n)
    out = v[:, torch.from_numpy(idx).to(v.device)]
    cj = torch.from_numpy(conj).to(v.device)
    out = torch.where(cj, out.conj(), out)
    return to_numpy(out).reshape(m.shape[:-1] + (n,))


def _neg_ifft_real(H: np.ndarray) -> np.ndarray:
    """Inverse of ``neg_fft`` for Hermitian rows H (H[n-1-k] = conj H[k]):
    the real polynomial, from the FFT-order half of H."""
    n = H.shape[-1]
    ctx = _emb_ring(n)
    fft_order = H[:, 0::2]   # k = 2k' has e = 4k' + 1
    return to_numpy(ctx.embed_inv(torch.from_numpy(np.ascontiguousarray(fft_order)), order=1))


def neg_ifft(V):
    """Inverse of ``neg_fft``: m_j = (1/n) sum_k V_k Phi^(-(2k+1) j), complex
    (verify_multi_slot_simd.py:42-52).  Any V splits into two Hermitian
    vectors, H = (V + V*)/2 and A = (V - V*)/(2i) with V*_k = conj V_(n-1-k),
    whose inverses are the real and imaginary parts of m; for a Hermitian V
    (real polynomials, encode with nslot <= n/2) A is zero."""
    V = np.asarray(V, dtype=np.complex128)
    n = V.shape[-1]
    rows = V.reshape(-1, n)
    Vs = np.conj(rows[:, ::-1])
    H = (rows + Vs) / 2
    A = (rows - Vs) / 2j
    both = _neg_ifft_real(np.concatenate([H, A]))
    r = rows.shape[0]
    m = both[:r] + 1j * both[r:]
    return m.reshape(V.shape[:-1] + (n,))


def encode(v, scale, n):
    """verify_multi_slot_simd.py:55-84"""
    nslot = len(v)
    V = np.zeros(n, dtype=np.complex128)
    for k in range(nslot):
        V[k] = scale * v[k]
        sk = (n - 1 - k) % n
        if sk != k and sk >= nslot:
            V[sk] = np.conj(V[k])
    return neg_ifft(V).real


def decode(m, scale, n, nslot):
    """verify_multi_slot_simd.py:87-91"""
    return neg_fft(np.asarray(m, dtype=np.float64)).real[:nslot] / scale


# ---------------------------------------------------------------------------
# Noise monitor and gear policy (the monitor's energy runs on the GPU; the
# policies are scalar host logic, as in the reference).

def noise_monitor_energy(coeffs, q):
    """(E, r_hat) with E = sum min(c, q - c)^2 / count and r_hat = sqrt(E)/q
    (scripts/verify_m4_h_multilayer.py:162-174; rtl/m4/noise_monitor.sv),
    for any modulus q < 2^45 and any count (the reference calls it with
    q = 113 on 32 coefficients, :260).  The sum of squares is the statistics
    kernel's exact 128-bit row sum."""
    c = np.asarray(coeffs, dtype=np.uint64).reshape(-1)
    n = c.size
    if n == 0:
        raise ZeroDivisionError("noise_monitor_energy of an empty coefficient list")
    if int(c.max()) >= q:
        raise ValueError("coefficients must lie in [0, q)")
    x = torch.from_numpy(c.copy()).to(_dev())
    st = torch.empty(5, dtype=torch.uint64, device=x.device)
    call("ckks_noise_stats_q", _ptr(x), _ptr(st), 1, n, int(q), _stream())
    w = [int(v) for v in to_numpy(st)]
    E = (w[1] + (w[2] << 64)) / n
    return E, math.sqrt(E) / q


def compute_snr(signal, noise):
    """10 log10(sum |signal|^2 / sum |noise|^2) in dB, inf for zero noise
    (scripts/verify_m4h2_prod_calibration.py:64-69)."""
    sp = float(np.sum(np.abs(np.asarray(signal)) ** 2))
    npw = float(np.sum(np.abs(np.asarray(noise)) ** 2))
    return math.inf if npw == 0 else 10.0 * math.log10(sp / npw)


def snr_to_gear(snr_db, threshold_db=15.0):
    """'Gear 3' above 20 dB, 'Gear 2' above threshold_db, 'Gear 1' above
    10 dB, else 'Gear 0' -- the reference's labels
    (scripts/verify_m4h2_prod_calibration.py:72-82)."""
    if snr_db > 20:
        return "Gear 3"
    if snr_db > threshold_db:
        return "Gear 2"
    if snr_db > 10:
        return "Gear 1"
    return "Gear 0"


# Largest noise ratio r_hat each gear tolerates, gear 0 (strict) to 3
# (scripts/exp_e11_ckks_e2e_real.py:182-187,
# scripts/verify_project_b_mlp_reference.py:245-250).
GEAR_R_MAX = (0.10, 0.20, 0.35, 0.55)


def estimate_r_hat(decrypted_val, true_val):
    """Relative decryption error, capped at 1 (exp_e11_ckks_e2e_real.py:169-176)."""
    return float(min(abs(decrypted_val - true_val) / (abs(true_val) + 1e-8), 1.0))


def select_gear(r_hat, safety_first=True):
    """Safety first: the strictest gear whose r_max tolerates r_hat, gear 0 if
    none does; otherwise the loosest such gear (exp_e11_ckks_e2e_real.py:189-200)."""
    ok = [g for g, r in enumerate(GEAR_R_MAX) if r >= r_hat]
    if not ok:
        return 0
    return ok[0] if safety_first else ok[-1]


class AdaptiveGearCtrl:
    """Safety-first gear choice with hysteresis: a new target gear must be
    seen on `hyst` + 1 consecutive steps before the gear changes; starts at
    the loosest gear (scripts/verify_project_b_mlp_reference.py:279-317)."""

    def __init__(self, hyst: int = 2, gear: int = 3):
        self.gear, self.hyst = gear, hyst
        self.hyst_cnt, self.hyst_target = 0, -1

    def step(self, r_hat: float) -> int:
        target = select_gear(r_hat, True)
        if target != self.gear:
            if target == self.hyst_target:
                self.hyst_cnt += 1
            else:
                self.hyst_target, self.hyst_cnt = target, 0
            if self.hyst_cnt >= self.hyst:
                self.gear, self.hyst_cnt, self.hyst_target = target, 0, -1
        return self.gear


class RtlGearCtrl:
    """The register transfer of rtl/m4/adaptive_gear_ctrl.sv:42-100, one
    step per flush_valid (hysteresis on, no manual gear).  Thresholds
    t1 > t2 > t3 map the monitor's energy to a raw gear (>= t1 -> 0,
    >= t2 -> 1, >= t3 -> 2, else 3).  A stricter raw gear is taken at once.
    A looser one is a candidate: the same candidate again counts up, a new
    one restarts the count at 1, and the gear moves to the reading when the
    count before this step had reached `hyst` (the switch test reads the
    register's old value).  A reading at the current gear clears the count."""

    def __init__(self, t1: float, t2: float, t3: float, hyst: int = 3):
        self.t = (t1, t2, t3)
        self.hyst = hyst
        self.gear, self.cnt, self.cand = 3, 0, 3      # reset values

    def raw(self, energy: float) -> int:
        return next((g for g, t in enumerate(self.t) if energy >= t), 3)

    def step(self, energy: float) -> int:
        r = self.raw(energy)
        gear, cnt = self.gear, self.cnt
        if r < gear:
            self.gear, self.cnt, self.cand = r, 0, r
        elif r == gear:
            self.cnt, self.cand = 0, r
        else:
            self.cnt = cnt + 1 if r == self.cand else 1
            self.cand = r
            if cnt >= self.hyst:
                self.gear, self.cnt, self.cand = r, 0, gear
        return self.gear


# ---------------------------------------------------------------------------
# Single-modulus RLWE helpers of the reference's experiments
# (scripts/exp_e24_rotation_bootstrap.py, scripts/exp_e11_ckks_e2e_real.py).
# The negacyclic products run on the GPU and are exact; the mask a is drawn
# from numpy's global generator with the reference's own calls, so a run
# seeded like the reference's gets the same a.  Any modulus q < 2^45.

def _residues(x, q) -> np.ndarray:
    return np.array([int(v) % q for v in np.asarray(x).reshape(-1)], dtype=np.uint64)


def _negconv(x, y, q) -> np.ndarray:
    n = len(x)
    return to_numpy(_prod(_residues(x, q), _residues(y, q), q)).reshape(n).astype(np.int64)


def neg_conv_modq(x, y, q):
    """Exact negacyclic product mod (X^n + 1, q) for any q < 2^45, int64
    (exp_e24_rotation_bootstrap.py:80-106)."""
    return _negconv(x, y, q)


def decrypt_exact_int(ct, s, q):
    """(b - a s) mod q centred to (-q/2, q/2], int64
    (exp_e24_rotation_bootstrap.py:126-133)."""
    d = (_residues(ct[0], q).astype(np.int64) - _negconv(ct[1], s, q)) % q
    return np.where(d > q // 2, d - q, d).astype(np.int64)


def _draw_a(n: int, q: int) -> np.ndarray:
    # exp_e24_rotation_bootstrap.py:217-220 / exp_e11_ckks_e2e_real.py:79-82
    if q <= 2 ** 31 - 1:
        return np.random.randint(0, q, size=n).astype(np.int64)
    return np.floor(np.random.random(size=n) * q).astype(np.int64)


def rlwe_encrypt_mint(m_int, s, e, q):
    """(b, a) with b = a s + e + m mod q for an integer message, int64
    (exp_e24_rotation_bootstrap.py:209-222)."""
    a = _draw_a(len(s), q)
    extra = (_residues(e, q).astype(np.int64) + _residues(m_int, q).astype(np.int64)) % q
    return (_negconv(a, s, q) + extra) % q, a


def reencrypt(ct, s_key, q, scale, rng):
    """DARE refresh: the decrypted integer message (scale kept, no rounding)
    encrypted afresh with ternary noise drawn from rng; returns
    ((b, a), m_new) (exp_e24_rotation_bootstrap.py:287-304)."""
    m_new = decrypt_exact_int(ct, s_key, q)
    e_new = rng.integers(-1, 2, size=len(s_key)).astype(np.int64)
    return rlwe_encrypt_mint(m_new, s_key, e_new, q), m_new


def rlwe_encrypt(m_poly, s, e, q, scale):
    """(b, a) with b = a s + e + scale m mod q, the fractional message part
    kept in float64 (exp_e11_ckks_e2e_real.py:72-88).  a s is exact and
    reduced mod q before the float64 sum, where the reference sums its
    float64 convolution unreduced: equal up to float64 rounding."""
    a = _draw_a(len(s), q)
    total = _negconv(a, s, q).astype(np.float64) + np.asarray(e, dtype=np.float64)
    return np.mod(total + scale * np.asarray(m_poly, dtype=np.float64), q), a


def rlwe_decrypt(ct, s, q):
    """(b - a s) mod q centred to [-q/2, q/2) in float64, i.e. scale m + e
    (exp_e11_ckks_e2e_real.py:66-70,91-99)."""
    v = np.mod(np.asarray(ct[0], dtype=np.float64) - _negconv(ct[1], s, q).astype(np.float64), q)
    return np.where(v > q / 2, v - q, v)


def decrypt_exact(ct, s, q):
    """decrypt_exact_int as float64 (exp_e24_rotation_bootstrap.py:109-123)."""
    return decrypt_exact_int(ct, s, q).astype(np.float64)


def mul_plain_exact(ct, p_poly, q):
    """(b p, a p) mod q, exact, int64 (exp_e24_rotation_bootstrap.py:136-140)."""
    return _negconv(ct[0], p_poly, q), _negconv(ct[1], p_poly, q)


def poly_x_power(nu, n):
    """The monomial X^(nu mod n) (exp_e24_rotation_bootstrap.py:225-229)."""
    p = np.zeros(n, dtype=np.int64)
    p[nu % n] = 1
    return p


def rotate_monomial(ct, nu, q):
    """ct times X^nu, i.e. a negacyclic shift of both polynomials: the
    reference's coefficient-domain `rotate` (exp_e24_rotation_bootstrap.py:232-238),
    named apart from the slot rotation of the CKKS API."""
    return mul_plain_exact(ct, poly_x_power(nu, len(ct[0])), q)


def scatter_w(msg_indices, w, scale, q, n):
    """Integer plaintext that puts round(scale w_j) at position n - msg_j
    (negated when msg_j = 0 mod n), summing collisions, reduced mod q
    (exp_e24_rotation_bootstrap.py:241-262)."""
    acc = np.zeros(n, dtype=object)
    for j, mi in enumerate(msg_indices):
        v = int(round(scale * float(w[j])))
        k = int(mi) % n
        acc[(n - k) % n] += -v if k == 0 else v
    return np.array([int(x) % q for x in acc], dtype=np.int64)


def inner_product(ct, w, s_key, q, scale, msg_indices=None):
    """sum_j x_(msg_j) w_j from one plaintext product and the constant term
    of its decryption, -dec[0] / scale^2 (exp_e24_rotation_bootstrap.py:265-275)."""
    w = np.asarray(w, dtype=np.float64)
    idx = range(len(w)) if msg_indices is None else msg_indices
    prod = mul_plain_exact(ct, scatter_w(idx, w, scale, q, len(ct[0])), q)
    return -float(decrypt_exact(prod, s_key, q)[0]) / (scale * scale)


def coef_sum(ct, dim, s_key, q, scale):
    """sum_(i < dim) x_i: the inner product with ones
    (exp_e24_rotation_bootstrap.py:278-284)."""
    return inner_product(ct, np.ones(dim), s_key, q, scale)
