"""CPU oracle for the CKKS evaluation path -- TEST INFRASTRUCTURE ONLY.

This module is a plain-Python (big-int) and numpy restatement of the
reference's golden models.  It exists to check the HIP library, never to
stand in for it: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker.
The product package never imports anything under ``oracle/``.

Every function cites the reference file:line it restates.  Conventions that
the reference leaves ambiguous are frozen here once:

* ring: negacyclic Z_q[X]/(X^N + 1) for CKKS; the reference's cyclic
  transform (``ntt``/``intt``) is kept as-is for its own call sites;
* ciphertext: ct = (c0, c1), decrypt = c0 + c1*s; relinearisation key
  b = a*s + e + s^2, f0 = d0 + d2*b, f1 = d1 - d2*a
  (scripts/exp_e33_real_ciphertext.py:153-175), generalised to a hybrid
  key switch with one digit per RNS prime and one special prime p;
* rescale: exact round-half-up modulus switch of the CRT value,
  scripts/verify_rescale.py:38-53, applied from Q_l to Q_{l-1};
* slots: canonical embedding at zeta^(5^j), zeta = exp(i*pi/N).

Pinning (what checks each function; see tests/test_oracle_pinning.py):
  ntt/intt/bit_rev_perm ..... tb/tb_ntt_n8.sv KAT, DFT closed form, round trip
  poly_mult_via_ntt ......... tb/tb_pm_n8.sv and tb/tb_pm_n16.sv KATs, direct
                              convolution
  negacyclic_ntt ............ closed form sum a_j psi^((2k+1)j), schoolbook
  power_table ............... rtl/m2/omega_table_chain4_q*.hex digests,
                              rtl/m5/tw/*.hex
  poly_mult (schoolbook) .... X^N = -1 monomial identities, NTT product
  rescale_coeff ............. verify_rescale.py invariants + big-int division
  rescale_toy ............... equals rescale_coeff at scale_bits 0 (odd q_in),
                              rescale_coeff >> scale_bits below q_in, exact
                              fractions on small inputs
  e2e_mult .................. composition of the pinned products; with the
                              E33 key (b = a s + e + s^2) it decrypts to the
                              rescaled plaintext product
  rns_rescale ............... CRT big-int rescale_coeff (independent path)
  keyswitch_hybrid .......... dnum=1,P=1 reduces to relin(); decrypt invariant
  automorphism .............. monomial closed form, decode(rot) = roll(decode)
  encode/decode ............. direct O(N^2) evaluation, SIMD multiplicativity
  noise monitor / gears ..... closed forms and the thresholds cited
"""
from __future__ import annotations

import math
import random
from dataclasses import dataclass, field
from typing import List, Sequence, Tuple

import numpy as np

# ---------------------------------------------------------------------------
# number theory (scripts/gen_modulus_chain.py, tb/ref/gen_ntt_reference.py)
# ---------------------------------------------------------------------------

_MR_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n: int) -> bool:
    """Deterministic Miller-Rabin for n < 3.3e24 (the reference uses a
    trial/MR test in scripts/gen_modulus_chain.py for the same purpose)."""
    if n < 2:
        return False
    for p in _MR_BASES:
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _MR_BASES:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def _factor_small(n: int) -> List[int]:
    fs, d = [], 2
    while d * d <= n:
        if n % d == 0:
            fs.append(d)
            while n % d == 0:
                n //= d
        d += 1 if d == 2 else 2
    if n > 1:
        fs.append(n)
    return fs


def primitive_root(q: int) -> int:
    """Smallest generator of Z_q^*  (tb/ref/gen_ntt_reference.py:38-61)."""
    fs = _factor_small(q - 1)
    g = 2
    while True:
        if all(pow(g, (q - 1) // f, q) != 1 for f in fs):
            return g
        g += 1


def root_of_unity(q: int, order: int) -> int:
    """g^((q-1)/order) with g the smallest generator -- the construction of
    the chain's omega (scripts/gen_modulus_chain.py:128-130)."""
    assert (q - 1) % order == 0
    return pow(primitive_root(q), (q - 1) // order, q)


def smallest_primitive_root_of_unity(q: int, order: int) -> int:
    """Smallest x in [2, q) of exact multiplicative order ``order`` (a power
    of two): x^(order/2) == q-1.  Defines psi for the N=4096 product."""
    assert order & (order - 1) == 0 and (q - 1) % order == 0
    # the primitive roots of order 2^k are exactly psi0^j for odd j
    psi0 = root_of_unity(q, order)
    sq = psi0 * psi0 % q
    x, best = psi0, q
    for _ in range(order // 2):
        best = min(best, x)
        x = x * sq % q
    return best


def largest_prime_1_mod(bits: int, modulus: int) -> int:
    """Largest prime below 2^bits that is 1 mod ``modulus``."""
    k = ((1 << bits) - 2) // modulus
    while k > 0:
        q = k * modulus + 1
        if is_prime(q):
            return q
        k -= 1
    raise ValueError("no prime")


def primes_2k_plus(bits: int, step: int, count: int, start_j: int = 1) -> List[int]:
    """Primes of the form 2^(bits-1) + step*j + 1, j >= start_j, ascending
    (the search space of scripts/gen_modulus_chain.py:83-95)."""
    out, j, base = [], start_j, 1 << (bits - 1)
    while len(out) < count:
        q = base + step * j + 1
        if is_prime(q):
            out.append(q)
        j += 1
    return out


def power_table(omega: int, count: int, q: int) -> List[int]:
    """omega^k mod q, k = 0..count-1 (scripts/gen_omega_tables.py:96-106)."""
    t, x = [], 1
    for _ in range(count):
        t.append(x)
        x = x * omega % q
    return t


# ---------------------------------------------------------------------------
# cyclic transform of the reference (tb/ref/gen_poly_mult_vectors_n12.py)
# ---------------------------------------------------------------------------

def bit_reverse(i: int, bits: int) -> int:
    r = 0
    for _ in range(bits):
        r = (r << 1) | (i & 1)
        i >>= 1
    return r


def bit_rev_perm(a: Sequence[int], n: int) -> List[int]:
    """Bit-reversal permutation of a copy (gen_poly_mult_vectors_n12.py:54-65)."""
    bits = n.bit_length() - 1
    assert 1 << bits == n and len(a) == n
    return [a[bit_reverse(i, bits)] for i in range(n)]


def ntt(a: Sequence[int], omega: int, N: int, q: int) -> List[int]:
    """Cyclic radix-2 DIT transform X[k] = sum_j a_j omega^(jk), natural
    order out, bit-reversed input order internally; stage roots are
    omega^(N/len) stepped by repeated multiplication
    (gen_poly_mult_vectors_n12.py:68-82)."""
    x = bit_rev_perm(a, N)
    half = 1
    while half < N:
        step = pow(omega, N // (2 * half), q)
        for base in range(0, N, 2 * half):
            w = 1
            for k in range(half):
                u = x[base + k]
                v = x[base + k + half] * w % q
                x[base + k] = (u + v) % q
                x[base + k + half] = (u - v) % q
                w = w * step % q
        half *= 2
    return x


def intt(a: Sequence[int], omega_inv: int, N: int, q: int) -> List[int]:
    """Inverse of ``ntt``: the same DIT with omega^-1, then times N^-1
    (gen_poly_mult_vectors_n12.py:85-100)."""
    x = ntt(a, omega_inv, N, q)
    n_inv = pow(N, -1, q)
    return [v * n_inv % q for v in x]


def poly_mult_via_ntt(a, b, N, q, omega, omega_inv):
    """Cyclic product by transform (gen_poly_mult_vectors_n12.py:103-107)."""
    ah = ntt(a, omega, N, q)
    bh = ntt(b, omega, N, q)
    return intt([x * y % q for x, y in zip(ah, bh)], omega_inv, N, q)


def cyclic_conv_direct(a, b, N, q):
    """O(N^2) cyclic convolution (gen_poly_mult_vectors_n12.py:110-117)."""
    return [sum(a[k] * b[(i - k) % N] for k in range(N)) % q for i in range(N)]


# ---------------------------------------------------------------------------
# negacyclic ring (scripts/verify_keyswitch.py:48-59; the cyclic ntt of
# tb/ref/gen_poly_mult_vectors_n12.py:68-82 wrapped with psi twists)
# ---------------------------------------------------------------------------

def negacyclic_ntt(a: Sequence[int], psi: int, N: int, q: int) -> List[int]:
    """Xneg[k] = sum_j a_j psi^((2k+1) j): the cyclic ``ntt`` of a_j psi^j
    with omega = psi^2.  Natural order."""
    tw = power_table(psi, N, q)
    return ntt([x * t % q for x, t in zip(a, tw)], psi * psi % q, N, q)


def negacyclic_intt(A: Sequence[int], psi: int, N: int, q: int) -> List[int]:
    """Inverse of ``negacyclic_ntt``."""
    psi_inv = pow(psi, -1, q)
    x = intt(A, psi_inv * psi_inv % q, N, q)
    tw = power_table(psi_inv, N, q)
    return [v * t % q for v, t in zip(x, tw)]


def poly_add(a, b, q):
    """verify_keyswitch.py:40-41"""
    return [(x + y) % q for x, y in zip(a, b)]


def poly_sub(a, b, q):
    """verify_keyswitch.py:44-45"""
    return [(x - y) % q for x, y in zip(a, b)]


def poly_mult(a, b, q):
    """Schoolbook negacyclic product mod (X^n + 1, q): the term a_i b_j goes
    to X^(i+j) with a sign flip when i + j >= n (verify_keyswitch.py:48-59)."""
    n = len(a)
    acc = [0] * (2 * n)
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                acc[i + j] += x * y
    return [(acc[k] - acc[k + n]) % q for k in range(n)]


def poly_mult_ntt(a, b, psi, q):
    """Negacyclic product through ``negacyclic_ntt`` (the fast path that the
    schoolbook product pins)."""
    n = len(a)
    A = negacyclic_ntt(a, psi, n, q)
    B = negacyclic_ntt(b, psi, n, q)
    return negacyclic_intt([x * y % q for x, y in zip(A, B)], psi, n, q)


def relin(e0, e1, e2, a, b, q):
    """Single-modulus relinearisation f0 = e0 + e2*b, f1 = e1 - e2*a
    (verify_keyswitch.py:62-68; key convention exp_e33:153-175)."""
    return poly_add(e0, poly_mult(e2, b, q), q), poly_sub(e1, poly_mult(e2, a, q), q)


# ---------------------------------------------------------------------------
# rescale (scripts/verify_rescale.py:38-57)
# ---------------------------------------------------------------------------

def rescale_coeff(c: int, q_in: int, q_out: int) -> int:
    """round-half-up(c * q_out / q_in) mod q_out, exact integer arithmetic
    (verify_rescale.py:38-53)."""
    t = c * q_out
    r, rem = divmod(t, q_in)
    if 2 * rem >= q_in:
        r += 1
    return r % q_out


def rescale_poly(c, q_in, q_out):
    """verify_rescale.py:56-57"""
    return [rescale_coeff(x, q_in, q_out) for x in c]


def rescale_toy(s: int, q_in: int, q_out: int, scale_bits: int) -> int:
    """The toy rescale floor((s q_out + floor(q_in/2)) / (q_in 2^scale_bits))
    mod q_out (scripts/verify_e2e_mult.py:70-75, also
    gen_tb_e2e_vec.py:59-63, gen_tb_rescale_vec.py:28-33)."""
    return (s * q_out + q_in // 2) // (q_in << scale_bits) % q_out


def e2e_mult(c0_1, c1_1, c0_2, c1_2, a, b, q_in, q_out, scale_bits=None):
    """Tensor product, relinearisation f0 = m0 + m2 b, f1 = m1 - m2 a, then
    the rescale of every coefficient (scripts/verify_e2e_mult.py:78-102):
    the toy rescale with scale_bits, the exact rescale_coeff when it is
    None.  Negacyclic schoolbook products, any q_in."""
    m0 = poly_mult(c0_1, c0_2, q_in)
    m1 = poly_add(poly_mult(c0_1, c1_2, q_in), poly_mult(c1_1, c0_2, q_in), q_in)
    m2 = poly_mult(c1_1, c1_2, q_in)
    f0 = poly_add(m0, poly_mult(m2, b, q_in), q_in)
    f1 = poly_sub(m1, poly_mult(m2, a, q_in), q_in)
    if scale_bits is None:
        return rescale_poly(f0, q_in, q_out), rescale_poly(f1, q_in, q_out)
    return ([rescale_toy(x, q_in, q_out, scale_bits) for x in f0],
            [rescale_toy(x, q_in, q_out, scale_bits) for x in f1])


def crt_compose(limbs: Sequence[Sequence[int]], primes: Sequence[int]) -> List[int]:
    """CRT value in [0, prod(primes)) for each coefficient."""
    Q = math.prod(primes)
    terms = []
    for qi in primes:
        Qi = Q // qi
        terms.append((Qi, pow(Qi % qi, -1, qi)))
    n = len(limbs[0])
    out = []
    for k in range(n):
        s = 0
        for (Qi, inv), qi, limb in zip(terms, primes, limbs):
            s += (limb[k] * inv % qi) * Qi
        out.append(s % Q)
    return out


def crt_split(values: Sequence[int], primes: Sequence[int]) -> List[List[int]]:
    return [[v % qi for v in values] for qi in primes]


def rns_rescale(limbs, primes):
    """Drop the last prime: out = [rescale_coeff(CRT(c), Q_l, Q_(l-1))]_(q_i)
    for i < l-1.  Big-int restatement of verify_rescale.py:38-53 on the CRT
    value."""
    Q_in = math.prod(primes)
    Q_out = Q_in // primes[-1]
    vals = crt_compose(limbs, primes)
    res = [rescale_coeff(v, Q_in, Q_out) for v in vals]
    return crt_split(res, primes[:-1])


def rns_rescale_per_limb(limbs, primes):
    """Per-limb form of ``rns_rescale``:
    out_i = ((c_i + h) - [(c_L + h) mod q_L]) * q_L^-1 mod q_i, h = (q_L-1)/2.
    An exact integer identity with the CRT form (checked in the tests)."""
    qL = primes[-1]
    h = qL // 2
    t = [(c + h) % qL for c in limbs[-1]]
    out = []
    for qi, limb in zip(primes[:-1], limbs[:-1]):
        inv = pow(qL, -1, qi)
        out.append([((c + h - tt) % qi) * inv % qi for c, tt in zip(limb, t)])
    return out


def mod_down_round(limbs_qp, primes, p):
    """round-half-up(X / p) mod Q for X given on the basis primes + [p]
    (limbs_qp[-1] is the p limb).  Big-int definition of ModDown."""
    vals = crt_compose(limbs_qp, list(primes) + [p])
    QP = math.prod(primes) * p
    res = [rescale_coeff(v, QP, QP // p) for v in vals]
    return crt_split(res, primes)


# ---------------------------------------------------------------------------
# hybrid key switching (generalises verify_keyswitch.relin,
# scripts/verify_keyswitch.py:62-68)
# ---------------------------------------------------------------------------

def gadget_factor(primes, p, j, i):
    """[P * Qhat_j * (Qhat_j^-1 mod q_j)] reduced mod basis prime i
    (i == len(primes) selects p).  It is p mod q_j on limb j and 0
    elsewhere."""
    if i == len(primes):
        return 0
    return p % primes[i] if i == j else 0


def keyswitch_hybrid(d2, key, primes, p, psis, N):
    """Hybrid key switch with one digit per prime.

    d2:  l limbs (coefficient form), digit j = [d2]_(q_j) lifted to Z.
    key: list over digits j of (b_j, a_j), each a list of l+1 limbs over
         (q_0..q_(l-1), p), with b_j = a_j s + e_j + g_j s_from.
    Returns (F0, F1) after ModDown: F0 = round(sum_j D_j b_j / p),
    F1 = -round(sum_j D_j a_j / p) ... computed as round(-sum/p) with the
    negation applied before the division (E33 sign convention:
    f1 = d1 - d2*a).
    """
    basis = list(primes) + [p]
    l = len(primes)
    acc0 = [[0] * N for _ in basis]
    acc1 = [[0] * N for _ in basis]
    for j in range(l):
        digit = d2[j]
        bj, aj = key[j]
        for i, qi in enumerate(basis):
            dj = [x % qi for x in digit]
            pb = poly_mult_ntt(dj, bj[i], psis[i], qi)
            pa = poly_mult_ntt(dj, aj[i], psis[i], qi)
            acc0[i] = poly_add(acc0[i], pb, qi)
            acc1[i] = poly_sub(acc1[i], pa, qi)
    return mod_down_round(acc0, primes, p), mod_down_round(acc1, primes, p)


def hmult(ct_a, ct_b, key, primes, p, psis, N):
    """HMult + relinearise + rescale on RNS ciphertexts (coefficient form).

    ct = (c0, c1), each a list of l limbs.  Tensor product
    d0 = a0 b0, d1 = a0 b1 + a1 b0, d2 = a1 b1 (verify_e2e_mult.py:78-92),
    key switch of d2 as above, f = d + ModDown(.), then ``rns_rescale``
    (the exact modulus switch that replaces the toy rescale at
    verify_e2e_mult.py:70-75)."""
    l = len(primes)
    d0, d1, d2 = [], [], []
    for i in range(l):
        qi, ps = primes[i], psis[i]
        a0, a1 = ct_a[0][i], ct_a[1][i]
        b0, b1 = ct_b[0][i], ct_b[1][i]
        d0.append(poly_mult_ntt(a0, b0, ps, qi))
        d1.append(poly_add(poly_mult_ntt(a0, b1, ps, qi), poly_mult_ntt(a1, b0, ps, qi), qi))
        d2.append(poly_mult_ntt(a1, b1, ps, qi))
    k0, k1 = keyswitch_hybrid(d2, key, primes, p, psis, N)
    f0 = [poly_add(d0[i], k0[i], primes[i]) for i in range(l)]
    f1 = [poly_add(d1[i], k1[i], primes[i]) for i in range(l)]
    return rns_rescale(f0, primes), rns_rescale(f1, primes)


# ---------------------------------------------------------------------------
# Galois automorphism (scripts/exp_e24_rotation_bootstrap.py:143-176)
# ---------------------------------------------------------------------------

def automorphism(f: Sequence[int], nu: int, n: int, q: int) -> List[int]:
    """f(X) -> f(X^nu) mod (X^n + 1): coefficient k lands on (nu*k) mod n
    with sign (-1)^floor(nu*k / n), accumulated (exp_e24:143-162)."""
    g = [0] * n
    for k, c in enumerate(f):
        if c:
            t = nu * k
            j = t % n
            g[j] = (g[j] + (c if (t // n) % 2 == 0 else -c)) % q
    return g


def galois_elt(step: int, N: int) -> int:
    """5^step mod 2N (negative steps rotate the other way)."""
    return pow(5, step % (N // 2), 2 * N)


# ---------------------------------------------------------------------------
# canonical embedding (scripts/verify_multi_slot_simd.py:32-91)
# ---------------------------------------------------------------------------

def neg_fft(m):
    """Direct evaluation at Phi^(2k+1), Phi = exp(i pi/N), natural k order
    (verify_multi_slot_simd.py:32-39)."""
    m = np.asarray(m, dtype=np.complex128)
    n = len(m)
    k = np.arange(n)
    W = np.exp(1j * np.pi * np.outer(2 * k + 1, np.arange(n)) / n)
    return W @ m


def neg_ifft(V):
    """Inverse of ``neg_fft`` (verify_multi_slot_simd.py:42-52)."""
    V = np.asarray(V, dtype=np.complex128)
    n = len(V)
    k = np.arange(n)
    W = np.exp(-1j * np.pi * np.outer(np.arange(n), 2 * k + 1) / n)
    return (W @ V) / n


def encode_natural(v, scale, n):
    """Slots at Phi^(2k+1) in natural order with the conjugate at slot
    n-1-k (verify_multi_slot_simd.py:55-84); returns real coefficients."""
    nslot = len(v)
    V = np.zeros(n, dtype=np.complex128)
    for k in range(nslot):
        V[k] = scale * v[k]
        sk = (n - 1 - k) % n
        if sk != k and sk >= nslot:
            V[sk] = np.conj(V[k])
    return neg_ifft(V).real


def decode_natural(m, scale, n, nslot):
    """verify_multi_slot_simd.py:87-91"""
    return neg_fft(np.asarray(m, dtype=np.float64)).real[:nslot] / scale


def slot_exponents(N: int) -> np.ndarray:
    """e_j = 5^j mod 2N, j < N/2: slot j sits at zeta^(e_j)."""
    e = np.empty(N // 2, dtype=np.int64)
    x = 1
    for j in range(N // 2):
        e[j] = x
        x = x * 5 % (2 * N)
    return e


def embed_direct(m, N):
    """z_j = m(zeta^(5^j)) by direct O(N^2) evaluation (FP64)."""
    e = slot_exponents(N)
    W = np.exp(1j * np.pi * np.outer(e, np.arange(N)) / N)
    return W @ np.asarray(m, dtype=np.float64)


def decode_fft(m, N):
    """z_j = m(zeta^(5^j)) through one length-N/2 complex FFT:
    u_i = m_i + i m_(i+N/2); z_j = sum_i (u_i zeta^i) w^(i k_j),
    w = exp(2 pi i/(N/2)), 4 k_j + 1 = 5^j mod 2N."""
    n = N // 2
    m = np.asarray(m, dtype=np.float64)
    u = m[:n] + 1j * m[n:]
    y = u * np.exp(1j * np.pi * np.arange(n) / N)
    V = np.fft.ifft(y) * n
    k = (slot_exponents(N) - 1) // 4
    return V[k]


def encode_fft(z, N):
    """Real coefficients m (unrounded) with m(zeta^(5^j)) = z_j."""
    n = N // 2
    V = np.zeros(n, dtype=np.complex128)
    V[(slot_exponents(N) - 1) // 4] = np.asarray(z, dtype=np.complex128)
    y = np.fft.fft(V) / n
    u = y * np.exp(-1j * np.pi * np.arange(n) / N)
    return np.concatenate([u.real, u.imag])


def encode_rns(z, scale, N, primes):
    """round(scale * embedding^-1(z)) reduced into each prime."""
    m = np.rint(encode_fft(np.asarray(z) * scale, N)).astype(np.int64)
    return [[int(x) % q for x in m] for q in primes]


def decode_rns(limbs, scale, N, primes):
    """Centered CRT value / scale, evaluated at the slots."""
    vals = crt_compose(limbs, primes) if len(primes) > 1 else list(limbs[0])
    Q = math.prod(primes)
    c = np.array([float(v - Q if v > Q // 2 else v) for v in vals]) / scale
    return decode_fft(c, N)


# ---------------------------------------------------------------------------
# RLWE primitives (exp_e33:75-85,151-191; exp_e24:209-222,287-304)
# ---------------------------------------------------------------------------

def decrypt(ct, s_limbs, primes, psis):
    """c0 + c1*s per limb (E33 convention, exp_e33:138-141)."""
    out = []
    for i, qi in enumerate(primes):
        out.append(poly_add(ct[0][i], poly_mult_ntt(ct[1][i], s_limbs[i], psis[i], qi), qi))
    return out


# ---------------------------------------------------------------------------
# single-modulus RLWE helpers of the reference's experiments (exact integers)
# ---------------------------------------------------------------------------

def neg_conv_modq(x, y, q):
    """Exact negacyclic product mod (X^n + 1, q) as int64
    (exp_e24_rotation_bootstrap.py:80-106; the reference splits operands
    into 10-bit pieces so float64 stays exact -- here: big integers)."""
    return np.array(poly_mult([int(v) % q for v in x], [int(v) % q for v in y], q), dtype=np.int64)


def decrypt_exact_int(ct, s, q):
    """(b - a s) mod q centred to (-q/2, q/2] as int64
    (exp_e24_rotation_bootstrap.py:126-133)."""
    b = [int(v) % q for v in ct[0]]
    a_s = neg_conv_modq(ct[1], s, q)
    d = [(x - int(y)) % q for x, y in zip(b, a_s)]
    return np.array([v - q if v > q // 2 else v for v in d], dtype=np.int64)


def draw_uniform_a(n, q):
    """The reference's draw of the mask a from numpy's global generator:
    randint below 2^31 - 1, floor(random * q) above
    (exp_e24_rotation_bootstrap.py:217-220, exp_e11_ckks_e2e_real.py:79-82)."""
    if q <= 2 ** 31 - 1:
        return np.random.randint(0, q, size=n).astype(np.int64)
    return np.floor(np.random.random(size=n) * q).astype(np.int64)


def rlwe_encrypt_mint(m_int, s, e, q):
    """b = a s + e + m mod q for an integer message (exp_e24:209-222)."""
    a = draw_uniform_a(len(s), q)
    a_s = neg_conv_modq(a, s, q)
    b = [(int(x) + int(ei) + int(mi)) % q for x, ei, mi in zip(a_s, e, m_int)]
    return np.array(b, dtype=np.int64), a


def reencrypt(ct, s_key, q, scale, rng):
    """DARE refresh: decrypt to the scale-carrying integer message and
    encrypt it afresh with ternary noise from rng (exp_e24:287-304).
    Returns ((b, a), m_new)."""
    m_new = decrypt_exact_int(ct, s_key, q)
    e_new = rng.integers(-1, 2, size=len(s_key)).astype(np.int64)
    return rlwe_encrypt_mint(m_new, s_key, e_new, q), m_new


def decrypt_exact(ct, s, q):
    """decrypt_exact_int as float64 (exp_e24_rotation_bootstrap.py:109-123)."""
    return decrypt_exact_int(ct, s, q).astype(np.float64)


def mul_plain_exact(ct, p_poly, q):
    """(b p, a p) mod q, exact (exp_e24_rotation_bootstrap.py:136-140)."""
    return neg_conv_modq(ct[0], p_poly, q), neg_conv_modq(ct[1], p_poly, q)


def poly_x_power(nu, n):
    """The monomial X^(nu mod n) (exp_e24_rotation_bootstrap.py:225-229)."""
    p = np.zeros(n, dtype=np.int64)
    p[nu % n] = 1
    return p


def rotate_monomial(ct, nu, q):
    """ct times X^nu: a negacyclic coefficient shift
    (exp_e24_rotation_bootstrap.py:232-238, `rotate`)."""
    return mul_plain_exact(ct, poly_x_power(nu, len(ct[0])), q)


def scatter_w(msg_indices, w, scale, q, n):
    """Integer plaintext with round(scale w_j) at n - msg_j (negated at 0),
    colliding positions summed, reduced mod q
    (exp_e24_rotation_bootstrap.py:241-262)."""
    m = [0] * n
    for j, mi in enumerate(msg_indices):
        v = int(round(scale * float(w[j])))
        if mi % n == 0:
            m[0] -= v
        else:
            m[n - mi % n] += v
    return np.array([x % q for x in m], dtype=np.int64)


def inner_product(ct, w, s_key, q, scale, msg_indices=None):
    """-dec(ct * scatter_w)[0] / scale^2 = sum_j x_(msg_j) w_j
    (exp_e24_rotation_bootstrap.py:265-275)."""
    w = np.asarray(w, dtype=np.float64)
    idx = list(range(len(w))) if msg_indices is None else msg_indices
    m_w = scatter_w(idx, w, scale, q, len(ct[0]))
    return -float(decrypt_exact(mul_plain_exact(ct, m_w, q), s_key, q)[0]) / (scale * scale)


def coef_sum(ct, dim, s_key, q, scale):
    """sum_(i < dim) x_i (exp_e24_rotation_bootstrap.py:278-284)."""
    return inner_product(ct, np.ones(dim), s_key, q, scale)


def rlwe_encrypt(m_poly, s, e, q, scale):
    """b = a s + e + scale m mod q with a fractional message part kept in
    float64 (exp_e11_ckks_e2e_real.py:72-88).  a s is exact here and taken
    mod q before the float64 sum (the reference sums it unreduced)."""
    a = draw_uniform_a(len(s), q)
    a_s = neg_conv_modq(a, s, q).astype(np.float64)
    b = np.mod(a_s + np.asarray(e, dtype=np.float64) + scale * np.asarray(m_poly, dtype=np.float64), q)
    return b, a


def rlwe_decrypt(ct, s, q):
    """(b - a s) mod q centred to [-q/2, q/2) in float64: scale m + e
    (exp_e11_ckks_e2e_real.py:66-70,91-99)."""
    b = np.asarray(ct[0], dtype=np.float64)
    a_s = neg_conv_modq(ct[1], s, q).astype(np.float64)
    v = np.mod(b - a_s, q)
    return np.where(v > q / 2, v - q, v)


# ---------------------------------------------------------------------------
# noise monitor and gear policy
# ---------------------------------------------------------------------------

def noise_monitor_energy(coeffs, q):
    """E = sum min(c, q-c)^2 / n, r_hat = sqrt(E)/q
    (scripts/verify_m4_h_multilayer.py:162-174; rtl/m4/noise_monitor.sv:44-78)."""
    total = 0
    for c in coeffs:
        a = c if c <= q - c else q - c
        total += a * a
    E = total / len(coeffs)
    return E, math.sqrt(E) / q


def compute_snr(signal, noise):
    """10 log10(sum|s|^2 / sum|n|^2) (scripts/verify_m4h2_prod_calibration.py:64-69)."""
    sp = float(np.sum(np.abs(np.asarray(signal)) ** 2))
    npw = float(np.sum(np.abs(np.asarray(noise)) ** 2))
    if npw == 0:
        return math.inf
    return 10.0 * math.log10(sp / npw)


def snr_to_gear(snr_db, threshold_db=15.0):
    """>20 dB -> 3, >threshold -> 2, >10 -> 1, else 0
    (verify_m4h2_prod_calibration.py:72-82)."""
    if snr_db > 20:
        return 3
    if snr_db > threshold_db:
        return 2
    if snr_db > 10:
        return 1
    return 0


GEAR_R_MAX = (0.10, 0.20, 0.35, 0.55)   # verify_project_b_mlp_reference.py:245-250
GEAR_COST = (1.00, 0.62, 0.36, 0.20)


@dataclass
class AdaptiveGearCtrl:
    """Safety-first gear choice with hysteresis
    (verify_project_b_mlp_reference.py:279-317)."""
    gear: int = 3
    hyst_cnt: int = 0
    hyst_target: int = -1
    hyst: int = 2

    @staticmethod
    def best_gear(r_hat: float) -> int:
        for g, r in enumerate(GEAR_R_MAX):
            if r >= r_hat:
                return g
        return 0

    def step(self, r_hat: float) -> int:
        target = self.best_gear(r_hat)
        if target != self.gear:
            if self.hyst_target != target:
                self.hyst_target, self.hyst_cnt = target, 0
            else:
                self.hyst_cnt += 1
            if self.hyst_cnt >= self.hyst:
                self.gear, self.hyst_cnt, self.hyst_target = target, 0, -1
        return self.gear


@dataclass
class RtlGearCtrl:
    """One flush_valid step of rtl/m4/adaptive_gear_ctrl.sv:42-100 with
    hysteresis_en = 1 and manual_en = 0 (the register transfer, non-blocking
    semantics: the switch test reads the counter's value before the step).
    raw: energy >= t1 -> 0, >= t2 -> 1, >= t3 -> 2, else 3 (:44-50).
    A stricter raw gear is taken at once (:77-80); a looser one moves the
    candidate/counter (:66-71) and the gear switches to it when the previous
    counter had reached hyst (:72-76), i.e. on the (hyst+1)-th consecutive
    reading of one candidate.  Reset: gear 3, counter 0, candidate 3."""
    t1: float
    t2: float
    t3: float
    hyst: int = 3
    gear: int = 3
    cnt: int = 0
    cand: int = 3

    def raw(self, energy: float) -> int:
        if energy >= self.t1:
            return 0
        if energy >= self.t2:
            return 1
        if energy >= self.t3:
            return 2
        return 3

    def step(self, energy: float) -> int:
        r = self.raw(energy)
        old_gear, old_cnt = self.gear, self.cnt
        if r > old_gear:
            if r == self.cand:
                self.cnt = old_cnt + 1
            else:
                self.cand, self.cnt = r, 1
            if old_cnt >= self.hyst:
                self.gear, self.cnt, self.cand = r, 0, old_gear
        elif r < old_gear:
            self.gear, self.cnt, self.cand = r, 0, r
        else:
            self.cnt, self.cand = 0, r
        return self.gear


def fast_red(x: int, q: int, k: int = 44) -> int:
    """x mod q for q = 2^k + c with small c, by the reference's three rounds
    of 2^k = -c folding (scripts/_verify_fastred.py:55-99):
    r1 = xl - c xh; r1 = h1 2^k + l1 (Euclidean); r2 = l1 - c h1;
    r2 = h2 2^k + l2; r3 = l2 - c h2; one conditional +q / -q.
    The reference fixes q = 2^44 + 131073 with 18-bit c; the same three
    rounds are exact for every c with c^2 < 2^k here (asserted)."""
    c = q - (1 << k)
    assert 0 < c and c * c < (1 << k) and 0 <= x < (1 << (2 * k + 2))
    mask = (1 << k) - 1
    r1 = (x & mask) - (x >> k) * c
    h1 = r1 >> k                       # floor, may be negative
    l1 = r1 - (h1 << k)
    r2 = l1 - h1 * c                   # -h1 >= 0
    assert r2 >= 0
    h2 = r2 >> k
    l2 = r2 - (h2 << k)
    r = l2 - h2 * c
    if r < 0:
        r += q
    if r >= q:
        r -= q
    assert 0 <= r < q
    return r


def apply_block_sparse_mask(weight_matrix, threshold_val, B):
    """Zero every full run of B flattened weights whose max |w| is below
    threshold_val; a ragged tail is never pruned.  Returns (masked,
    skipped, blocks, kept share) (exp_e10_sparsity_sweep_v2.py:102-124)."""
    flat = np.array(weight_matrix, dtype=np.float64).ravel()
    out = flat.copy()
    nb = flat.size // B
    skip = 0
    for i in range(nb):
        if np.max(np.abs(flat[i * B:(i + 1) * B])) < threshold_val:
            out[i * B:(i + 1) * B] = 0.0
            skip += 1
    return out.reshape(np.shape(weight_matrix)), skip, nb, (nb - skip) / nb


def block_sparse_mask(weight_matrix, threshold_frac, B):
    """The same rule with the threshold threshold_frac * max |w|; returns
    (masked, skipped, blocks) (exp_e7_sparsity_sweep.py:72-90)."""
    t = float(np.max(np.abs(np.asarray(weight_matrix, dtype=np.float64)))) * threshold_frac
    w, skip, nb, _ = apply_block_sparse_mask(weight_matrix, t, B)
    return w, skip, nb


def block_prune_mask(w: np.ndarray, frac: float, block: int = 64) -> np.ndarray:
    """Zero the ``frac`` share of size-``block`` blocks (row-major runs of
    the flattened matrix) with the smallest max-abs -- the quantile form of
    the max|w| < threshold rule (scripts/exp_e10_sparsity_sweep_v2.py:102-124)."""
    flat = np.abs(np.asarray(w, dtype=np.float64)).ravel()
    nb = (flat.size + block - 1) // block
    pad = np.zeros(nb * block)
    pad[: flat.size] = flat
    bmax = pad.reshape(nb, block).max(axis=1)
    k = int(round(frac * nb))
    keep = np.ones(nb, dtype=bool)
    if k > 0:
        keep[np.argsort(bmax, kind="stable")[:k]] = False
    return np.repeat(keep, block)[: flat.size].reshape(np.shape(w))
