"""Pins the CPU oracle against the reference's known answers and fixtures,
against closed forms, and the C restatement against the Python one.  CPU only."""
import hashlib
import math
import random

import numpy as np
import pytest

import ckks_oracle as O
import oracle_c as OC

CHAIN4 = [17592186634241, 17592186372097, 17592186273793, 17592186175489]
SPECIAL_P = 17592186765313


def _digest(vals):
    return hashlib.sha256(np.asarray(vals, dtype="<u8").tobytes()).hexdigest()


# ---------------------------------------------------------------- KATs

def test_kat_ntt_n8(fixtures):
    k = fixtures["kats"]["ntt_n8"]
    assert O.bit_rev_perm(k["input"], 8) == k["bitrev_input"]
    omega = fixtures["tables"]["rtl/m5/tw/tw_omega_toy_n8.hex"]["values"][1]
    assert omega == 18
    assert O.ntt(k["input"], omega, 8, k["q"]) == k["expected"]
    assert list(OC.ntt_cyclic(k["input"], 8, k["q"], omega)) == k["expected"]


def test_kat_polymult_n8(fixtures):
    k = fixtures["kats"]["pm_n8"]
    q, w = k["q"], 18
    wi = pow(w, -1, q)
    assert O.poly_mult_via_ntt(k["a"], k["b"], 8, q, w, wi) == k["expected"]
    assert O.cyclic_conv_direct(k["a"], k["b"], 8, q) == k["expected"]


def test_kat_polymult_n16(fixtures):
    k = fixtures["kats"]["pm_n16"]
    q = k["q"]
    tab = fixtures["tables"]["rtl/m5/tw/tw_omega_toy_n16.hex"]["values"]
    itab = fixtures["tables"]["rtl/m5/tw/tw_omegainv_toy_n16.hex"]["values"]
    w, wi = tab[1], itab[1]
    assert O.power_table(w, 8, q) == tab and O.power_table(wi, 8, q) == itab
    assert pow(16, -1, q) == k["n_inv"]
    assert O.poly_mult_via_ntt(k["a"], k["b"], 16, q, w, wi) == k["expected"]
    c = OC.intt_cyclic(OC.ntt_cyclic(k["a"], 16, q, w) * OC.ntt_cyclic(k["b"], 16, q, w) % q, 16, q, wi)
    assert list(c) == k["expected"]


# ---------------------------------------------------------------- chain + twiddle fixtures

@pytest.mark.parametrize("i", range(4))
def test_chain4_roots_and_tables(fixtures, i):
    e = fixtures["chain4"][i]
    q = e["q"]
    assert q == CHAIN4[i] and O.is_prime(q) and (q - 1) % (1 << 15) == 0
    assert O.primitive_root(q) == e["primitive_root"]
    w = O.root_of_unity(q, 16384)
    assert w == e["omega"] and pow(w, -1, q) == e["omega_inv"]
    assert pow(16384, -1, q) == e["N_INV"]
    t = fixtures["tables"][f"rtl/m2/omega_table_chain4_q{i}.hex"]
    ti = fixtures["tables"][f"rtl/m2/omega_inv_table_chain4_q{i}.hex"]
    assert _digest(O.power_table(w, t["count"], q)) == t["sha256"]
    assert _digest(O.power_table(pow(w, -1, q), ti["count"], q)) == ti["sha256"]
    # the table is exactly the psi^k pre-twist of an N=8192 negacyclic transform
    assert pow(w, 8192, q) == q - 1


def test_prod_tables_are_q3(fixtures):
    q = CHAIN4[3]
    w = O.root_of_unity(q, 16384)
    assert fixtures["tables"]["rtl/m2/omega_table_prod.hex"]["sha256"] == _digest(O.power_table(w, 8192, q))


@pytest.mark.parametrize("name", ["p1", "p2", "p3", "q3"])
@pytest.mark.parametrize("n", [16, 4096, 16384])
def test_rns_twiddle_roms(fixtures, name, n):
    key = f"rtl/m5/tw/tw_omega_{name}_n{n}.hex"
    if key not in fixtures["tables"]:
        pytest.skip("table not shipped by the reference")
    primes = {"p1": 77309411329, "p2": 206158430209, "p3": 2061584302081, "q3": CHAIN4[3]}
    q = primes[name]
    t = fixtures["tables"][key]
    ti = fixtures["tables"][f"rtl/m5/tw/tw_omegainv_{name}_n{n}.hex"]
    assert t["count"] == n // 2
    w = (t.get("values") or t["head"])[1]
    wi = (ti.get("values") or ti["head"])[1]
    assert O.is_prime(q) and w * wi % q == 1
    assert pow(w, n, q) == 1 and pow(w, n // 2, q) == q - 1   # primitive n-th root
    assert _digest(O.power_table(w, n // 2, q)) == t["sha256"]
    assert _digest(O.power_table(wi, n // 2, q)) == ti["sha256"]


def test_single_prime_q_and_psi():
    q = O.largest_prime_1_mod(45, 8192)
    assert q == 35184371884033 and q < 1 << 45
    # v2(q-1) = 13: psi of order 8192 exists, 16384 does not
    assert (q - 1) % 8192 == 0 and (q - 1) % 16384 != 0
    psi = O.smallest_primitive_root_of_unity(q, 8192)
    assert psi == 7388705119
    assert pow(psi, 4096, q) == q - 1


def test_special_prime():
    assert O.is_prime(SPECIAL_P) and (SPECIAL_P - 1) % (1 << 15) == 0
    assert SPECIAL_P not in CHAIN4 and SPECIAL_P < 1 << 45
    assert O.primes_2k_plus(45, 32768, 1, 22)[0] == SPECIAL_P


def test_chain5_extra_prime():
    """The fifth ciphertext prime is the next prime 2^44 + 32768 j + 1 above
    the special prime (j = 48), brute force over every j in between."""
    from cipherpu.params import CHAIN5
    extra = CHAIN5[4]
    assert CHAIN5[:4] == tuple(CHAIN4)
    assert O.is_prime(extra) and (extra - 1) % (1 << 15) == 0 and extra < 1 << 45
    js = [j for j in range(23, 49) if O.is_prime(2 ** 44 + 32768 * j + 1)]
    assert js == [48] and extra == 2 ** 44 + 32768 * 48 + 1
    # every prime of the set is below twice every other one (the key switch
    # then skips reducing a digit's residues into another prime)
    allp = list(CHAIN5) + [SPECIAL_P]
    assert max(allp) < 2 * min(allp)


# ---------------------------------------------------------------- transforms

def test_ntt_closed_form_and_roundtrip():
    rng = random.Random(5)
    q = 7681   # 7681 = 15*512 + 1
    for N in (8, 16, 32):
        w = O.root_of_unity(q, N)
        a = [rng.randrange(q) for _ in range(N)]
        X = O.ntt(a, w, N, q)
        assert X == [sum(a[j] * pow(w, j * k, q) for j in range(N)) % q for k in range(N)]
        assert O.intt(X, pow(w, -1, q), N, q) == a
        psi = O.root_of_unity(q, 2 * N)
        Xn = O.negacyclic_ntt(a, psi, N, q)
        assert Xn == [sum(a[j] * pow(psi, (2 * k + 1) * j, q) for j in range(N)) % q for k in range(N)]
        assert O.negacyclic_intt(Xn, psi, N, q) == a


def test_negacyclic_schoolbook_monomials():
    q, N = 97, 8
    x = [0] * N
    x[1] = 1
    f = list(range(1, N + 1))
    # f * X: coefficients shift up, the top one wraps with a sign flip
    assert O.poly_mult(f, x, q) == [(-f[-1]) % q] + f[:-1]
    xn1 = [0] * N
    xn1[N - 1] = 1
    assert O.poly_mult(x, xn1, q) == [q - 1] + [0] * (N - 1)   # X^N = -1


@pytest.mark.parametrize("N", [16, 64, 256])
def test_polymult_ntt_vs_schoolbook(N):
    rng = random.Random(N)
    q = CHAIN4[0]
    psi = O.root_of_unity(q, 2 * N)
    a = [rng.randrange(q) for _ in range(N)]
    b = [rng.randrange(q) for _ in range(N)]
    ref = O.poly_mult(a, b, q)
    assert O.poly_mult_ntt(a, b, psi, q) == ref
    assert list(OC.polymul_schoolbook(a, b, q)) == ref
    assert list(OC.polymul_ntt(a, b, q, psi)) == ref


def test_c_oracle_transforms_match_python():
    rng = random.Random(11)
    q = CHAIN4[2]
    N = 1024
    psi = O.root_of_unity(q, 2 * N)
    a = [rng.randrange(q) for _ in range(N)]
    assert list(OC.ntt_negacyclic(a, N, q, psi)) == O.negacyclic_ntt(a, psi, N, q)
    w = psi * psi % q
    assert list(OC.ntt_cyclic(a, N, q, w)) == O.ntt(a, w, N, q)


def test_c_schoolbook_vs_ntt_4096():
    """The N=4096 dual-path check of gen_poly_mult_vectors_n12.py:123-143,
    negacyclic form, on the single-prime product's prime."""
    q, psi = 35184371884033, 7388705119
    r = np.random.default_rng(1)
    a = r.integers(0, q, 4096, dtype=np.uint64)
    b = r.integers(0, q, 4096, dtype=np.uint64)
    assert np.array_equal(OC.polymul_schoolbook(a, b, q), OC.polymul_ntt(a, b, q, psi))


# ---------------------------------------------------------------- rescale

@pytest.mark.parametrize("i", range(3))
def test_rescale_coeff_invariants(i):
    """verify_rescale.py:60-128: range, round trip within 2, monotonicity."""
    rng = random.Random(i)
    q_in, q_out = CHAIN4[i], CHAIN4[i + 1]
    pts = [0, 1, 2, q_in - 3, q_in - 2, q_in - 1, q_out - 1, q_out, q_out + 1]
    for c in pts:
        if 0 <= c < q_in:
            assert 0 <= O.rescale_coeff(c, q_in, q_out) < q_out
    for _ in range(2000):
        c = rng.randrange(q_in)
        o = O.rescale_coeff(c, q_in, q_out)
        assert 0 <= o < q_out
        assert abs(round(o * q_in / q_out) - c) <= 2
        # closed form: nearest integer to c*q_out/q_in, ties up
        assert o == (2 * c * q_out + q_in) // (2 * q_in) % q_out
    for _ in range(500):
        a, b = sorted((rng.randrange(q_in), rng.randrange(q_in)))
        assert O.rescale_coeff(a, q_in, q_out) <= O.rescale_coeff(b, q_in, q_out)


def test_rescale_toy_against_rescale_coeff():
    """The toy rescale of verify_e2e_mult.py:70-75 against what fixes it
    independently of its formula: at scale_bits 0 and odd q_in it is the
    pinned rescale_coeff (floor((x + floor(q/2)) / q) rounds x / q half up for
    odd and even q alike); below q_in it is rescale_coeff >> scale_bits,
    since floor(floor(x / a) / b) = floor(x / (a b)) and nothing wraps mod
    q_out.  A dropped q_in // 2 or a shift taken before the division fails."""
    rng = random.Random(11)
    for i in range(3):
        q_in, q_out = CHAIN4[i], CHAIN4[i + 1]
        for s in [0, 1, q_in // 2, q_in - 1] + [rng.randrange(q_in) for _ in range(500)]:
            r = O.rescale_coeff(s, q_in, q_out)
            assert O.rescale_toy(s, q_in, q_out, 0) == r
            for sb in (1, 13, 30, 44):
                assert O.rescale_toy(s, q_in, q_out, sb) == r >> sb
    for q_in, q_out in ((2, 3), (4, 7), (113, 113), (6, 5)):
        for s in range(3 * q_in):
            assert O.rescale_toy(s, q_in, q_out, 0) == O.rescale_coeff(s, q_in, q_out)


def test_e2e_mult_decrypts_to_the_rescaled_product():
    """verify_e2e_mult.py:78-102 composed by the oracle: with the E33 key
    b = a s + s^2 (no error, so the single-modulus relinearisation is exact)
    and E33 ciphertexts, f0' + f1' s decrypts to round(m1 m2 q_out / q_in)
    within the rounding of f0', f1' (1/2 + ||s||_1 / 2).  The exact
    (scale_bits None) and toy (0) rescales agree at odd q_in.  A flipped
    relinearisation sign or a swapped key half fails by ~q."""
    rng = random.Random(17)
    n, q_in, q_out = 16, CHAIN4[0], CHAIN4[1]
    s = [rng.randrange(-1, 2) for _ in range(n)]
    sq = [v % q_in for v in s]
    s2 = O.poly_mult(sq, sq, q_in)
    a_k = [rng.randrange(q_in) for _ in range(n)]
    b_k = O.poly_add(O.poly_mult(a_k, sq, q_in), s2, q_in)

    def enc(m):
        a = [rng.randrange(q_in) for _ in range(n)]
        c0 = O.poly_add(O.poly_mult(a, sq, q_in), [v % q_in for v in m], q_in)
        return c0, [(-v) % q_in for v in a]

    for _ in range(4):
        m1 = [rng.randrange(-2 ** 16, 2 ** 16) for _ in range(n)]
        m2 = [rng.randrange(-2 ** 16, 2 ** 16) for _ in range(n)]
        (c01, c11), (c02, c12) = enc(m1), enc(m2)
        f0, f1 = O.e2e_mult(c01, c11, c02, c12, a_k, b_k, q_in, q_out, None)
        assert (f0, f1) == O.e2e_mult(c01, c11, c02, c12, a_k, b_k, q_in, q_out, 0)
        d = O.poly_add(f0, O.poly_mult(f1, [v % q_out for v in s], q_out), q_out)
        prod = O.poly_mult([v % q_in for v in m1], [v % q_in for v in m2], q_in)
        for got, p in zip(d, prod):
            p = p - q_in if p > q_in // 2 else p
            got = got - q_out if got > q_out // 2 else got
            assert abs(got - p * q_out / q_in) <= 0.5 + sum(map(abs, s)) / 2 + 1e-6
        bad = O.poly_add(O.poly_mult(c11, c12, q_in), [0] * n, q_in)
        f1_bad = O.poly_add(O.poly_add(O.poly_mult(c01, c12, q_in), O.poly_mult(c11, c02, q_in), q_in),
                            O.poly_mult(bad, a_k, q_in), q_in)
        assert f1_bad != O.relin([0] * n, O.poly_add(O.poly_mult(c01, c12, q_in), O.poly_mult(c11, c02, q_in),
                                                      q_in), bad, a_k, b_k, q_in)[1]


def test_rns_rescale_identity():
    rng = random.Random(3)
    N = 64
    for l in (2, 3, 4):
        primes = CHAIN4[:l]
        limbs = [[rng.randrange(q) for _ in range(N)] for q in primes]
        a = O.rns_rescale(limbs, primes)
        assert a == O.rns_rescale_per_limb(limbs, primes)
        assert OC.rescale(limbs, primes).tolist() == a
        # exact division check: out * q_L ~= CRT value (centered difference <= q_L/2)
        Q = math.prod(primes)
        vals = O.crt_compose(limbs, primes)
        outs = O.crt_compose(a, primes[:-1]) if l > 2 else a[0]
        for v, o in zip(vals, outs):
            assert abs(v - o * primes[-1]) <= primes[-1] // 2 or abs(v - o * primes[-1] - Q) <= primes[-1] // 2


def test_mod_down_identity():
    rng = random.Random(4)
    N = 32
    primes = CHAIN4
    limbs = [[rng.randrange(q) for _ in range(N)] for q in primes + [SPECIAL_P]]
    a = O.mod_down_round(limbs, primes, SPECIAL_P)
    assert OC.mod_down(limbs, primes, SPECIAL_P).tolist() == a


# ---------------------------------------------------------------- key switch / HMult

def _keygen(rng, N, primes, p, s_from, s):
    """Hybrid key on primes + [p] in coefficient form (oracle side):
    b_j = a_j s + e_j + g_j s_from."""
    basis = primes + [p]
    psis = [O.root_of_unity(q, 2 * N) for q in basis]
    key = []
    for j in range(len(primes)):
        e = [rng.choice([-2, -1, 0, 1, 2]) for _ in range(N)]
        bj, aj = [], []
        for i, qi in enumerate(basis):
            a = [rng.randrange(qi) for _ in range(N)]
            b = O.poly_mult_ntt(a, [x % qi for x in s], psis[i], qi)
            b = O.poly_add(b, [x % qi for x in e], qi)
            g = O.gadget_factor(primes, p, j, i)
            if g:
                b = O.poly_add(b, [g * (x % qi) % qi for x in s_from], qi)
            bj.append(b)
            aj.append(a)
        key.append((bj, aj))
    return key, psis


def test_keyswitch_bigint_route():
    """ModDown(sum_j D_j b_j) through per-limb NTT products equals the
    big-int route: integer schoolbook products over Z, reduce mod Q*p,
    divide by p with round-half-up."""
    rng = random.Random(9)
    N = 16
    primes = CHAIN4[:2]
    p = SPECIAL_P
    s = [rng.choice([-1, 0, 1]) for _ in range(N)]
    key, psis = _keygen(rng, N, primes, p, O.poly_mult(s, s, 1 << 200), s)
    d2 = [[rng.randrange(q) for _ in range(N)] for q in primes]
    f0, f1 = O.keyswitch_hybrid(d2, key, primes, p, psis, N)
    QP = primes[0] * primes[1] * p
    basis = primes + [p]
    acc0 = [0] * N
    acc1 = [0] * N
    for j in range(2):
        bj = O.crt_compose(key[j][0], basis)
        aj = O.crt_compose(key[j][1], basis)
        pb = O.poly_mult(d2[j], bj, QP)
        pa = O.poly_mult(d2[j], aj, QP)
        acc0 = [(x + y) % QP for x, y in zip(acc0, pb)]
        acc1 = [(x - y) % QP for x, y in zip(acc1, pa)]
    Q = primes[0] * primes[1]
    r0 = [O.rescale_coeff(v, QP, Q) for v in acc0]
    r1 = [O.rescale_coeff(v, QP, Q) for v in acc1]
    assert O.crt_compose(f0, primes) == r0
    assert O.crt_compose(f1, primes) == r1


def test_hmult_decrypt_invariant_and_c_parity():
    """dec(HMult(ct_a, ct_b)) ~= dec(ct_a) * dec(ct_b) / q_L: the tensor,
    key switch and rescale together preserve the message; a dropped term or
    wrong sign breaks the bound by orders of magnitude.  Also C == Python."""
    rng = random.Random(21)
    N, l = 32, 3
    primes = CHAIN4[:l]
    p = SPECIAL_P
    s = [rng.choice([-1, 0, 1]) for _ in range(N)]
    s2 = O.poly_mult(s, s, 1 << 200)
    s2 = [x if x < (1 << 199) else x - (1 << 200) for x in s2]
    key, psis = _keygen(rng, N, primes, p, s2, s)

    def enc(m):
        c0, c1 = [], []
        e = [rng.choice([-3, -1, 0, 1, 3]) for _ in range(N)]
        for i, qi in enumerate(primes):
            a = [rng.randrange(qi) for _ in range(N)]
            as_ = O.poly_mult_ntt(a, [x % qi for x in s], psis[i], qi)
            c0.append([(x + y + z) % qi for x, y, z in zip(as_, e, m)])
            c1.append([(-x) % qi for x in a])
        return [c0, c1]

    m1 = [rng.randrange(-(1 << 40), 1 << 40) for _ in range(N)]
    m2 = [rng.randrange(-(1 << 40), 1 << 40) for _ in range(N)]
    cta, ctb = enc(m1), enc(m2)
    out = O.hmult(cta, ctb, key, primes, p, psis, N)
    keyarr = np.array([[kb, ka] for kb, ka in key], dtype=np.uint64)
    c_out = OC.hmult(np.array(cta, dtype=np.uint64), np.array(ctb, dtype=np.uint64), keyarr,
                     primes, p, psis[:l], psis[l])
    assert c_out.tolist() == [out[0], out[1]]
    dec = O.decrypt(out, [[x % q for x in s] for q in primes[:-1]], primes[:-1], psis)
    Q2 = primes[0] * primes[1]
    vals = [v - Q2 if v > Q2 // 2 else v for v in O.crt_compose(dec, primes[:-1])]
    want = O.poly_mult(m1, m2, 1 << 400)
    want = [x if x < (1 << 399) else x - (1 << 400) for x in want]
    err = max(abs(v - w / primes[-1]) for v, w in zip(vals, want))
    assert err < 2 ** 20, err


def test_automorphism_closed_form():
    q, N = 97, 16
    # X^k -> X^(g k) with sign (-1)^floor(gk/N)
    for g in (3, 5, 25 % 32, 31):
        for k in range(N):
            f = [0] * N
            f[k] = 1
            out = O.automorphism(f, g, N, q)
            t = g * k % (2 * N)
            exp = [0] * N
            exp[t % N] = 1 if t < N else q - 1
            assert out == exp
            assert list(OC.automorphism(f, q, g)) == exp


def test_rotation_moves_slots():
    N = 32
    rng = np.random.default_rng(0)
    z = rng.uniform(-1, 1, N // 2) + 1j * rng.uniform(-1, 1, N // 2)
    m = O.encode_fft(z, N)
    for r in (1, 2, 5):
        g = O.galois_elt(r, N)
        # automorphism on real coefficients (float path of exp_e24:143-162)
        mm = np.zeros(N)
        for k in range(N):
            t = g * k % (2 * N)
            mm[t % N] += m[k] if t < N else -m[k]
        np.testing.assert_allclose(O.decode_fft(mm, N), np.roll(z, -r), atol=1e-9)


# ---------------------------------------------------------------- encode / decode

def test_embedding_fft_vs_direct():
    N = 64
    rng = np.random.default_rng(2)
    m = rng.normal(size=N)
    np.testing.assert_allclose(O.decode_fft(m, N), O.embed_direct(m, N), atol=1e-10)
    z = rng.normal(size=N // 2) + 1j * rng.normal(size=N // 2)
    np.testing.assert_allclose(O.embed_direct(O.encode_fft(z, N), N), z, atol=1e-10)


def test_embedding_simd_multiplicative():
    """verify_multi_slot_simd.py:113-168: the negacyclic product of encodings
    decodes to the slot-wise product (and commutes / associates)."""
    N = 16
    rng = np.random.default_rng(3)
    z1 = rng.uniform(-1, 1, N // 2)
    z2 = rng.uniform(-1, 1, N // 2)
    m1, m2 = O.encode_fft(z1, N), O.encode_fft(z2, N)
    prod = np.zeros(2 * N)
    for i in range(N):
        prod[i:i + N] += m1[i] * m2
    prod = prod[:N] - prod[N:]
    np.testing.assert_allclose(O.decode_fft(prod, N), z1 * z2, atol=1e-10)


def test_natural_embedding_matches_reference_semantics():
    v = np.array([1.0, 2.0, 3.0, 4.0])
    m = O.encode_natural(v, 2.0, 8)
    np.testing.assert_allclose(O.decode_natural(m, 2.0, 8, 4), v, atol=1e-9)
    back = O.neg_ifft(O.neg_fft(np.array([1.0, 2.0, -1.0, 0.5, 0.0, 0.25, -0.5, 1.5])))
    np.testing.assert_allclose(back.real, [1.0, 2.0, -1.0, 0.5, 0.0, 0.25, -0.5, 1.5], atol=1e-12)


def test_encode_rns_roundtrip():
    N = 64
    primes = CHAIN4[:2]
    z = np.random.default_rng(4).uniform(-1, 1, N // 2)
    limbs = O.encode_rns(z, 2.0 ** 40, N, primes)
    np.testing.assert_allclose(O.decode_rns(limbs, 2.0 ** 40, N, primes).real, z, atol=1e-9)


# ---------------------------------------------------------------- monitor / gears

def test_noise_monitor_energy_closed_form():
    q = 101
    E, r = O.noise_monitor_energy([0, 1, 100, 50, 51], q)
    assert E == (0 + 1 + 1 + 2500 + 2500) / 5
    assert r == math.sqrt(E) / q


def test_snr_and_gears():
    assert O.compute_snr([3, 4], [0, 0]) == math.inf
    assert abs(O.compute_snr([10.0], [1.0]) - 20.0) < 1e-12
    assert [O.snr_to_gear(x) for x in (25, 20, 16, 15, 11, 10, 0)] == [3, 2, 2, 1, 1, 0, 0]
    c = O.AdaptiveGearCtrl()
    seq = [c.step(r) for r in (0.5, 0.5, 0.5, 0.05, 0.05, 0.05)]
    assert seq == [3, 3, 3, 3, 3, 0]
    assert O.AdaptiveGearCtrl.best_gear(0.9) == 0
    # adaptive_gear_ctrl.sv traced by hand, register by register (HYST_CNT 3):
    # 600 -> raw 0, down at once; three readings of raw 3 only count 1, 2, 3;
    # the next looser reading (raw 1) switches, to raw 1, on the old count 3.
    rc = O.RtlGearCtrl(500, 200, 50)
    assert [rc.step(e) for e in (600, 10, 10, 10, 300)] == [0, 0, 0, 0, 1]
    assert (rc.cnt, rc.cand) == (0, 0)
    # a change of candidate restarts the count at 1
    rc = O.RtlGearCtrl(500, 200, 50)
    assert [rc.step(e) for e in (600, 100, 100, 10, 10, 10, 10)] == [0, 0, 0, 0, 0, 0, 3]
    # readings at the current gear clear the count
    rc = O.RtlGearCtrl(500, 200, 50)
    assert [rc.step(e) for e in (600, 10, 10, 600, 10, 10, 10, 10)] == [0, 0, 0, 0, 0, 0, 0, 3]


def test_block_prune_mask():
    w = np.arange(256, dtype=float).reshape(2, 128)
    m = O.block_prune_mask(w, 0.5, 64)
    assert m.sum() == 128 and not m.ravel()[:128].any() and m.ravel()[128:].all()


def test_fast_red_reference_cases():
    """The special-form reduction against x % q on the reference's own edge
    set for q = 2^44 + 131073 (scripts/_verify_fastred.py:108-118) and its
    hand checks (:48-51), then on every chain prime with random products."""
    import random
    Q, C = 17592186175489, 131073
    assert Q == CHAIN4[3] and Q - (1 << 44) == C
    assert O.fast_red(Q, Q) == 0 and O.fast_red(Q - 1, Q) == Q - 1
    assert O.fast_red(0, Q) == 0 and O.fast_red(1, Q) == 1
    edges = [0, 1, 2, Q - 1, Q, Q + 1, 2 * Q, 3 * Q, 2 * Q - 1, 1 << 44, (1 << 44) - 1, (1 << 44) + 1,
             C, 2 * C, 3 * C, (1 << 90) - 1, (1 << 90) - Q, Q * ((1 << 44) - 1), (1 << 45) * C, (1 << 46) * C,
             (1 << 46) * C - 1, (1 << 64) + (1 << 44), ((1 << 44) - 1) * ((1 << 18) - 1)]
    for e in edges:
        assert O.fast_red(e, Q) == e % Q, e
    rng = random.Random(20261003)
    from cipherpu.params import CHAIN5
    for q in list(CHAIN5) + [SPECIAL_P]:
        for _ in range(3000):
            a, b = rng.randrange(q), rng.randrange(q)
            assert O.fast_red(a * b, q) == a * b % q


def test_eval_domain_automorphism_permutation():
    """Position i of the (bit-reversed) forward transform holds
    a(psi^(2 brv(i) + 1)); the automorphism X -> X^g then permutes positions
    by j(i) = brv(((g (2 brv(i) + 1)) mod 2N - 1) / 2) -- the map the hoisted
    rotation applies in the evaluation domain."""
    import random
    N, q = 64, 7681          # 2N | q - 1
    psi = O.smallest_primitive_root_of_unity(q, 2 * N)
    logn = N.bit_length() - 1
    brv = lambda i: int(format(i, f"0{logn}b")[::-1], 2)
    rng = random.Random(3)
    a = [rng.randrange(q) for _ in range(N)]
    A = O.negacyclic_ntt(a, psi, N, q)                   # natural order: A[k] = a(psi^(2k+1))
    for k in range(N):
        assert A[k] == sum(x * pow(psi, (2 * k + 1) * j, q) for j, x in enumerate(a)) % q
    A_br = [A[brv(i)] for i in range(N)]
    for g in (5, 25, 2 * N - 1, pow(5, 7, 2 * N)):
        B = O.negacyclic_ntt(O.automorphism(a, g, N, q), psi, N, q)
        B_br = [B[brv(i)] for i in range(N)]
        for i in range(N):
            e = (g * (2 * brv(i) + 1)) % (2 * N)
            assert B_br[i] == A_br[brv((e - 1) // 2)]


def test_rotate_hoisted_oracle_decrypts():
    """The hoisted rotation's oracle (digits = automorphism over Z of the
    canonical lifts) decrypts to the rotated message like the one-step
    rotation, with noise of the same size, though not bit-identically."""
    import oracle_c as OC
    N, l = 1024, 2
    pr, p = list(CHAIN4[:l]), SPECIAL_P
    psis = [O.smallest_primitive_root_of_unity(q, 2 * N) for q in pr]
    pp = O.smallest_primitive_root_of_unity(p, 2 * N)
    rng = np.random.default_rng(1)
    s = rng.integers(-1, 2, N)
    rns = lambda x, q: (np.asarray(x, dtype=object) % q).astype(np.uint64)
    g = pow(5, 3, 2 * N)
    sg = np.zeros(N, dtype=np.int64)
    for k in range(N):
        t = (g * k) % (2 * N)
        sg[t % N] += s[k] if t < N else -s[k]
    allq = pr + [p]
    key = np.zeros((l, 2, l + 1, N), dtype=np.uint64)
    for j in range(l):
        ej = rng.integers(-3, 4, N)          # one error polynomial per digit, in every limb
        for i, q in enumerate(allq):
            a = rng.integers(0, q, N, dtype=np.uint64)
            psi = psis[i] if i < l else pp
            b = (OC.polymul_ntt(a, rns(s, q), q, psi).astype(object) + rns(ej, q)) % q
            if i == j:
                b = (b + (p % q) * rns(sg, q).astype(object)) % q
            key[j, 0, i], key[j, 1, i] = b.astype(np.uint64), a
    m = rng.integers(-1000, 1000, N)
    a = [rng.integers(0, q, N, dtype=np.uint64) for q in pr]
    e = rng.integers(-3, 4, N)
    c0 = np.stack([((OC.polymul_ntt(a[i], rns(s, q), q, psis[i]).astype(object) + rns(e + m, q)) % q)
                   .astype(np.uint64) for i, q in enumerate(pr)])
    c1 = np.stack([((q - a[i].astype(object)) % q).astype(np.uint64) for i, q in enumerate(pr)])
    ct = np.stack([c0, c1])
    q0 = pr[0]

    def dec(c):
        v = (c[0, 0].astype(object) + OC.polymul_ntt(c[1, 0], rns(s, q0), q0, psis[0]).astype(object)) % q0
        return np.array([int(x) - q0 if x > q0 // 2 else int(x) for x in v])
    mrot = np.zeros(N, dtype=np.int64)
    for k in range(N):
        t = (g * k) % (2 * N)
        mrot[t % N] += m[k] if t < N else -m[k]
    one = OC.rotate(ct, key, pr, p, psis, pp, g)
    hoisted = OC.rotate_hoisted(ct, key, pr, p, psis, pp, g)
    assert np.max(np.abs(dec(one) - mrot)) < 2000
    assert np.max(np.abs(dec(hoisted) - mrot)) < 2000
    assert not np.array_equal(one, hoisted)


def test_single_modulus_rlwe_helpers():
    """exp_e24/exp_e11 RLWE helpers: the mask is numpy's global draw (same
    seed, same a), b - a s recovers e + m exactly (a dropped term or sign
    fails), DARE's refresh decrypts to the old decryption plus ternary
    noise, and the float64 pair round-trips scale m + e."""
    for q, n in ((998244353, 64), (17592186634241, 32)):
        rng = np.random.default_rng(q % 1000)
        s = rng.integers(-1, 2, n)
        e = rng.integers(-3, 4, n)
        m = rng.integers(-10 ** 6, 10 ** 6, n)
        np.random.seed(5)
        b, a = O.rlwe_encrypt_mint(m, s, e, q)
        np.random.seed(5)
        want_a = (np.random.randint(0, q, size=n) if q <= 2 ** 31 - 1
                  else np.floor(np.random.random(size=n) * q)).astype(np.int64)
        assert np.array_equal(a, want_a)
        a_s = O.poly_mult([int(v) for v in a], [int(v) % q for v in s], q)
        assert [(int(x) - y) % q for x, y in zip(b, a_s)] == [int(v) % q for v in e + m]
        assert np.array_equal(O.decrypt_exact_int((b, a), s, q), e + m)
        (b2, a2), m_new = O.reencrypt((b, a), s, q, 2.0 ** 20, np.random.default_rng(9))
        assert np.array_equal(m_new, e + m)
        noise = O.decrypt_exact_int((b2, a2), s, q) - m_new
        assert set(np.unique(noise)) <= {-1, 0, 1}
        assert np.array_equal(noise, np.random.default_rng(9).integers(-1, 2, size=n))
        scale = 2.0 ** 12
        mf = rng.uniform(-1, 1, n)
        bf, af = O.rlwe_encrypt(mf, s, e, q, scale)
        # float64 sums of values up to q carry rounding of q 2^-52 per step
        assert np.max(np.abs(O.rlwe_decrypt((bf, af), s, q) - (scale * mf + e))) < q * 2.0 ** -49


def test_coefficient_encoding_primitives():
    """exp_e24's coefficient-encoding primitives: X^nu multiplication is the
    negacyclic shift, scatter_w places and sums weights as traced by hand,
    and inner_product / coef_sum decrypt to the plain dot product and sum
    (an index or sign slip in the placement fails them)."""
    q, n = 17592186634241, 32
    rng = np.random.default_rng(11)
    c = (rng.integers(0, q, n), rng.integers(0, q, n))
    for nu in (0, 1, 5, 31, 33):
        r = O.rotate_monomial(c, nu, q)
        for src, dst in ((c[0], r[0]), (c[1], r[1])):
            want = [0] * n
            for j in range(n):
                k = j + nu % n
                want[k % n] = int(src[j]) if k < n else (-int(src[j])) % q
            assert list(map(int, dst)) == want
    assert list(O.scatter_w([0, 1, 1, 3], [0.5, 0.25, 0.25, -1.0], 4.0, 97, 8)) == [95, 0, 0, 0, 0, 93, 0, 2]
    scale = 2.0 ** 14
    s = rng.integers(-1, 2, n)
    e = rng.integers(-3, 4, n)
    x = rng.uniform(-1, 1, n)
    np.random.seed(4)
    ct = O.rlwe_encrypt_mint(np.round(scale * x).astype(np.int64), s, e, q)
    w = rng.uniform(-1, 1, 20)
    assert abs(O.inner_product(ct, w, s, q, scale) - float(x[:20] @ w)) < 0.02
    idx = [3, 7, 7, 0, 31]
    assert abs(O.inner_product(ct, w[:5], s, q, scale, idx) - float(sum(x[i] * wj for i, wj in zip(idx, w[:5])))) < 0.02
    assert abs(O.coef_sum(ct, 25, s, q, scale) - float(x[:25].sum())) < 0.02
