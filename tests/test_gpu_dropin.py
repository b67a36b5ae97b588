"""GPU parity of the drop-in surface at the reference's own call shapes.

Each test calls a golden_api function (or the C ABI under it) the way the
reference's scripts call the function it replaces, and compares the words
with the oracle's restatement:
  - e2e_mult with the positional scale_bits of scripts/verify_e2e_mult.py
    (:173, :205, :220) and its toy rescale (:70-75);
  - poly_mult / neg_conv_modq / cyclic_conv_direct / relin for moduli that
    admit no transform (q = 113, composite q, q = 2), through the exact
    three-prime product;
  - noise_monitor_energy at q = 113 on 16 and 32 coefficients
    (scripts/verify_m4_h_multilayer.py:260);
  - encode / neg_ifft with more real slots than n/2 (a non-Hermitian V,
    scripts/verify_multi_slot_simd.py:55-84).
"""
import ctypes
import random

import numpy as np
import pytest

import ckks_oracle as O
import oracle_c as OC

pytestmark = pytest.mark.gpu

CHAIN4 = [17592186634241, 17592186372097, 17592186273793, 17592186175489]


@pytest.fixture(scope="module")
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from cipherpu import _lib
    _lib.load()
    return torch


def _toy_gpu(torch, vals, q_in, q_out, sb):
    from cipherpu import _lib
    x = torch.tensor(np.asarray(vals, dtype=np.uint64).view(np.int64)).cuda()
    out = torch.empty_like(x)
    _lib.call("ckks_mod_switch_scaled", ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(out.data_ptr()), x.numel(),
              q_in, q_out, sb, None)
    torch.cuda.synchronize()
    return [int(v) for v in out.cpu().numpy().view(np.uint64)]


@pytest.mark.parametrize("q_in,q_out", [(CHAIN4[0], CHAIN4[1]), (113, 113), (2, 3), ((1 << 45) - 1, 1000003),
                                        (CHAIN4[3], (1 << 45) - 1)])
def test_toy_rescale_kernel_every_shift(gpu, q_in, q_out):
    """ckks_mod_switch_scaled against the oracle's Python-integer formula on
    values below q_in, above it, and up to 2^64 - 1, for shifts 0..127."""
    rng = random.Random(q_in ^ q_out)
    vals = [0, 1, q_in - 1, q_in, q_in // 2, q_in // 2 + 1, (1 << 64) - 1, (1 << 63)]
    vals += [rng.randrange(q_in) for _ in range(200)] + [rng.randrange(1 << 64) for _ in range(200)]
    for sb in (0, 1, 7, 30, 44, 63, 64, 65, 90, 127):
        got = _toy_gpu(gpu, vals, q_in, q_out, sb)
        want = [O.rescale_toy(v, q_in, q_out, sb) for v in vals]
        assert got == want, (q_in, q_out, sb)


def test_toy_rescale_rejects_bad_arguments(gpu):
    from cipherpu import _lib
    torch = gpu
    x = torch.zeros(4, dtype=torch.int64, device="cuda")
    p = ctypes.c_void_p(x.data_ptr())
    for args in ((1, 5, 0), (5, 1, 0), (1 << 45, 5, 0), (5, 5, -1), (5, 5, 128)):
        with pytest.raises(_lib.CkksError):
            _lib.call("ckks_mod_switch_scaled", p, p, 4, *args, None)


def test_e2e_mult_reference_call_shape(gpu):
    """The flow of verify_e2e_mult.py main(): chain q0 -> q1, N = 16,
    SCALE_BITS = 30, random.Random(42) draws, the three tests' call shapes
    (random ciphertexts, the simplified encryptions, C * C), positional
    scale_bits; then scale_bits = None (exact modulus switch)."""
    from cipherpu import golden_api as G
    q_in, q_out, n, sb = CHAIN4[0], CHAIN4[1], 16, 30
    rng = random.Random(42)
    s = [rng.randrange(q_in) for _ in range(n)]
    a_ks = [rng.randrange(q_in) for _ in range(n)]
    e_ks = [rng.randrange(100) for _ in range(n)]
    b_ks = O.poly_add(O.poly_mult(a_ks, s, q_in), e_ks, q_in)
    assert G.poly_add(G.poly_mult(a_ks, s, q_in), e_ks, q_in) == b_ks
    cases = []
    for _ in range(10):
        cases.append([[rng.randrange(q_in) for _ in range(n)] for _ in range(4)])
    for _ in range(3):
        m1 = [rng.randrange(10) for _ in range(n)]
        m2 = [rng.randrange(10) for _ in range(n)]
        noise = [rng.randrange(100) for _ in range(n)]
        a_pk = [rng.randrange(q_in) for _ in range(n)]
        cases.append([O.poly_add(m1, noise, q_in), a_pk, O.poly_add(m2, noise, q_in), a_pk])
    c0 = [rng.randrange(q_in) for _ in range(n)]
    c1 = [rng.randrange(q_in) for _ in range(n)]
    cases.append([c0, c1, c0, c1])
    for c in cases:
        got = G.e2e_mult(*c, a_ks, b_ks, q_in, q_out, sb)
        want = O.e2e_mult(*c, a_ks, b_ks, q_in, q_out, sb)
        assert got[0] == want[0] and got[1] == want[1]
        assert all(0 <= v < q_out for v in got[0] + got[1])
        got = G.e2e_mult(*c, a_ks, b_ks, q_in, q_out)
        want = O.e2e_mult(*c, a_ks, b_ks, q_in, q_out, None)
        assert got[0] == want[0] and got[1] == want[1]


@pytest.mark.parametrize("q,n", [(113, 16), (113, 32), (2, 16), (3, 64), (1000003, 1024), ((1 << 45) - 1, 256),
                                 (998244353, 2048)])
def test_any_modulus_products(gpu, q, n):
    """poly_mult / neg_conv_modq / cyclic_conv_direct / relin / e2e_mult for
    moduli with no root of unity of the needed order: the exact product
    through three primes and CRT equals the schoolbook oracle."""
    from cipherpu import golden_api as G
    rng = random.Random(q * 7 + n)
    a = [rng.randrange(q) for _ in range(n)]
    b = [rng.randrange(q) for _ in range(n)]
    a[0], b[-1] = q - 1, q - 1
    if n <= 64:
        assert G.poly_mult(a, b, q) == O.poly_mult(a, b, q)
        assert G.cyclic_conv_direct(a, b, n, q) == O.cyclic_conv_direct(a, b, n, q)
        e0, e1, e2 = ([rng.randrange(q) for _ in range(n)] for _ in range(3))
        assert G.relin(e0, e1, e2, a, b, q) == O.relin(e0, e1, e2, a, b, q)
        cs = [[rng.randrange(q) for _ in range(n)] for _ in range(4)]
        for sb in (0, 3):
            got = G.e2e_mult(*cs, a, b, q, q, sb)
            assert got == O.e2e_mult(*cs, a, b, q, q, sb)
    else:
        assert G.poly_mult(a, b, q) == [int(v) for v in OC.polymul_schoolbook(a, b, q)]
    x = np.array([rng.randrange(-10 ** 6, 10 ** 6) for _ in range(n)], dtype=np.int64)
    y = np.array([rng.randrange(-1, 2) for _ in range(n)], dtype=np.int64)
    assert np.array_equal(G.neg_conv_modq(x, y, q), O.neg_conv_modq(x, y, q))


def test_any_modulus_product_at_full_length(gpu):
    """N = 16384 with q = 2^45 - 1 (odd, composite): the worst coefficient
    bound N (q-1)^2 < 2^104 the three-prime product must hold, against the
    C oracle's schoolbook; and extreme inputs (every coefficient q - 1)."""
    from cipherpu import golden_api as G
    q, n = (1 << 45) - 1, 16384
    rng = np.random.default_rng(4)
    a = rng.integers(0, q, n, dtype=np.uint64)
    b = rng.integers(0, q, n, dtype=np.uint64)
    assert G.poly_mult(list(a), list(b), q) == [int(v) for v in OC.polymul_schoolbook(a, b, q)]
    full = np.full(n, q - 1, dtype=np.uint64)
    # (q-1)^2 sum_j (+-1): coefficient k of the negacyclic square of the all-(q-1) row is (2k + 2 - n)(q-1)^2
    want = [((2 * k + 2 - n) * (q - 1) ** 2) % q for k in range(n)]
    assert G.poly_mult(list(full), list(full), q) == want


def test_noise_monitor_at_the_reference_toy_modulus(gpu):
    """noise_monitor_energy(f0p + f1p, TOY_Q=113) on 32 coefficients and its
    f0p / f1p halves on 16 (verify_m4_h_multilayer.py:260-262)."""
    from cipherpu import golden_api as G
    rng = random.Random(260)
    for n in (16, 32, 1, 7, 1000):
        c = [rng.randrange(113) for _ in range(n)]
        assert G.noise_monitor_energy(c, 113) == O.noise_monitor_energy(c, 113)
    c = [0, 56, 57, 112] * 8
    assert G.noise_monitor_energy(c, 113) == O.noise_monitor_energy(c, 113)


@pytest.mark.parametrize("n", [8, 16, 64])
def test_encode_with_more_slots_than_half(gpu, n):
    """encode with nslot > n/2 builds a non-Hermitian V; the drop-in's full
    complex inverse matches the oracle's direct evaluation."""
    from cipherpu import golden_api as G
    rng = np.random.default_rng(n)
    for nslot in (n // 2, n // 2 + 1, n - 1, n):
        v = rng.uniform(-1, 1, nslot)
        np.testing.assert_allclose(G.encode(v, 2.0, n), O.encode_natural(v, 2.0, n), atol=1e-9)
    V = rng.normal(size=n) + 1j * rng.normal(size=n)
    np.testing.assert_allclose(G.neg_ifft(V), O.neg_ifft(V), atol=1e-10)
    np.testing.assert_allclose(G.neg_fft(G.neg_ifft(V).real), O.neg_fft(O.neg_ifft(V).real), atol=1e-9)


def test_one_ckks_object_on_two_streams(gpu):
    """HMult and rotation of one CKKS object issued on two streams at once
    give the words of the same calls made one after the other: each stream
    has its own workspace."""
    torch = gpu
    from cipherpu.ckks import CKKS, CKKSParams
    ck = CKKS(CKKSParams(), seed=5)
    sk = ck.keygen()
    rlk = ck.relin_key(sk)
    gk = ck.galois_key(sk, 1)
    rng = np.random.default_rng(2)
    a = ck.encrypt(ck.encode(rng.uniform(-1, 1, (8, 4096))), sk)
    b = ck.encrypt(ck.encode(rng.uniform(-1, 1, (8, 4096))), sk)
    want_h = ck.hmult(a, b, rlk).data.clone()
    want_r = ck.rotate(a, gk).data.clone()
    want_h2 = ck.hmult(b, a, rlk).data.clone()
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.current_stream().synchronize()
    outs = []
    for _ in range(3):
        with torch.cuda.stream(s1):
            h = ck.hmult(a, b, rlk)
            h2 = ck.hmult(b, a, rlk)
        with torch.cuda.stream(s2):
            r = ck.rotate(a, gk)
        outs.append((h, h2, r))
    torch.cuda.synchronize()
    for h, h2, r in outs:
        assert torch.equal(h.data, want_h) and torch.equal(h2.data, want_h2) and torch.equal(r.data, want_r)


def test_ckks_from_the_reference_chain_file(gpu, fixtures, tmp_path):
    """CKKSParams.from_chain_json on the reference's chain schema gives a
    context whose HMult words equal those of the built-in parameters."""
    import json
    from cipherpu.ckks import CKKS, CKKSParams
    path = tmp_path / "modulus_chain_chain4.json"
    path.write_text(json.dumps(fixtures["chain4"]))
    a, b = CKKS(CKKSParams(), seed=3), CKKS(CKKSParams.from_chain_json(str(path)), seed=3)
    assert a.ring.roots == b.ring.roots
    outs = []
    for ck in (a, b):
        sk = ck.keygen()
        rlk = ck.relin_key(sk)
        z = np.random.default_rng(4).uniform(-1, 1, (2, 4096))
        ct = ck.encrypt(ck.encode(z), sk)
        outs.append(ck.hmult(ct, ct, rlk).data.cpu())
    assert (outs[0] == outs[1]).all()
