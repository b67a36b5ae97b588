"""GPU parity: the HIP library through its C ABI against the CPU oracle.

Integer work is compared bit-exactly; encode/decode (the only floating
point) within stated tolerances.  Sizes are chosen so the oracle finishes
in seconds; full-size batches are checked through size-independent
properties (decrypt error bounds) in test_gpu_fullsize.py, and the drop-in
functions at the reference's own call shapes in test_gpu_dropin.py."""
import random

import numpy as np
import pytest

import ckks_oracle as O
import oracle_c as OC

pytestmark = pytest.mark.gpu

CHAIN4 = [17592186634241, 17592186372097, 17592186273793, 17592186175489]
SPECIAL_P = 17592186765313
Q1, PSI1 = 35184371884033, 7388705119


@pytest.fixture(scope="module")
def gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from cipherpu import _lib
    _lib.load()
    return torch


def test_kats_through_drop_in_api(gpu, fixtures):
    from cipherpu import golden_api as G
    k = fixtures["kats"]["ntt_n8"]
    assert G.bit_rev_perm(k["input"], 8) == k["bitrev_input"]
    assert G.ntt(k["input"], 18, 8, 113) == k["expected"]
    assert G.intt(k["expected"], pow(18, -1, 113), 8, 113) == k["input"]
    k8 = fixtures["kats"]["pm_n8"]
    assert G.poly_mult_via_ntt(k8["a"], k8["b"], 8, 113, 18, pow(18, -1, 113)) == k8["expected"]
    k16 = fixtures["kats"]["pm_n16"]
    assert G.poly_mult_via_ntt(k16["a"], k16["b"], 16, 113, 40, pow(40, -1, 113)) == k16["expected"]


@pytest.mark.parametrize("N", [16, 256, 2048, 4096, 8192, 16384])
def test_cyclic_ntt_matches_reference_ntt(gpu, N):
    from cipherpu import golden_api as G
    q = CHAIN4[3]
    w = O.root_of_unity(q, N)
    rng = np.random.default_rng(N)
    a = rng.integers(0, q, N, dtype=np.uint64)
    X = G.ntt(list(a), w, N, q)
    assert X == list(OC.ntt_cyclic(a, N, q, w))
    assert G.intt(X, pow(w, -1, q), N, q) == list(a)


@pytest.mark.parametrize("N,threads", [(8, 0), (64, 0), (1024, 0), (4096, 128), (4096, 256), (4096, 512),
                                       (8192, 256), (8192, 512), (8192, 1024), (16384, 512), (16384, 1024)])
def test_negacyclic_ntt_batched(gpu, N, threads):
    torch = gpu
    from cipherpu.ring import RingContext, bitrev_, to_numpy
    primes = CHAIN4 + [SPECIAL_P]
    ring = RingContext(N, primes)
    ring.set_threads(threads, 0)
    rng = np.random.default_rng(N + threads)
    x = rng.integers(0, 1 << 44, (3, 5, N), dtype=np.uint64) % np.array(primes, dtype=np.uint64)[None, :, None]
    X = ring.ntt(x)
    Xn = to_numpy(bitrev_(X.clone()))
    for p in range(3):
        for i in range(5):
            want = OC.ntt_negacyclic(x[p, i], N, primes[i], ring.roots[i])
            assert np.array_equal(Xn[p, i], want), (p, i)
    back = to_numpy(ring.intt(X))
    assert np.array_equal(back, x)
    # natural-order flag gives the same as the explicit permutation
    assert np.array_equal(to_numpy(ring.ntt(x, natural=True)), Xn)


def test_ntt_edge_values(gpu):
    from cipherpu.ring import RingContext, to_numpy
    N = 8192
    ring = RingContext(N, CHAIN4[:1])
    q = CHAIN4[0]
    for fill in (0, 1, q - 1):
        x = np.full((1, 1, N), fill, dtype=np.uint64)
        X = ring.ntt(x)
        assert np.array_equal(to_numpy(ring.intt(X)), x)
    x = np.zeros((1, 1, N), dtype=np.uint64)
    x[0, 0, 0] = 1   # delta -> all ones
    assert np.all(to_numpy(ring.ntt(x)) == 1)
    empty = ring.tensor(np.zeros((0, 1, N), dtype=np.uint64))
    ring.ntt_(empty)


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("threads", [0, 128, 512])
def test_single_prime_product_bit_exact(gpu, threads, fp64):
    """Single-prime product: N=4096, q = largest 45-bit prime = 1 mod 8192, psi =
    smallest primitive 8192-th root; forward NTT -> point-wise -> INTT, on
    64-bit integer words and on the FP64 pipeline."""
    from cipherpu.ring import RingContext, to_numpy
    ring = RingContext(4096, [Q1], roots=[PSI1])
    ring.set_threads(threads, 0)
    ring.set_arith(fp64)
    assert ring.roots == [PSI1]
    rng = np.random.default_rng(123)
    a = rng.integers(0, Q1, (6, 1, 4096), dtype=np.uint64)
    b = rng.integers(0, Q1, (6, 1, 4096), dtype=np.uint64)
    a[4, 0] = 0
    a[4, 0, 0] = 1   # delta: product = b
    b[5] = Q1 - 1
    c = to_numpy(ring.poly_mul(ring.tensor(a), ring.tensor(b)))
    for i in range(6):
        assert np.array_equal(c[i, 0], OC.polymul_ntt(a[i, 0], b[i, 0], Q1, PSI1)), i
    assert np.array_equal(c[4, 0], b[4, 0])
    assert np.array_equal(c[0, 0], OC.polymul_schoolbook(a[0, 0], b[0, 0], Q1))


def test_poly_mult_drop_in_small(gpu):
    from cipherpu import golden_api as G
    rng = random.Random(5)
    q = CHAIN4[1]
    for n in (16, 32):
        a = [rng.randrange(q) for _ in range(n)]
        b = [rng.randrange(q) for _ in range(n)]
        assert G.poly_mult(a, b, q) == O.poly_mult(a, b, q)
        assert G.poly_add(a, b, q) == O.poly_add(a, b, q)
        assert G.poly_sub(a, b, q) == O.poly_sub(a, b, q)


def test_relin_drop_in(gpu):
    """verify_keyswitch.relin semantics at N=16 against the schoolbook oracle."""
    from cipherpu import golden_api as G
    rng = random.Random(42)
    q, n = CHAIN4[0], 16
    s = [rng.randrange(q) for _ in range(n)]
    for _ in range(5):
        e0, e1, e2, a = ([rng.randrange(q) for _ in range(n)] for _ in range(4))
        b = O.poly_add(O.poly_mult(a, s, q), [rng.randrange(100) for _ in range(n)], q)
        assert G.relin(e0, e1, e2, a, b, q) == O.relin(e0, e1, e2, a, b, q)


def test_rescale_coeff_drop_in(gpu):
    from cipherpu import golden_api as G
    rng = random.Random(1)
    for i in range(3):
        q_in, q_out = CHAIN4[i], CHAIN4[i + 1]
        pts = [0, 1, 2, q_in - 3, q_in - 2, q_in - 1, q_out - 1, q_out, q_out + 1]
        pts += [rng.randrange(q_in) for _ in range(4000)]
        pts = [p for p in pts if 0 <= p < q_in]
        assert G.rescale_poly(pts, q_in, q_out) == O.rescale_poly(pts, q_in, q_out)
        assert G.rescale_coeff(q_in - 1, q_in, q_out) == O.rescale_coeff(q_in - 1, q_in, q_out)
    # the other direction and a non-chain pair
    assert G.rescale_poly([5, 10 ** 12], CHAIN4[1], CHAIN4[0]) == O.rescale_poly([5, 10 ** 12], CHAIN4[1], CHAIN4[0])


@pytest.mark.parametrize("N", [4096, 8192])
def test_rns_rescale(gpu, N):
    from cipherpu.ring import RingContext, to_numpy
    ring = RingContext(N, CHAIN4)
    rng = np.random.default_rng(N)
    for l in (2, 3, 4):
        x = np.stack([rng.integers(0, q, (3, N), dtype=np.uint64) for q in CHAIN4[:l]], axis=1)
        x[0, :, :4] = 0
        x[0, :, 4:8] = np.array(CHAIN4[:l], dtype=np.uint64)[:, None] - 1
        got = to_numpy(ring.rescale(ring.tensor(x)))
        for p in range(3):
            assert np.array_equal(got[p], OC.rescale(x[p], CHAIN4[:l]))
        small = [list(map(int, x[1, i, :32])) for i in range(l)]
        assert got[1, :, :32].tolist() == O.rns_rescale(small, CHAIN4[:l])


@pytest.mark.parametrize("N", [64, 128, 2048, 8192, 16384])
def test_automorphism(gpu, N):
    """The Galois map against the C oracle: the gather kernel (N < 128) and
    the LDS-staged row kernel (64 to 1024 threads, up to 128 KB of LDS)."""
    from cipherpu.ring import RingContext, to_numpy
    from cipherpu import golden_api as G
    ring = RingContext(N, CHAIN4)
    rng = np.random.default_rng(3)
    x = np.stack([rng.integers(0, q, (2, N), dtype=np.uint64) for q in CHAIN4], axis=1)
    x[0, :, :3] = 0                      # zero coefficients stay zero when negated
    for g in (5, 25, 5 ** 7 % (2 * N), 2 * N - 1, 3):
        got = to_numpy(ring.automorphism(ring.tensor(x), g))
        for p in range(2):
            for i in range(4):
                assert np.array_equal(got[p, i], OC.automorphism(x[p, i], CHAIN4[i], g)), (g, p, i)
    n, q = 16, 97
    ct = ([rng.integers(0, q) for _ in range(n)], [rng.integers(0, q) for _ in range(n)])
    a0, a1 = G.automorphism_exact(ct, 5, n, q)
    assert list(a0) == O.automorphism(ct[0], 5, n, q) and list(a1) == O.automorphism(ct[1], 5, n, q)


def test_noise_stats(gpu):
    from cipherpu.ring import RingContext
    N = 4096
    ring = RingContext(N, CHAIN4[:2])
    rng = np.random.default_rng(9)
    x = np.stack([rng.integers(0, q, (2, N), dtype=np.uint64) for q in CHAIN4[:2]], axis=1)
    st = ring.noise_stats(ring.tensor(x))
    for p in range(2):
        for i in range(2):
            q = CHAIN4[i]
            c = [int(v) for v in x[p, i]]
            a = [min(v, q - v) for v in c]
            row = st[p * 2 + i]
            assert row[0] == sum(a) and row[1] == sum(v * v for v in a) and row[2] == max(a)
            assert row[3] == sum(v if v <= q - v else v - q for v in c)
            E, r = O.noise_monitor_energy(c, q)
            assert row[1] / N == E


def test_noise_monitor_energy_drop_in(gpu):
    """golden_api.noise_monitor_energy (the statistics kernel's exact sums)
    against the oracle's restatement of verify_m4_h_multilayer.py:162-174 on
    ragged lengths, edge values and a prime of each chain."""
    from cipherpu import golden_api as G
    rng = np.random.default_rng(12)
    for q, n in ((CHAIN4[3], 1), (CHAIN4[3], 5), (CHAIN4[0], 1000), (CHAIN4[1], 4096), (CHAIN4[2], 8191)):
        c = [int(v) for v in rng.integers(0, q, n, dtype=np.uint64)]
        c[0] = q - 1
        if n > 2:
            c[1], c[2] = 0, q // 2
        assert G.noise_monitor_energy(c, q) == O.noise_monitor_energy(c, q), (q, n)
    with pytest.raises(ValueError):
        G.noise_monitor_energy([CHAIN4[0]], CHAIN4[0])


# ---------------------------------------------------------------- CKKS ops

@pytest.fixture(scope="module")
def ck(gpu):
    from cipherpu.ckks import CKKS, CKKSParams
    c = CKKS(CKKSParams(), seed=11)
    sk = c.keygen()
    return c, sk


def test_keygen_bit_exact(ck):
    """b_j = a_j s + e_j + g_j s^2 per limb, against the oracle's products."""
    from cipherpu.ring import to_numpy
    c, sk = ck
    rng = np.random.default_rng(4)
    L, K, N = 4, 5, c.N
    a = np.stack([rng.integers(0, q, (L, N), dtype=np.uint64) for q in c.primes], axis=1)
    e = rng.integers(-6, 7, (L, N)).astype(np.int64)
    rlk = c.relin_key(sk, a=c.ring.tensor(a), e=e)
    got = to_numpy(rlk.coeff)
    s = sk.s
    for j in (0, 3):
        for i in range(K):
            q = c.primes[i]
            srns = np.where(s < 0, q + s, s).astype(np.uint64)
            want = OC.polymul_ntt(a[j, i], srns, q, c.ring.roots[i])
            want = (want + np.where(e[j] < 0, q + e[j], e[j]).astype(np.uint64)) % np.uint64(q)
            if i == j:
                s2 = OC.polymul_ntt(srns, srns, q, c.ring.roots[i])
                want = np.array([(int(w) + (SPECIAL_P % q) * int(v)) % q for w, v in zip(want, s2)], dtype=np.uint64)
            assert np.array_equal(got[j, 0, i], want), (j, i)
            assert np.array_equal(got[j, 1, i], a[j, i])


def _oracle_args(c):
    return c.primes[:4], c.primes[4], c.ring.roots[:4], c.ring.roots[4]


def test_relin_keyswitch_bit_exact(ck):
    from cipherpu.ring import to_numpy
    c, sk = ck
    rlk = c.relin_key(sk)
    rng = np.random.default_rng(6)
    N = c.N
    for l in (4, 2):
        d = [np.stack([rng.integers(0, q, (2, N), dtype=np.uint64) for q in c.primes[:l]], axis=1) for _ in range(3)]
        out = to_numpy(c.relin(*(c.ring.tensor(v) for v in d), rlk))
        key = to_numpy(rlk.coeff)
        pr, p, ps, pp = _oracle_args(c)
        for b in range(2):
            k0, k1 = OC.keyswitch(d[2][b], key, pr[:l], p, ps[:l], pp)
            want0 = (k0 + d[0][b]) % np.array(pr[:l], dtype=np.uint64)[:, None]
            want1 = (k1 + d[1][b]) % np.array(pr[:l], dtype=np.uint64)[:, None]
            assert np.array_equal(out[b, 0], want0) and np.array_equal(out[b, 1], want1), (l, b)


@pytest.mark.parametrize("l", [4, 3, 2])
def test_hmult_bit_exact(ck, l):
    from cipherpu.ring import to_numpy
    c, sk = ck
    rlk = c.relin_key(sk)
    rng = np.random.default_rng(l)
    z = rng.uniform(-1, 1, (3, c.N // 2))
    cta = c.encrypt(c.encode(z, level=l), sk)
    ctb = c.encrypt(c.encode(z[::-1].copy(), level=l), sk)
    out = c.hmult(cta, ctb, rlk)
    got = to_numpy(out.data)
    key = to_numpy(rlk.coeff)
    pr, p, ps, pp = _oracle_args(c)
    for b in range(3):
        want = OC.hmult(to_numpy(cta.data[b]), to_numpy(ctb.data[b]), key, pr[:l], p, ps[:l], pp)
        assert np.array_equal(got[b], want), b


def test_hmult_uniform_inputs_and_small_n(gpu):
    """Uniformly random (non-RLWE) inputs at N=4096: pure datapath parity,
    as the reference's 'uniform' regime (exp_e33_real_ciphertext.py:160-170)."""
    from cipherpu.ckks import CKKS, CKKSParams, Ciphertext
    from cipherpu.ring import to_numpy
    c = CKKS(CKKSParams(N=4096), seed=2)
    sk = c.keygen()
    rlk = c.relin_key(sk)
    rng = np.random.default_rng(8)
    mk = lambda: np.stack([rng.integers(0, q, (2, 2, 4096), dtype=np.uint64) for q in c.primes[:4]], axis=2)
    a, b = mk(), mk()
    out = to_numpy(c.hmult(Ciphertext(c.ring.tensor(a), 1.0), Ciphertext(c.ring.tensor(b), 1.0), rlk).data)
    key = to_numpy(rlk.coeff)
    pr, p, ps, pp = _oracle_args(c)
    for i in range(2):
        assert np.array_equal(out[i], OC.hmult(a[i], b[i], key, pr, p, ps, pp))


@pytest.mark.parametrize("N,batch,threads", [(4096, 523, 512), (8192, 261, 1024)])
def test_looping_kernels_over_many_items(gpu, N, batch, threads):
    """Batches that give the looping front and key-switch kernels several
    items per workgroup, with item counts that are not multiples of 8 (the
    XCD deal falls back to the plain order): HMult under both key-switch
    placements and both fronts, and the rotation's key switch, give the same
    words, and the first, a middle and the last ciphertext equal the oracle."""
    from cipherpu.ckks import CKKS, CKKSParams, Ciphertext
    from cipherpu.ring import to_numpy
    c = CKKS(CKKSParams(N=N), seed=29)
    c.ring.set_threads(0, threads, threads)   # paired shapes: the looping kernels
    sk = c.keygen()
    rlk = c.relin_key(sk)
    gk = c.galois_key(sk, 3)
    rng = np.random.default_rng(batch)
    mk = lambda: np.stack([rng.integers(0, q, (batch, 2, N), dtype=np.uint64) for q in c.primes[:4]], axis=2)
    a, b = mk(), mk()
    ca, cb = Ciphertext(c.ring.tensor(a), 1.0), Ciphertext(c.ring.tensor(b), 1.0)
    ref = to_numpy(c.hmult(ca, cb, rlk).data)
    for front, order in ((1, 0), (0, 1), (0, 0)):
        c.ring.set_front(bool(front))
        c.ring.set_ks_order(bool(order))
        assert np.array_equal(to_numpy(c.hmult(ca, cb, rlk).data), ref), (front, order)
    rot = to_numpy(c.rotate(ca, gk).data)
    key, gkey = to_numpy(rlk.coeff), to_numpy(gk.coeff)
    pr, p, ps, pp = _oracle_args(c)
    for i in (0, batch // 2, batch - 1):
        assert np.array_equal(ref[i], OC.hmult(a[i], b[i], key, pr, p, ps, pp)), i
        assert np.array_equal(rot[i], OC.rotate(a[i], gkey, pr, p, ps, pp, gk.galois)), i


@pytest.mark.parametrize("N,threads,l", [(8192, 1024, 4), (8192, 512, 3), (4096, 512, 2), (4096, 256, 4)])
def test_squaring_gives_the_words_of_the_product(gpu, N, threads, l):
    """HMult of a ciphertext batch with itself (one pointer: the squaring
    front, or one forward transform on the unfused path) gives the words of
    HMult with a copy of it, on both fronts and in both arithmetics, and
    the first and last ciphertexts equal the oracle's product."""
    from cipherpu.ckks import CKKS, CKKSParams, Ciphertext
    from cipherpu.ring import to_numpy
    c = CKKS(CKKSParams(N=N), seed=37)
    c.ring.set_threads(threads, threads, threads)
    sk = c.keygen()
    rng = np.random.default_rng(N + l)
    batch = 37
    a = np.stack([rng.integers(0, q, (batch, 2, N), dtype=np.uint64) for q in c.primes[:l]], axis=2)
    ca = Ciphertext(c.ring.tensor(a), 1.0)
    cb = Ciphertext(ca.data.clone(), 1.0)
    for fp64 in (True, False):
        c.ring.set_arith(fp64)
        rlk = c.relin_key(sk)
        for front in (True, False):
            c.ring.set_front(front)
            want = to_numpy(c.hmult(ca, cb, rlk).data)
            assert np.array_equal(to_numpy(c.hmult(ca, ca, rlk).data), want), (fp64, front)
    key = to_numpy(rlk.coeff)
    pr, p, ps, pp = _oracle_args(c)
    for i in (0, batch - 1):
        assert np.array_equal(want[i], OC.hmult(a[i], a[i], key, pr[:l], p, ps[:l], pp)), i


@pytest.mark.parametrize("l", [4, 2, 1])
def test_rotate_add_gives_the_words_of_rotate_then_add(gpu, l):
    """ckks_rotate_add (the input added in the rotation's ModDown pass)
    against ckks_rotate followed by the point-wise sum, in both arithmetics,
    and the decoded slots of ct + rot(ct)."""
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu.ring import to_numpy
    c = CKKS(CKKSParams(), seed=41)
    sk = c.keygen()
    z = np.random.default_rng(l).uniform(-1, 1, (5, c.N // 2))
    ct = c.encrypt(c.encode(z, level=l), sk)
    for fp64 in (True, False):
        c.ring.set_arith(fp64)
        for step in (1, -128, 512):
            gk = c.galois_key(sk, step)
            want = to_numpy(c.add(ct, c.rotate(ct, gk)).data)
            got = c.rotate_add(ct, gk)
            assert np.array_equal(to_numpy(got.data), want), (fp64, step)
            dec = to_numpy(c.decode(c.decrypt(got, sk), got.scale)).real
            assert np.max(np.abs(dec - (z + np.roll(z, -step, axis=1)))) < 1e-5, (fp64, step)


@pytest.mark.parametrize("N", [4096, 8192, 16384])
def test_fused_encryption_gives_the_words_of_the_separate_kernels(gpu, N):
    """ckks_encrypt_sk (mask, product, error and sums in one kernel) and
    ckks_decrypt_sk against
    the separate device samplers, ckks_poly_mul and the point-wise sums under
    the same seeds, call after call and at several levels, in words; and the
    fused ciphertexts decrypt to the message."""
    import torch
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu.ring import to_numpy
    fused, split = CKKS(CKKSParams(N=N), seed=51), CKKS(CKKSParams(N=N), seed=51)
    split.fused_encrypt = split.fused_decrypt = False
    skf, sks = fused.keygen(), split.keygen()
    rng = np.random.default_rng(N)
    for l in (4, 3, 1):
        z = rng.uniform(-1, 1, (3, N // 2))
        for _ in range(2):
            a = fused.encrypt(fused.encode(z, level=l), skf)
            b = split.encrypt(split.encode(z, level=l), sks)
            assert torch.equal(a.data, b.data), (N, l)
            assert torch.equal(fused.decrypt(a, skf), split.decrypt(b, sks)), (N, l)   # ckks_decrypt_sk
        dec = to_numpy(fused.decode(fused.decrypt(a, skf), a.scale)).real
        assert np.max(np.abs(dec - z)) < 1e-7, (N, l)


def test_rotate_bit_exact_and_slots_move(ck):
    from cipherpu.ring import to_numpy
    c, sk = ck
    rng = np.random.default_rng(10)
    z = rng.uniform(-1, 1, (2, c.N // 2))
    ct = c.encrypt(c.encode(z), sk)
    pr, p, ps, pp = _oracle_args(c)
    for step in (1, 16):
        gk = c.galois_key(sk, step)
        rot = c.rotate(ct, gk)
        got = to_numpy(rot.data)
        key = to_numpy(gk.coeff)
        for b in range(2):
            want = OC.rotate(to_numpy(ct.data[b]), key, pr, p, ps, pp, gk.galois)
            assert np.array_equal(got[b], want), (step, b)
        dec = to_numpy(c.decode(c.decrypt(rot, sk), rot.scale))
        # key-switch noise: coefficient std ~ 2 sigma q sqrt(N / 3) / p ~ 330, but the digits are
        # the canonical lifts in [0, q), whose mean q / 2 gives a term (q / 2) sum_k zeta^k that
        # peaks near zeta = +-1: the worst slot reaches ~2e-6 at scale 2^40 (the C oracle alone
        # shows 1.7e-6 on one draw)
        assert np.max(np.abs(dec.real - np.roll(z, -step, axis=1))) < 1e-5


@pytest.mark.parametrize("N", [4096, 8192])
def test_rotate_many_equals_one_rotation_per_ciphertext(gpu, N):
    """ckks_rotate_many with a different step (and key) per ciphertext gives,
    word for word, what ckks_rotate gives for each ciphertext alone, at two
    levels, with repeated and distinct steps; one of them against the
    oracle."""
    from cipherpu.ckks import CKKS, CKKSParams, Ciphertext
    from cipherpu.ring import to_numpy
    c = CKKS(CKKSParams(N=N), seed=71)
    sk = c.keygen()
    steps = [1, 8, 1, 256, 3, N // 4 - 1]
    gks = {s: c.galois_key(sk, s) for s in set(steps)}
    rng = np.random.default_rng(N + 71)
    pr, p, ps, pp = _oracle_args(c)
    for l in (4, 2):
        ct = c.encrypt(c.encode(rng.uniform(-1, 1, (len(steps), N // 2)), level=l), sk)
        got = to_numpy(c.rotate_many(ct, [gks[s] for s in steps]).data)
        for b, s in enumerate(steps):
            one = to_numpy(c.rotate(Ciphertext(ct.data[b:b + 1].contiguous(), ct.scale), gks[s]).data)[0]
            assert np.array_equal(got[b], one), (l, b, s)
        want = OC.rotate(to_numpy(ct.data[3]), to_numpy(gks[256].coeff), pr[:l], p, ps[:l], pp, gks[256].galois)
        assert np.array_equal(got[3], want), l


def test_encode_decode_vs_fp64_oracle(ck):
    from cipherpu.ring import to_numpy
    c, sk = ck
    rng = np.random.default_rng(12)
    z = rng.uniform(-1, 1, (2, c.N // 2)) + 1j * rng.uniform(-1, 1, (2, c.N // 2))
    pt = to_numpy(c.encode(z))
    for b in range(2):
        m = np.rint(O.encode_fft(z[b] * 2.0 ** 40, c.N)).astype(np.int64)
        want = np.where(m < 0, np.uint64(c.primes[0]) - (-m).astype(np.uint64), m.astype(np.uint64))
        # FP64 rounding ties may differ by one unit in rare coefficients
        diff = (pt[b, 0].astype(np.int64) - want.astype(np.int64)) % c.primes[0]
        diff = np.minimum(diff, c.primes[0] - diff)
        assert diff.max() <= 1 and (diff != 0).mean() < 1e-3
    dec = to_numpy(c.decode(c.ring.tensor(pt), 2.0 ** 40))
    assert np.max(np.abs(dec - z)) < 1e-9
    # decode of a multi-limb plaintext equals the oracle's exact-CRT decode
    want = O.decode_rns([list(map(int, pt[0, i])) for i in range(4)], 2.0 ** 40, c.N, c.primes[:4])
    assert np.max(np.abs(dec[0] - want)) < 1e-9


def test_natural_order_embedding_drop_in(gpu):
    from cipherpu import golden_api as G
    v = np.array([1.0, 2.0, 3.0, 4.0])
    m = G.encode(v, 2.0, 8)
    np.testing.assert_allclose(m, O.encode_natural(v, 2.0, 8), atol=1e-12)
    np.testing.assert_allclose(G.decode(m, 2.0, 8, 4), v, atol=1e-12)
    rng = np.random.default_rng(0)
    mm = rng.normal(size=64)
    np.testing.assert_allclose(G.neg_fft(mm), O.neg_fft(mm), atol=1e-10)


def test_error_codes(gpu):
    from cipherpu import _lib
    from cipherpu.ring import RingContext
    with pytest.raises(_lib.CkksError) as e:
        RingContext(8192, [Q1])     # 2N = 16384 does not divide q-1
    assert e.value.code == -2
    ring = RingContext(4096, CHAIN4[:2])
    x = ring.tensor(np.zeros((1, 3, 4096), dtype=np.uint64))
    with pytest.raises(ValueError):
        ring.ntt_(x)                # 3 limbs on a 2-prime context


def test_pair_buffers_need_16_byte_alignment(ck):
    """Prepared keys and plaintext pairs are read as 16-byte pairs: a key or
    plaintext pointer 8 bytes off is rejected before anything is launched,
    while 8-byte alignment stays enough for ciphertext data."""
    import ctypes
    import torch
    from cipherpu._lib import CkksError, call
    c, sk = ck
    rlk = c.relin_key(sk)
    p = lambda t, off=0: ctypes.c_void_p(t.data_ptr() + off)
    B, l = 1, c.L
    d = torch.zeros(B, 3, l, c.N, dtype=torch.uint64, device=c.device)
    out = torch.zeros(B, 2, l - 1, c.N, dtype=torch.uint64, device=c.device)
    words = (call("ckks_workspace_bytes", c.ring.handle, B, l) + 7) // 8
    ws = torch.zeros(words + 2, dtype=torch.uint64, device=c.device)
    key = torch.zeros(rlk.prepared.numel() + 2, dtype=torch.uint64, device=c.device)
    with pytest.raises(CkksError) as e:
        call("ckks_hmult", c.ring.handle, p(d), p(d), p(key, 8), c.L + 1, p(out), B, l, c.sp, p(ws),
             c.ring.stream())
    assert e.value.code == -3
    E = torch.zeros(1, B, 2, l, c.N, dtype=torch.uint64, device=c.device)
    D = torch.zeros(1, 1, l, c.N, 2 + 1, dtype=torch.uint64, device=c.device)
    o = torch.zeros(1, B, 2, l, c.N, dtype=torch.uint64, device=c.device)
    with pytest.raises(CkksError) as e:
        call("ckks_lincomb", c.ring.handle, p(E), p(D, 8), p(o), 1, 1, B, l, c.ring.stream())
    assert e.value.code == -3
    call("ckks_lincomb", c.ring.handle, p(E), p(D), p(o), 1, 1, B, l, c.ring.stream())   # aligned: accepted
    torch.cuda.synchronize()


def test_committed_golden_vectors(gpu):
    """The HIP path against tests/golden/golden_vectors.npz (made by
    tests/golden/make_golden.py from the oracle)."""
    import hashlib
    import os
    import sys
    from cipherpu.ckks import CKKS, CKKSParams, Ciphertext
    from cipherpu.ring import RingContext, to_numpy
    here = os.path.dirname(os.path.abspath(__file__))
    g = np.load(os.path.join(here, "golden", "golden_vectors.npz"))
    ring1 = RingContext(4096, [Q1], roots=[PSI1])
    c = to_numpy(ring1.poly_mul(ring1.tensor(g["cfg1_a"][None, None]), ring1.tensor(g["cfg1_b"][None, None])))
    assert np.array_equal(c[0, 0], g["cfg1_c"])
    ring = RingContext(8192, CHAIN4)
    assert np.array_equal(to_numpy(ring.ntt(g["ntt8192_q0_in"][None, None], natural=True))[0, 0],
                          g["ntt8192_q0_out"])
    r64 = RingContext(64, CHAIN4)
    assert np.array_equal(to_numpy(r64.rescale(r64.tensor(g["rescale_in"][None]))[0]), g["rescale_out"])
    sys.path.insert(0, os.path.join(here, "golden"))
    import make_golden
    s, key, cta, ctb, _ = make_golden.hmult_case(8192, 4, int(g["hmult8192_seed"][0]))
    ck = CKKS(CKKSParams())
    rlk = ck.load_switch_key(key)
    out = ck.hmult(Ciphertext(ck.ring.tensor(cta[None]), 1.0), Ciphertext(ck.ring.tensor(ctb[None]), 1.0), rlk)
    d = hashlib.sha256(np.ascontiguousarray(to_numpy(out.data)[0], dtype="<u8").tobytes()).digest()
    assert d == bytes(g["hmult8192_out_sha256"])


def test_serialisation_roundtrip(ck, tmp_path):
    from cipherpu import ckks as CK
    from cipherpu.ring import to_numpy
    c, sk = ck
    rlk = c.relin_key(sk)
    z = np.random.default_rng(1).uniform(-1, 1, (1, c.N // 2))
    ct = c.encrypt(c.encode(z), sk)
    p = str(tmp_path / "state.npz")
    CK.save(p, ct=ct, rlk=rlk, sk=sk)
    back = CK.load(p, c)
    assert np.array_equal(to_numpy(back["ct"].data), to_numpy(ct.data)) and back["ct"].scale == ct.scale
    assert np.array_equal(to_numpy(back["rlk"].prepared), to_numpy(rlk.prepared))
    out1 = to_numpy(c.hmult(ct, ct, rlk).data)
    out2 = to_numpy(c.hmult(back["ct"], back["ct"], back["rlk"]).data)
    assert np.array_equal(out1, out2)


def test_special_form_product_matches_barrett(ck):
    """Pointwise op 4 (special-form reduction) equals op 0 (Barrett) and
    Python's a*b % q on every chain prime; a non-special prime is rejected."""
    import ctypes
    import torch
    from cipherpu._lib import CkksError
    from cipherpu.ring import RingContext, to_numpy
    c, _ = ck
    rng = np.random.default_rng(31)
    a = np.stack([rng.integers(0, q, (3, c.N), dtype=np.uint64) for q in c.primes], axis=1)
    b = np.stack([rng.integers(0, q, (3, c.N), dtype=np.uint64) for q in c.primes], axis=1)
    for i, q in enumerate(c.primes):          # extreme residues
        a[0, i, 0], b[0, i, 0] = q - 1, q - 1
        a[0, i, 1], b[0, i, 1] = 0, q - 1
    A, B = c.ring.tensor(a), c.ring.tensor(b)
    got4 = to_numpy(c.ring._pw(A, B, 4, 0, None))
    got0 = to_numpy(c.ring.mul(A, B))
    assert np.array_equal(got4, got0)
    for i, q in enumerate(c.primes):
        want = (a[1, i].astype(object) * b[1, i].astype(object)) % q
        assert np.array_equal(got4[1, i].astype(object), want)
    r1 = RingContext(4096, [Q1], roots=[PSI1])
    x = r1.tensor(np.ones((1, 1, 4096), dtype=np.uint64))
    with pytest.raises(CkksError):
        r1._pw(x, x, 4, 0, None)


def test_rotate_hoisted_bit_exact_and_slots_move(ck):
    """Hoisted rotations against the oracle's hoisted restatement, bit for
    bit, and decoded slots rolled by each step."""
    from cipherpu.ring import to_numpy
    c, sk = ck
    rng = np.random.default_rng(14)
    z = rng.uniform(-1, 1, (2, c.N // 2))
    pr, p, ps, pp = _oracle_args(c)
    for l in (4, 3, 2):
        ct = c.encrypt(c.encode(z, level=l), sk)
        steps = [1, 5, 64, -3]
        gks = [c.galois_key(sk, st) for st in steps]
        outs = c.rotate_hoisted(ct, gks)
        for st, gk, out in zip(steps, gks, outs):
            got = to_numpy(out.data)
            for b in range(2):
                want = OC.rotate_hoisted(to_numpy(ct.data[b]), to_numpy(gk.coeff), pr[:l], p, ps[:l], pp,
                                         gk.galois)
                assert np.array_equal(got[b], want), (l, st, b)
            dec = to_numpy(c.decode(c.decrypt(out, sk), out.scale))
            assert np.max(np.abs(dec.real - np.roll(z, -st, axis=1))) < 1e-5, (l, st)   # as above


@pytest.mark.parametrize("N", [4096, 8192, 16384])
def test_every_launch_shape_gives_the_same_words(gpu, N):
    """HMult (fused and separate front), rotation and hoisted rotation under
    every built thread count of the transform, tensor/front and key-switch
    kernels, and both key-switch placements, give the same words as the
    default shape, whose HMult equals the C oracle: the launch shape (single
    or paired transforms, 8 to 32 coefficients per thread) never changes a
    result."""
    import itertools
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu.ring import to_numpy
    from cipherpu.tune import BUILT, BUILT_FUSED
    c = CKKS(CKKSParams(N=N), seed=21)
    sk = c.keygen()
    rlk = c.relin_key(sk)
    gks = [c.galois_key(sk, s) for s in (1, 3)]
    rng = np.random.default_rng(N)
    z = rng.uniform(-1, 1, (2, N // 2))
    a, b = (c.encrypt(c.encode(v), sk) for v in (z, z[::-1].copy()))
    run = lambda: (to_numpy(c.hmult(a, b, rlk).data), to_numpy(c.rotate(a, gks[0]).data),
                   [to_numpy(r.data) for r in c.rotate_hoisted(a, gks)] +
                   [to_numpy(r.data) for r in c.rotate_hoisted(a, gks, eval_form=True)])
    ref = run()
    key = to_numpy(rlk.coeff)
    pr, p, ps, pp = _oracle_args(c)
    assert np.array_equal(ref[0][1], OC.hmult(to_numpy(a.data[1]), to_numpy(b.data[1]), key, pr, p, ps, pp))
    logN = N.bit_length() - 1
    for front, ntt, tensor, ks in itertools.product((0, 1), BUILT[logN], BUILT_FUSED[logN], BUILT_FUSED[logN]):
        c.ring.set_front(bool(front))
        c.ring.set_threads(ntt, ks, tensor)
        c.ring.set_ks_order(bool((ntt + ks) & 256))   # both key-switch placements across the grid
        got = run()
        shape = (front, ntt, tensor, ks)
        assert np.array_equal(got[0], ref[0]), ("hmult",) + shape
        assert np.array_equal(got[1], ref[1]), ("rotate",) + shape
        assert all(np.array_equal(g, r) for g, r in zip(got[2], ref[2])), ("hoisted",) + shape


@pytest.mark.parametrize("N", [4096, 8192, 16384])
def test_integer_arithmetic_gives_the_same_words(gpu, N):
    """The 64-bit integer kernels (Shoup and Barrett products on
    __umul64hi, CKKS_ARITH_INT64) against the FP64 residue pipeline, which
    the oracle pins: HMult (fused and separate front) under every built
    tensor/front and key-switch shape, relinearisation, rotation, mixed-step
    rotation and hoisted rotations (both output forms) give the same words.
    Keys are prepared again for the integer arithmetic from the same
    coefficients, and a key of the other form is refused."""
    import itertools
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu.ring import prepared_pairs, to_numpy
    from cipherpu.tune import BUILT_FUSED
    c = CKKS(CKKSParams(N=N), seed=23)
    sk = c.keygen()
    rng = np.random.default_rng(N + 1)
    z = rng.uniform(-1, 1, (3, N // 2))
    a, b = (c.encrypt(c.encode(v), sk) for v in (z, z[::-1].copy()))
    d = [c.ring.tensor(np.stack([rng.integers(0, q, (3, N), dtype=np.uint64) for q in c.primes[:3]], axis=1))
         for _ in range(3)]

    def run(rlk, gks):
        return ([to_numpy(c.hmult(a, b, rlk).data), to_numpy(c.relin(*d, rlk)), to_numpy(c.rotate(a, gks[0]).data),
                 to_numpy(c.rotate_many(a, [gks[0], gks[1], gks[0]]).data)] +
                [to_numpy(r.data) for r in c.rotate_hoisted(a, gks)] +
                [to_numpy(r.data) for r in c.rotate_hoisted(a, gks, eval_form=True)])

    rlk = c.relin_key(sk)
    gks = [c.galois_key(sk, s) for s in (1, 3)]
    ref = run(rlk, gks)
    pr, p, ps, pp = _oracle_args(c)
    assert np.array_equal(ref[0][2], OC.hmult(to_numpy(a.data[2]), to_numpy(b.data[2]), to_numpy(rlk.coeff),
                                              pr, p, ps, pp))
    c.ring.set_arith(False)
    with pytest.raises(ValueError):
        c.hmult(a, b, rlk)
    q_all = np.array(c.primes, dtype=np.uint64).reshape(1, 1, -1, 1)
    irlk = c.load_switch_key(rlk.coeff)
    igks = [c.load_switch_key(g.coeff, g.galois) for g in gks]
    # the same words in the integer form: (w, floor(w 2^63 / q))
    words = to_numpy(rlk.prepared).view(np.float64)     # FP64 key rows: the words as doubles
    assert np.array_equal(words, np.floor(words)) and np.all(words < q_all.astype(np.float64))
    words = words.astype(np.uint64)
    assert np.array_equal(to_numpy(irlk.prepared), prepared_pairs(words, q_all, fp64=False))
    logN = N.bit_length() - 1
    for front, tensor, ks in itertools.product((0, 1), BUILT_FUSED[logN], BUILT_FUSED[logN]):
        c.ring.set_front(bool(front))
        c.ring.set_threads(0, ks, tensor)
        c.ring.set_ks_order(bool(ks & 512))
        got = run(irlk, igks)
        for k, (g, r) in enumerate(zip(got, ref)):
            assert np.array_equal(g, r), (k, front, tensor, ks)


WIDE = [CHAIN4[0], 2199023288321, CHAIN4[2], 4398047232001]   # 2^44.., 2^41.., 2^44.., 2^42.. (= 1 mod 2^15)


def test_wide_chain_bit_exact(gpu):
    """A chain whose primes differ by more than a factor of two (2^41 to
    2^44, special prime 2^44): a residue mod one prime is no longer below
    twice another, so the digit lift, ModDown and rescale take their full
    reductions instead of one conditional subtraction.  HMult at two levels,
    a rotation and hoisted rotations against the oracle, bit for bit; the
    evaluation-form hoisted rotation against the transform of the
    coefficient form."""
    from cipherpu.ckks import CKKS, CKKSParams, Ciphertext
    from cipherpu.ring import to_numpy
    for q in WIDE[1::2]:
        assert all(pow(a, q - 1, q) == 1 for a in (2, 3, 5, 7)) and q % (1 << 15) == 1
    c = CKKS(CKKSParams(N=4096, primes=WIDE), seed=51)
    assert max(c.primes) >= 2 * min(c.primes)
    sk = c.keygen()
    rlk = c.relin_key(sk)
    pr, p, ps, pp = _oracle_args(c)
    key = to_numpy(rlk.coeff)
    rng = np.random.default_rng(51)
    for l in (4, 3):
        mk = lambda: np.stack([rng.integers(0, q, (2, 2, c.N), dtype=np.uint64) for q in c.primes[:l]], axis=2)
        a, b = mk(), mk()
        out = to_numpy(c.hmult(Ciphertext(c.ring.tensor(a), 1.0), Ciphertext(c.ring.tensor(b), 1.0), rlk).data)
        for i in range(2):
            assert np.array_equal(out[i], OC.hmult(a[i], b[i], key, pr[:l], p, ps[:l], pp)), (l, i)
    ct = c.encrypt(c.encode(rng.uniform(-1, 1, (2, c.N // 2))), sk)
    gks = [c.galois_key(sk, s) for s in (1, 9)]
    rot = to_numpy(c.rotate(ct, gks[0]).data)
    for i in range(2):
        want = OC.rotate(to_numpy(ct.data[i]), to_numpy(gks[0].coeff), pr, p, ps, pp, gks[0].galois)
        assert np.array_equal(rot[i], want), i
    coef = c.rotate_hoisted(ct, gks)
    ev = c.rotate_hoisted(ct, gks, eval_form=True)
    for gk, co, e in zip(gks, coef, ev):
        got = to_numpy(co.data)
        for i in range(2):
            want = OC.rotate_hoisted(to_numpy(ct.data[i]), to_numpy(gk.coeff), pr, p, ps, pp, gk.galois)
            assert np.array_equal(got[i], want), (gk.galois, i)
        assert np.array_equal(to_numpy(e.data).reshape(-1, 4, c.N),
                              to_numpy(c.ring.ntt(co.data.reshape(-1, 4, c.N)))), gk.galois


def _is_prime(n):
    if n < 2:
        return False
    d, s = n - 1, 0
    while d % 2 == 0:
        d, s = d // 2, s + 1
    for a in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
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


def test_long_chain_hmult_bit_exact(gpu):
    """HMult on a chain of 10 primes (2^44 + 2^15 j + 1) at levels 10, 9 and
    8 against the oracle: ModDown + rescale runs its compile-time limb-count
    kernel for up to 8 chain limbs and the general kernel above that, and
    both must give the oracle's words."""
    from cipherpu.ckks import CKKS, CKKSParams, Ciphertext
    from cipherpu.ring import to_numpy
    primes = [q for q in ((1 << 44) + (j << 15) + 1 for j in range(1, 400)) if _is_prime(q)][:11]
    assert len(primes) == 11
    chain, special = primes[:10], primes[10]
    c = CKKS(CKKSParams(N=4096, primes=chain, special=special), seed=61)
    sk = c.keygen()
    rlk = c.relin_key(sk)
    key = to_numpy(rlk.coeff)
    roots = c.ring.roots
    rng = np.random.default_rng(61)
    for l in (10, 9, 8):
        mk = lambda: np.stack([rng.integers(0, q, (2, 2, c.N), dtype=np.uint64) for q in chain[:l]], axis=2)
        a, b = mk(), mk()
        out = to_numpy(c.hmult(Ciphertext(c.ring.tensor(a), 1.0), Ciphertext(c.ring.tensor(b), 1.0), rlk).data)
        assert out.shape == (2, 2, l - 1, c.N)
        for i in range(2):
            want = OC.hmult(a[i], b[i], key, chain[:l], special, roots[:l], roots[10])
            assert np.array_equal(out[i], want), (l, i)


def test_hoisted_placement_never_changes_a_word(gpu):
    """The hoisted key switch under every workgroup placement (tile sizes 1
    to the whole batch, ragged last tiles, with and without the per-XCD
    deal, on grids that are and are not multiples of 8) gives the words of
    the default placement, in both output forms."""
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu.ring import to_numpy
    c = CKKS(CKKSParams(N=8192), seed=41)
    sk = c.keygen()
    rng = np.random.default_rng(41)
    for B, steps in ((5, (1, 2, 7)), (4, (3, 9))):          # grids 5*3*5 = 75 and 5*2*4 = 40
        gks = [c.galois_key(sk, s) for s in steps]
        ct = c.encrypt(c.encode(rng.uniform(-1, 1, (B, c.N // 2))), sk)
        c.ring.set_hoist_order(0, False)
        ref = [[to_numpy(r.data) for r in c.rotate_hoisted(ct, gks, eval_form=ev)] for ev in (False, True)]
        for btile in range(1, B + 2):
            for xcd in (False, True):
                c.ring.set_hoist_order(btile, xcd)
                for ev in (False, True):
                    got = [to_numpy(r.data) for r in c.rotate_hoisted(ct, gks, eval_form=ev)]
                    assert all(np.array_equal(g, r) for g, r in zip(got, ref[ev])), (B, btile, xcd, ev)
        c.ring.set_hoist_order(0, False)


@pytest.mark.parametrize("N", [4096, 8192, 16384])
def test_rotate_hoisted_eval_form_is_ntt_of_coefficient_form(gpu, N):
    """Hoisted rotations with ModDown in the evaluation domain give the
    forward transform of the coefficient-form hoisted rotations, word for
    word (the evaluation-domain ModDown is the image of the coefficient one
    under the linear map NTT_i), at every level and for several steps."""
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu.ring import to_numpy
    c = CKKS(CKKSParams(N=N), seed=31)
    sk = c.keygen()
    gks = [c.galois_key(sk, s) for s in (1, 5, N // 4 - 1)]
    rng = np.random.default_rng(N + 7)
    for l in (4, 3, 2, 1):
        ct = c.encrypt(c.encode(rng.uniform(-1, 1, (3, N // 2)), level=l), sk)
        coef = c.rotate_hoisted(ct, gks)
        ev = c.rotate_hoisted(ct, gks, eval_form=True)
        for r in range(len(gks)):
            want = to_numpy(c.ring.ntt(coef[r].data.reshape(-1, l, N)))
            assert np.array_equal(to_numpy(ev[r].data).reshape(-1, l, N), want), (N, l, r)


def test_single_modulus_rlwe_drop_ins(gpu):
    """golden_api's exp_e24/exp_e11 RLWE helpers (products on the GPU)
    against the oracle under the same numpy seeds: bit-exact for the integer
    forms, equal floats for the float64 pair; both mask-draw branches
    (q below and above 2^31) at the reference's N = 1024."""
    from cipherpu import golden_api as G
    for q in (998244353, CHAIN4[0]):
        n = 1024
        rng = np.random.default_rng(q % 977)
        s = rng.integers(-1, 2, n)
        e = rng.integers(-3, 4, n)
        m = rng.integers(-10 ** 6, 10 ** 6, n)
        assert np.array_equal(G.neg_conv_modq(m, s, q), O.neg_conv_modq(m, s, q))
        np.random.seed(21)
        ct = G.rlwe_encrypt_mint(m, s, e, q)
        np.random.seed(21)
        want = O.rlwe_encrypt_mint(m, s, e, q)
        assert np.array_equal(ct[0], want[0]) and np.array_equal(ct[1], want[1])
        assert np.array_equal(G.decrypt_exact_int(ct, s, q), O.decrypt_exact_int(want, s, q))
        np.random.seed(22)
        (c2, m2) = G.reencrypt(ct, s, q, 2.0 ** 20, np.random.default_rng(3))
        np.random.seed(22)
        (w2, wm2) = O.reencrypt(want, s, q, 2.0 ** 20, np.random.default_rng(3))
        assert np.array_equal(m2, wm2) and np.array_equal(c2[0], w2[0]) and np.array_equal(c2[1], w2[1])
        mf = rng.uniform(-1, 1, n)
        np.random.seed(23)
        cf = G.rlwe_encrypt(mf, s, e, q, 2.0 ** 12)
        np.random.seed(23)
        wf = O.rlwe_encrypt(mf, s, e, q, 2.0 ** 12)
        assert np.array_equal(cf[0], wf[0]) and np.array_equal(cf[1], wf[1])
        assert np.array_equal(G.rlwe_decrypt(cf, s, q), O.rlwe_decrypt(wf, s, q))
    x = np.arange(1024) - 512
    assert np.array_equal(G.neg_conv_modq(x, x[::-1], 1000003), O.neg_conv_modq(x, x[::-1], 1000003))


def test_coefficient_encoding_drop_ins(gpu):
    """golden_api's exp_e24 coefficient-encoding primitives against the
    oracle at the reference's N = 1024, q = 998244353 and at a chain prime."""
    from cipherpu import golden_api as G
    for q, scale in ((998244353, 2.0 ** 8), (CHAIN4[1], 2.0 ** 14)):
        n = 1024
        rng = np.random.default_rng(q % 331)
        s = rng.integers(-1, 2, n)
        e = rng.integers(-3, 4, n)
        x = rng.uniform(-1, 1, n)
        np.random.seed(8)
        ct = O.rlwe_encrypt_mint(np.round(scale * x).astype(np.int64), s, e, q)
        p = rng.integers(-50, 50, n)
        got, want = G.mul_plain_exact(ct, p, q), O.mul_plain_exact(ct, p, q)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])
        for nu in (1, 77, 1023, 2047):
            got, want = G.rotate_monomial(ct, nu, q), O.rotate_monomial(ct, nu, q)
            assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), nu
        assert np.array_equal(G.decrypt_exact(ct, s, q), O.decrypt_exact(ct, s, q))
        w = rng.uniform(-0.1, 0.1, 100)
        idx = list(rng.integers(0, n, 100))
        assert np.array_equal(G.scatter_w(idx, w, scale, q, n), O.scatter_w(idx, w, scale, q, n))
        assert G.inner_product(ct, w, s, q, scale) == O.inner_product(ct, w, s, q, scale)
        assert G.inner_product(ct, w, s, q, scale, idx) == O.inner_product(ct, w, s, q, scale, idx)
        assert G.coef_sum(ct, 64, s, q, scale) == O.coef_sum(ct, 64, s, q, scale)


@pytest.mark.parametrize("fp64", [False, True])
@pytest.mark.parametrize("N,threads", [(4096, 256), (4096, 512), (8192, 256), (8192, 512), (8192, 1024),
                                       (16384, 512), (16384, 1024)])
def test_transforms_and_products_over_many_rows(gpu, N, threads, fp64):
    """Every built transform shape over 512 rows at once (enough workgroups
    that every CU runs several in turn): the forward-inverse round trip
    returns every row, sampled forward rows and products equal the C oracle.
    Catches LDS ordering faults between exchanges, which show only when
    waves of one workgroup drift apart."""
    from cipherpu.ring import RingContext, bitrev_, to_numpy
    q = CHAIN4[0]
    ring = RingContext(N, [q])
    ring.set_threads(threads, 0, 0)
    ring.set_arith(fp64)
    rng = np.random.default_rng(N + threads)
    rows = 512
    a = rng.integers(0, q, (rows, 1, N), dtype=np.uint64)
    b = rng.integers(0, q, (rows, 1, N), dtype=np.uint64)
    A = ring.tensor(a)
    X = ring.ntt_(A.clone())
    assert np.array_equal(to_numpy(ring.intt_(X.clone())), a)
    psi = ring.roots[0]
    Xn = to_numpy(bitrev_(X.view(rows, N).clone()))
    for r in (0, rows // 2, rows - 1):
        assert np.array_equal(Xn[r], OC.ntt_negacyclic(a[r, 0], N, q, psi)), r
    P = to_numpy(ring.poly_mul(A, ring.tensor(b)))
    for r in (0, 137, rows - 1):
        assert np.array_equal(P[r, 0], OC.polymul_ntt(a[r, 0], b[r, 0], q, psi)), r


@pytest.mark.parametrize("N,threads", [(4096, 128), (4096, 512), (8192, 512), (8192, 1024), (16384, 1024)])
def test_fp64_products_at_extreme_operands(gpu, N, threads):
    """The FP64 pipeline's products at the operands that push its bounds
    hardest, for the 45-bit prime just below 2^45 and a chain prime: all
    coefficients q - 1, alternating 0 / q - 1, a single q - 1, and random
    rows; every product equals the integer pipeline's and sampled ones the
    C oracle's."""
    from cipherpu.ring import RingContext, to_numpy
    for q, psi in ((Q1, None), (CHAIN4[3], None)):
        if N != 4096 and q == Q1:
            continue
        ring = RingContext(N, [q], roots=[PSI1] if q == Q1 else None)
        ring.set_threads(threads, 0, 0)
        rng = np.random.default_rng(N * 7 + threads)
        rows = 8
        a = rng.integers(0, q, (rows, 1, N), dtype=np.uint64)
        b = rng.integers(0, q, (rows, 1, N), dtype=np.uint64)
        a[0], b[0] = q - 1, q - 1
        a[1, 0, ::2], a[1, 0, 1::2] = 0, q - 1
        b[1] = q - 1
        a[2] = 0
        a[2, 0, N - 1] = q - 1
        b[3] = 0
        A, B = ring.tensor(a), ring.tensor(b)
        ring.set_arith(False)
        want = to_numpy(ring.poly_mul(A, B))
        ring.set_arith(True)
        got = to_numpy(ring.poly_mul(A, B))
        assert np.array_equal(got, want)
        for r in (0, 1, 2, 5):
            assert np.array_equal(got[r, 0], OC.polymul_ntt(a[r, 0], b[r, 0], q, ring.roots[0])), (q, r)


def test_prepared_pairs_are_correctly_rounded(ck):
    """The device's prepared pairs (ckks_pt_prepare, ckks_key_prepare) hold
    w and RN(w / q) bit for bit as IEEE double division on the host gives
    them, for random rows and the words 0, 1, q - 2 and q - 1."""
    import ctypes
    import torch
    from cipherpu._lib import call
    from cipherpu.ring import prepared_pairs, to_numpy
    c = ck[0]
    l = c.L
    rng = np.random.default_rng(77)
    pt = np.stack([rng.integers(0, q, (3, c.N), dtype=np.uint64) for q in c.primes[:l]], axis=1)
    for i, q in enumerate(c.primes[:l]):
        pt[0, i, :4] = [0, 1, q - 2, q - 1]
    t = c.ring.tensor(pt)
    prep = torch.empty(*t.shape, 2, dtype=torch.uint64, device=c.device)
    p = lambda x: ctypes.c_void_p(x.data_ptr())
    call("ckks_pt_prepare", c.ring.handle, p(t), p(prep), 3, l, 0, c.ring.stream())
    ev = to_numpy(t)        # transformed in place: the evaluation-form words
    got = to_numpy(prep)
    q = np.array(c.primes[:l], dtype=np.float64).reshape(1, l, 1)
    assert np.array_equal(got, prepared_pairs(ev, q))
    # the words themselves, placed directly
    raw = c.ring.tensor(pt)
    call("ckks_pt_prepare", c.ring.handle, p(raw), p(prep), 3, l, 0, c.ring.stream())
    assert np.array_equal(to_numpy(prep)[..., 0].view(np.float64), to_numpy(raw).astype(np.float64))
