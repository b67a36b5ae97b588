"""The matrix-core probe's arithmetic, on the CPU: the tables of
cipherpu.dft64_probe and a word-exact emulation of both kernels of
k_mfma_probe.hip (64-bit wrapping arithmetic as the device does it) against
the 64-point negacyclic transform's closed form."""
import numpy as np
import pytest

import ckks_oracle as O
from cipherpu import dft64_probe as P
from cipherpu.params import CHAIN4, SPECIAL_P

M64 = (1 << 64) - 1


def _tables(q):
    return P.build_tables(q, O.root_of_unity(q, 128))


@pytest.fixture(scope="module")
def T():
    return _tables(CHAIN4[0])


def test_matrix_is_the_negacyclic_transform_in_bit_reversed_order(T):
    q, psi = T.q, T.psi
    for i in range(64):
        e = 2 * P.bitrev6(i) + 1
        assert T.G[i] == [pow(psi, e * k, q) for k in range(64)]


def test_balanced_digits_reassemble(T):
    w = np.array([1 << (8 * t) for t in range(P.PLANES)], dtype=object)
    assert all(int((T.fdig[i, k].astype(object) * w).sum()) == T.G[i][k] for i in range(64) for k in range(64))
    assert T.fdig[..., :-1].min() >= -128 and T.fdig[..., :-1].max() <= 127


def _perm(a, b, sel):
    v = (a << 32) | b
    return sum(((v >> (8 * ((sel >> (8 * k)) & 0xff))) & 0xff) << (8 * k) for k in range(4))


def test_byte_planes_transpose():
    """The kernel's 12-permute transpose (v_perm_b32 semantics) against the
    byte extraction it replaces."""
    rng = np.random.default_rng(1)
    for _ in range(50):
        w = [int(x) for x in rng.integers(0, 1 << 48, 4, dtype=np.uint64)]
        lo = [x & 0xffffffff for x in w]
        hi = [x >> 32 for x in w]
        a01, b01 = _perm(lo[1], lo[0], 0x05010400), _perm(lo[1], lo[0], 0x07030602)
        a23, b23 = _perm(lo[3], lo[2], 0x05010400), _perm(lo[3], lo[2], 0x07030602)
        c01, c23 = _perm(hi[1], hi[0], 0x05010400), _perm(hi[3], hi[2], 0x05010400)
        planes = [_perm(a23, a01, 0x05040100), _perm(a23, a01, 0x07060302), _perm(b23, b01, 0x05040100),
                  _perm(b23, b01, 0x07060302), _perm(c23, c01, 0x05040100), _perm(c23, c01, 0x07060302)]
        for s in range(6):
            assert planes[s] == sum(((w[b] >> (8 * s)) & 0xff) << (8 * b) for b in range(4))


def _mulhi63_lazy(a, ws):
    a0, a1, w0, w1 = a & 0xffffffff, a >> 32, ws & 0xffffffff, ws >> 32
    m = (a1 * w0 + a0 * w1) & M64
    return (((a1 * w1) << 1) + (m >> 31)) & M64


def _sx(w):
    w &= 0xffffffff
    return (w - (1 << 32) if w >> 31 else w) & M64


def emulate_mfma(T, x, reps):
    """k_dft64_mfma word for word: z = x + BIAS, six byte planes as signed
    digits, the 11 int32 sums, the recombination and lazy Shoup product."""
    q, nq = T.q, (1 << 64) - T.q
    z = [[int(v) + P.BIAS for v in col] for col in x.T]
    for _ in range(reps):
        d = np.array([[[((v >> (8 * s)) & 0xff) - 128 for s in range(6)] for v in col] for col in z],
                     dtype=np.int64)                                 # [V][64][6]
        A = np.zeros((11, 64, len(z)), dtype=np.int64)
        for t in range(6):
            for s in range(6):
                A[t + s] += T.fdig[:, :, t] @ d[:, :, s].T
        assert np.abs(A).max() < 1 << 23
        out = []
        for c in range(len(z)):
            col = []
            for i in range(64):
                a = [int(A[k, i, c]) & 0xffffffff for k in range(11)]
                p0, p1 = (a[0] + (a[1] << 8)) & 0xffffffff, (a[2] + (a[3] << 8)) & 0xffffffff
                L = (_sx(p0) + (_sx(p1) << 16) + (a[4] << 32)) & M64
                q0 = (a[5] + (a[6] << 8)) & 0xffffffff
                q1 = (a[7] + (a[8] << 8)) & 0xffffffff
                q2 = (a[9] + (a[10] << 8)) & 0xffffffff
                Mw = (_sx(q0) + (_sx(q1) << 16) + (q2 << 32)) & M64
                lo = L & ((1 << 40) - 1)
                Ls = L - (1 << 64) if L >> 63 else L
                mp = (Mw + ((Ls >> 40) & M64) + T.kq) & M64
                assert 0 < mp < 1 << 62
                r = ((mp << 40) + _mulhi63_lazy(mp, T.ws40) * nq) & M64
                assert r < 4 * q
                col.append(r + lo)
            out.append(col)
        z = out
        assert max(max(c) for c in z) < 1 << 48
    return np.array([[(v + (-P.BIAS) % q) % q for v in col] for col in z], dtype=np.uint64).T


def emulate_radix2(T, x, reps):
    """k_dft64_radix2 word for word (ct_bfly_lazy with mul_shoup4_nq)."""
    q = T.q
    nq, four_q = (1 << 64) - q, 4 * q
    tw = [(pow(T.psi, P.bitrev6(k), q), (pow(T.psi, P.bitrev6(k), q) << 63) // q) for k in range(64)]
    out = []
    for col in x.T:
        a = [int(v) for v in col]
        for _ in range(reps):
            for m in (1, 2, 4, 8, 16, 32):
                t = 32 // m
                for i in range(m):
                    w, ws = tw[m + i]
                    for j in range(2 * i * t, 2 * i * t + t):
                        y = a[j + t]
                        prod = (y * w + _mulhi63_lazy(y, ws) * nq) & M64
                        a[j], a[j + t] = a[j] + prod, a[j] + four_q - prod
            assert max(a) < 1 << 62
        out.append([v % q for v in a])
    return np.array(out, dtype=np.uint64).T


@pytest.mark.parametrize("q", [CHAIN4[0], CHAIN4[3], SPECIAL_P])
@pytest.mark.parametrize("reps", [1, 3])
def test_emulated_kernels_match_exact_product(q, reps):
    T = _tables(q)
    rng = np.random.default_rng(reps)
    x = rng.integers(0, q, (64, 6), dtype=np.uint64)
    x[:, 0] = q - 1
    x[:, 1] = 0
    want = P.matmul_mod(T.G, x, q, reps)
    assert np.array_equal(emulate_mfma(T, x, reps), want)
    assert np.array_equal(emulate_radix2(T, x, reps), want)


def test_table_blob_layout(T):
    blob = T.blob
    assert blob.size == P.TABLE_BYTES
    tw = blob[:1024].view(np.uint64).reshape(64, 2)
    assert int(tw[1, 0]) == pow(T.psi, 32, T.q) and int(tw[1, 1]) == (int(tw[1, 0]) << 63) // T.q
    cst = blob[1024:2048].view(np.uint64)
    assert [int(v) for v in cst[:3]] == [T.kq, (-P.BIAS) % T.q, P.BIAS]
    fpk = blob[2048:].view(np.int8).reshape(2, 4, 6, 64, 16)
    for cm in (0, 1):
        for lane in (0, 17, 63):
            for j in (0, 5, 15):
                row = 16 * 2 + (lane & 15)
                assert fpk[cm, 2, 3, lane, j] == T.fdig[row, P.pi_index(cm, lane >> 4, j), 3]


def test_rejects_bad_parameters():
    q = CHAIN4[0]
    with pytest.raises(ValueError):
        P.build_tables(q, O.root_of_unity(q, 64))
    with pytest.raises(ValueError):
        P.build_tables(1 << 45 | 1, 3)
