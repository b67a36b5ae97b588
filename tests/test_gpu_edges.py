"""Edge cases of the C ABI on the GPU: zero-row calls, ragged batch sizes,
extreme coefficient values and the hoisted-rotation count limits, each
against the oracle where a result is produced."""
import ctypes

import numpy as np
import pytest

import oracle_c as OC

pytestmark = pytest.mark.gpu

SENTINEL = 0x25A5A5A5A5A5A5A5


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from cipherpu.ckks import CKKS, CKKSParams
    c = CKKS(CKKSParams(N=4096), seed=61)
    sk = c.keygen()
    return c, sk, c.relin_key(sk)


def _ptr(t, off=0):
    return ctypes.c_void_p(t.data_ptr() + off)


def _oracle_args(c):
    return c.primes[:4], c.primes[4], c.ring.roots[:4], c.ring.roots[4]


def _filled(c, *shape):
    """An int64 tensor of SENTINEL words (compared through this view; the
    library sees the same memory as uint64)."""
    import torch
    return torch.full(shape, SENTINEL, dtype=torch.int64, device=c.device)


def test_zero_row_calls_return_ok_and_touch_nothing(env):
    """Every batched entry point accepts a zero-row call, returns 0 and
    leaves its output (or in-place data) as it was."""
    import torch
    from cipherpu._lib import call
    c, sk, rlk = env
    l, N, st = 4, c.N, c.ring.stream()
    gk = c.galois_key(sk, 1)
    keys = torch.stack([c.hoist_key(gk), c.hoist_key(gk)]).contiguous()
    gal = (ctypes.c_uint64 * 2)(gk.galois, gk.galois)
    src = _filled(c, 2, 2, l, N)
    out = _filled(c, 2, 2, l, N)
    ws = torch.zeros(1 << 16, dtype=torch.uint64, device=c.device)
    calls = [
        ("ckks_ntt_fwd", (c.ring.handle, _ptr(out), 0, l, 0, 0, st)),
        ("ckks_ntt_inv", (c.ring.handle, _ptr(out), 0, l, 0, 0, st)),
        ("ckks_poly_mul", (c.ring.handle, _ptr(src), _ptr(src), _ptr(out), 0, l, 0, st)),
        ("ckks_rescale", (c.ring.handle, _ptr(src), _ptr(out), 0, l, 0, st)),
        ("ckks_hmult", (c.ring.handle, _ptr(src), _ptr(src), _ptr(rlk.prepared), c.L + 1, _ptr(out), 0, l, c.sp,
                        _ptr(ws), st)),
        ("ckks_relin", (c.ring.handle, _ptr(src), _ptr(src), _ptr(rlk.prepared), c.L + 1, _ptr(out), 0, l, c.sp,
                        _ptr(ws), st)),
        ("ckks_rotate", (c.ring.handle, _ptr(src), _ptr(gk.prepared), c.L + 1, _ptr(out), 0, l, c.sp, gk.galois,
                         _ptr(ws), st)),
        ("ckks_rotate_hoisted", (c.ring.handle, _ptr(src), _ptr(keys), c.L + 1, gal, 2, _ptr(out), 0, l, c.sp,
                                 _ptr(ws), st)),
        ("ckks_rotate_hoisted_eval", (c.ring.handle, _ptr(src), _ptr(keys), c.L + 1, gal, 2, _ptr(out), 0, l, c.sp,
                                      _ptr(ws), st)),
        ("ckks_automorphism", (c.ring.handle, _ptr(src), _ptr(out), 0, l, 0, 5, st)),
    ]
    for name, args in calls:
        assert call(name, *args) == 0, name
        torch.cuda.synchronize()
        assert bool((out == SENTINEL).all()), name


def test_hmult_ragged_batches_bit_exact(env):
    """Batches of 1, 3 and 7 pairs (grids that fill no round of the chip),
    at levels 4 and 2, against the oracle for the first and last pair."""
    from cipherpu.ckks import Ciphertext
    from cipherpu.ring import to_numpy
    c, sk, rlk = env
    key = to_numpy(rlk.coeff)
    pr, p, ps, pp = _oracle_args(c)
    rng = np.random.default_rng(62)
    for B in (1, 3, 7):
        for l in (4, 2):
            mk = lambda: np.stack([rng.integers(0, q, (B, 2, c.N), dtype=np.uint64) for q in c.primes[:l]], axis=2)
            a, b = mk(), mk()
            got = to_numpy(c.hmult(Ciphertext(c.ring.tensor(a), 1.0), Ciphertext(c.ring.tensor(b), 1.0), rlk).data)
            for i in {0, B - 1}:
                assert np.array_equal(got[i], OC.hmult(a[i], b[i], key, pr[:l], p, ps[:l], pp)), (B, l, i)


def test_extreme_coefficients_bit_exact(env):
    """Inputs at the ends of the residue range -- every coefficient q - 1,
    every coefficient 0, and q - 1 / 0 alternating -- through HMult and a
    rotation, against the oracle: the lazy bounds of the transforms and the
    key products are reached with these, not with random residues."""
    from cipherpu.ckks import Ciphertext
    from cipherpu.ring import to_numpy
    c, sk, rlk = env
    gk = c.galois_key(sk, 3)
    key = to_numpy(rlk.coeff)
    pr, p, ps, pp = _oracle_args(c)
    l, N = 4, c.N
    qs = np.array(c.primes[:l], dtype=np.uint64)[None, :, None]
    alt = (np.arange(N) % 2 == 0)[None, None, :]
    patterns = {
        "max": np.broadcast_to(qs - 1, (2, l, N)),
        "zero": np.zeros((2, l, N), dtype=np.uint64),
        "alternating": np.broadcast_to(np.where(alt, qs - 1, 0), (2, l, N)),
    }
    cts = np.stack([np.ascontiguousarray(v) for v in patterns.values()]).astype(np.uint64)   # [3][2][l][N]
    ct = Ciphertext(c.ring.tensor(cts), 1.0)
    prod = to_numpy(c.hmult(ct, Ciphertext(c.ring.tensor(cts[::-1].copy()), 1.0), rlk).data)
    rot = to_numpy(c.rotate(ct, gk).data)
    for i, name in enumerate(patterns):
        want = OC.hmult(cts[i], cts[2 - i], key, pr, p, ps, pp)
        assert np.array_equal(prod[i], want), ("hmult", name)
        want = OC.rotate(cts[i], to_numpy(gk.coeff), pr, p, ps, pp, gk.galois)
        assert np.array_equal(rot[i], want), ("rotate", name)


def test_hoisted_rotation_count_limits(env):
    """32 rotations in one hoisted call are accepted and each matches the
    oracle's hoisted rotation; 33 and 0 are rejected as bad shapes."""
    import torch
    from cipherpu._lib import CkksError, call
    from cipherpu.ring import to_numpy
    c, sk, rlk = env
    rng = np.random.default_rng(63)
    ct = c.encrypt(c.encode(rng.uniform(-1, 1, (1, c.N // 2)), level=2), sk)
    gks = [c.galois_key(sk, s) for s in range(1, 33)]
    outs = c.rotate_hoisted(ct, gks)
    pr, p, ps, pp = _oracle_args(c)
    for r in (0, 13, 31):
        want = OC.rotate_hoisted(to_numpy(ct.data[0]), to_numpy(gks[r].coeff), pr[:2], p, ps[:2], pp,
                                 gks[r].galois)
        assert np.array_equal(to_numpy(outs[r].data[0]), want), r
    keys = torch.stack([c.hoist_key(g) for g in gks] + [c.hoist_key(gks[0])]).contiguous()
    gal = (ctypes.c_uint64 * 33)(*([g.galois for g in gks] + [gks[0].galois]))
    out = torch.empty(33, 1, 2, 2, c.N, dtype=torch.uint64, device=c.device)
    words = (call("ckks_workspace_hoisted_bytes", c.ring.handle, 1, 2, 32) + 7) // 8
    ws = torch.empty(2 * words, dtype=torch.uint64, device=c.device)
    for R in (33, 0):
        with pytest.raises(CkksError) as e:
            call("ckks_rotate_hoisted", c.ring.handle, _ptr(ct.data), _ptr(keys), c.L + 1, gal, R, _ptr(out), 1, 2,
                 c.sp, _ptr(ws), c.ring.stream())
        assert e.value.code == -4, R
