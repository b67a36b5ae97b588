"""Parity at the benchmark's full sizes.

  C2: 256 HMult pairs at N = 8192, L = 4 -- every decrypted product within
      1e-6 of the FP64 slot-wise product, and the first, middle and last
      pair bit-exact against the C oracle.
  C3: the bench's 2048 mixed operations (N = 4096 / 8192, levels 1..4,
      HMult / rotation / plaintext product) -- the first, middle and last
      row of every bucket bit-exact against the C oracle, and the captured
      graph's replay equal to the eager launches.
  C4 / C5: the encrypted MLP and CNN on 4096 images at gear 0 -- output SNR
      above 60 dB against the FP64 networks; and the CNN's committed
      learned controller (gearnet_cnn.npz) meeting its target with every
      chunk either cleared by its prediction plus the margin or decided by
      the plain monitor.
"""
import numpy as np
import pytest

import oracle_c as OC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from cipherpu import _lib
    _lib.load()
    return torch


def test_c2_full_batch(torch_gpu):
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu.ring import to_numpy
    B = 256
    ck = CKKS(CKKSParams(), seed=2024)
    sk = ck.keygen()
    rlk = ck.relin_key(sk)
    rng = np.random.default_rng(7)
    za = rng.uniform(-1, 1, (B, 4096))
    zb = rng.uniform(-1, 1, (B, 4096))
    cta = ck.encrypt(ck.encode(za), sk)
    ctb = ck.encrypt(ck.encode(zb), sk)
    out = ck.hmult(cta, ctb, rlk)
    dec = to_numpy(ck.decode(ck.decrypt(out, sk), out.scale)).real
    err = np.abs(dec - za * zb).max(axis=1)
    print("C2 max-abs error over 256 pairs:", float(err.max()))
    assert err.max() < 1e-6
    o, a, b, key = to_numpy(out.data), to_numpy(cta.data), to_numpy(ctb.data), to_numpy(rlk.coeff)
    for i in (0, B // 2, B - 1):
        want = OC.hmult(a[i], b[i], key, ck.primes[:4], ck.primes[4], ck.ring.roots[:4], ck.ring.roots[4])
        assert np.array_equal(o[i], want), i


def _rows(bk):
    return sorted({0, bk.batch // 2, bk.batch - 1})


def test_c3_full_mixed_batch(torch_gpu):
    torch = torch_gpu
    from cipherpu.ring import to_numpy
    from cipherpu.sched import BatchScheduler, make_workload
    ops = make_workload(2048, seed=0)
    s = BatchScheduler(ops, seed=1)
    s.fill_random(seed=2)
    s.tune(reps=1)
    s.run_eager()
    torch.cuda.synchronize()
    assert sum(bk.batch for bk in s.buckets) == 2048
    checked = 0
    for bk in s.buckets:
        ck = s.ck[bk.N]
        l = bk.limbs
        pr, p, ps, pp = ck.primes[:l], ck.primes[ck.L], ck.ring.roots[:l], ck.ring.roots[ck.L]
        out, a = to_numpy(bk.out), to_numpy(bk.a)
        for r in _rows(bk):
            if bk.kind == "hmult":
                want = OC.hmult(a[r], to_numpy(bk.b)[r], to_numpy(s.rlk[bk.N].coeff), pr, p, ps, pp)
            elif bk.kind == "rotate":
                gk = s.gk[(bk.N, bk.steps[r])]
                want = OC.rotate(a[r], to_numpy(gk.coeff), pr, p, ps, pp, gk.galois)
            else:
                b = to_numpy(bk.b)
                want = np.stack([np.stack([OC.polymul_ntt(a[r, c, i], b[2 * r + c, i], pr[i], ps[i])
                                           for i in range(l)]) for c in range(2)])
            assert np.array_equal(out[r], want), (bk.kind, bk.N, bk.level, r)
            checked += 1
    print("C3 rows checked against the oracle:", checked, "in", len(s.buckets), "buckets")
    ref = [o.clone() for o in s.outputs()]
    s.capture()
    for rep in range(2):
        for o in s.outputs():
            o.fill_(0)
        torch.cuda.synchronize()
        s.replay()
        torch.cuda.synchronize()
        for i, (o, r) in enumerate(zip(s.outputs(), ref)):
            assert torch.equal(o, r), (rep, i)


@pytest.mark.parametrize("kind", ["mlp", "cnn"])
def test_network_full_batch_gear0(torch_gpu, kind):
    from cipherpu import gear, nn
    from cipherpu.ckks import CKKS, CKKSParams
    ck = CKKS(CKKSParams(), seed=99)
    sk = ck.keygen()
    if kind == "mlp":
        net = nn.EncryptedMLP(ck, sk, *nn.make_mlp(0))
    else:
        from cipherpu import cnn
        net = cnn.EncryptedCNN(ck, sk, *cnn.make_cnn(0))
    X = nn.synthetic_images(4096, seed=4096)
    ref = net.reference_outputs(X)
    r = gear.static_loop(net, X, 0, chunk_cts=64)
    snr = gear.snr_db(ref, r.outputs)
    print(kind, "4096 samples, gear 0: SNR dB", round(snr, 2))
    assert r.outputs.shape == (4096, 10)
    assert snr > 60


def test_cnn_learned_controller(torch_gpu):
    """The committed CNN controller: it loads, and on 1024 fresh images at a
    3 dB target every chunk's gear either had a predicted SNR of at least
    target + margin (3 dB) or was decided by the plain monitor, and the
    measured output SNR meets the target."""
    from cipherpu import cnn, gear, nn
    from cipherpu.ckks import CKKS, CKKSParams
    ck = CKKS(CKKSParams(), seed=99)
    sk = ck.keygen()
    net = cnn.EncryptedCNN(ck, sk, *cnn.make_cnn(0))
    ctl = gear.GearNet(seed=0)
    assert gear.load_trained(ctl, gear.TRAINED_PATHS["cnn"]) == "gearnet_cnn.npz"
    X = nn.synthetic_images(1024, seed=77)
    target = 3.0
    chunks = gear._setup(net, X, 64)
    table = gear._gear_table(net, ctl, chunks)
    r = gear.learned_loop(net, ctl, X, target, chunk_cts=64)
    prev = 0
    for i, g in enumerate(r.gears):
        pred = table[prev, i]
        chosen = gear.choose_gear(pred, target)
        if chosen is None:
            assert r.fallbacks > 0
        else:
            assert g == chosen and pred[g] >= target + gear.MARGIN_DB
        prev = g
    ref = net.reference_outputs(X)
    snr = gear.snr_db(ref, r.outputs)
    print("CNN learned loop: gears", r.gears, "SNR dB", round(snr, 2), "fallbacks", r.fallbacks)
    assert snr >= target
    assert r.monitor_calls == 0 or r.fallbacks > 0
