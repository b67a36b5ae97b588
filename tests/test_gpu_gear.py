"""The learned gear controller on the GPU: forward/backward/Adam kernels
against the CPU oracle, training on monitor sweeps, and the ablation
against the plain decrypt-based monitor gear loop."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_gpu():
    import torch
    assert torch.cuda.is_available()
    return torch


def _rand_net(torch, bf16, seed=0):
    from cipherpu.gear import GearNet
    net = GearNet(seed=seed, bf16=bf16)
    g = torch.Generator().manual_seed(seed + 1)
    for k in range(3):      # non-zero biases so that every term is exercised
        net.b[k] = (torch.randn(net.b[k].shape, generator=g) * 0.1).to(net.dev)
    return net


@pytest.mark.parametrize("bf16", [False, True])
def test_forward_backward_match_oracle(torch_gpu, bf16):
    import gear_oracle as go
    torch = torch_gpu
    net = _rand_net(torch, bf16)
    rng = np.random.default_rng(5)
    B = 200
    X = rng.normal(0, 1, (B, 16)).astype(np.float32)
    T = rng.normal(0, 1, (B, 4)).astype(np.float32)
    params = [(W.cpu().numpy(), b.cpu().numpy()) for W, b in net.params()]
    Y, acts = net.forward(torch.from_numpy(X).cuda())
    Yo, acts_o = go.forward(params, X, bf16_hidden=bf16)
    y = Y.cpu().numpy()
    # FP32 sums in another order; with BF16 storage an activation can land on
    # the other side of a rounding boundary (one BF16 ulp, 2^-8 relative)
    tol = 2e-2 if bf16 else 1e-4
    assert np.max(np.abs(y - Yo)) <= tol * max(1.0, np.max(np.abs(Yo)))
    # gradients: the oracle's backward on the GPU's own stored activations
    acts_gpu = [a.float().cpu().numpy() for a in acts]
    if bf16:
        for a in acts_gpu[1:]:
            assert np.array_equal(a, go.bf16_round(a))
    dY, rl = net.loss_grad(Y, torch.from_numpy(T).cuda())
    grads = net.backward(acts, dY)
    gref = go.backward(params, acts_gpu, y, T)
    assert abs(float(rl.sum()) / B - go.mse(y, T)) <= 1e-5 * max(1.0, go.mse(y, T))
    for k in range(3):
        for which in (0, 1):
            got = grads[k][which].cpu().numpy()
            ref = gref[k][which]
            assert np.max(np.abs(got - ref)) <= 1e-4 * max(1e-3, np.max(np.abs(ref))), (k, which)


def test_adam_step_matches_formula(torch_gpu):
    torch = torch_gpu
    net = _rand_net(torch, False)
    rng = np.random.default_rng(6)
    p0 = [p.clone() for p in net.W + net.b]
    grads = [(torch.from_numpy(rng.normal(0, 1, W.shape).astype(np.float32)).cuda(),
              torch.from_numpy(rng.normal(0, 1, b.shape).astype(np.float32)).cuda()) for W, b in net.params()]
    lr, b1, b2, eps = 1e-2, 0.9, 0.999, 1e-8
    net.adam(grads, lr=lr)
    net.adam(grads, lr=lr)
    gs = [g[0] for g in grads] + [g[1] for g in grads]
    for p, q, g in zip(net.W + net.b, p0, gs):
        g = g.cpu().double().numpy()
        m = v = 0.0
        x = q.cpu().double().numpy()
        for t in (1, 2):
            m = b1 * m + (1 - b1) * g
            v = b2 * v + (1 - b2) * g * g
            x = x - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
        assert np.max(np.abs(p.cpu().double().numpy() - x)) < 1e-5


def test_training_fits_a_known_function(torch_gpu):
    torch = torch_gpu
    from cipherpu.gear import GearNet
    rng = np.random.default_rng(7)
    X = rng.normal(0, 1, (8192, 16))
    T = np.stack([np.abs(X[:, 0]) + X[:, 1], X[:, 2] * 2, np.maximum(X[:, 3], 0), X[:, 4] - X[:, 5]], 1)
    net = GearNet(seed=3)
    hist = net.fit(X, T, epochs=40, batch=512, lr=3e-3)
    assert hist[-1] < 0.05 * hist[0]
    pred = net.predict(X[:512]).cpu().numpy()
    assert np.sqrt(np.mean((pred - T[:512]) ** 2)) < 0.25 * T.std()


@pytest.fixture(scope="module")
def mlp(torch_gpu):
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu import nn
    ck = CKKS(CKKSParams(), seed=11)
    sk = ck.keygen()
    W1, W2 = nn.make_mlp(0)
    return nn.EncryptedMLP(ck, sk, W1, W2)


def test_monitor_features_and_gear_costs(mlp):
    from cipherpu import gear, nn
    import time
    import torch
    X = nn.synthetic_images(256, seed=21)
    ct = mlp.encrypt_inputs(X)
    netf = gear.network_features(mlp)
    f = gear.ciphertext_features(mlp.ck, ct, prev_gear=2, netf=netf).cpu().numpy()
    assert f.shape == (64, 16)
    # uniform residues: rms(c')/q -> 1/sqrt(12), mean(c')/q -> 1/4
    assert np.allclose(f[:, [0, 1]], 1 / math.sqrt(12), atol=0.01)
    assert np.allclose(f[:, 11], 0.25, atol=0.01)
    assert np.allclose(f[:, 2:11], netf[None, :].astype(np.float32))
    assert netf.shape == (9,) and np.all(netf[6:] <= 1.0)
    assert np.allclose(f[:, 12], 1.0) and np.allclose(f[:, 14], 2 / 3)
    snrs, costs = [], []
    for g in range(4):
        mlp.set_gears(g, g)
        mlp.forward(ct)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = mlp.forward(ct)
        torch.cuda.synchronize()
        costs.append(time.perf_counter() - t0)
        s, per = gear.monitor(mlp, out, X)
        snrs.append(s)
        assert per.shape == (64,)
    print("gear SNR dB", [round(s, 2) for s in snrs], "seconds per 256 samples", [round(c, 4) for c in costs])
    assert snrs[0] > 60                       # exact arithmetic: only CKKS noise
    assert snrs[0] > snrs[1] > snrs[3]        # pruning costs SNR


def test_learned_controller_beats_plain_monitor_loop(mlp):
    """Ablation: the same inputs, the same target SNR; the learned loop must
    need fewer evaluations and monitor calls, reach an equal or better
    measured output SNR, and run faster."""
    from cipherpu import gear, nn
    F, T = gear.sweep(mlp, 384, seed=1)
    assert F.shape == (384, 16) and T.shape == (384, 4)
    net = gear.GearNet(seed=0)
    hist = net.fit(F, T, epochs=60, batch=128, lr=2e-3)
    assert hist[-1] < hist[0]
    pred = net.predict(F).cpu().numpy()
    err = np.abs(pred[:, 1:] - T[:, 1:]).mean()
    print("predictor mean abs error (dB) on gears 1-3:", round(float(err), 3))
    X = nn.synthetic_images(1024, seed=99)
    target = 3.0
    ref = nn.plain_forward(mlp.W1, mlp.W2, X)
    plain = gear.plain_monitor_loop(mlp, X, target, chunk_cts=64)
    learned = gear.learned_loop(mlp, net, X, target, chunk_cts=64)
    s_plain, s_learned = gear.snr_db(ref, plain.outputs), gear.snr_db(ref, learned.outputs)
    print("plain  ", plain.gears, plain.evaluations, plain.monitor_calls, round(plain.samples_per_s, 1), s_plain)
    print("learned", learned.gears, learned.evaluations, learned.monitor_calls, round(learned.samples_per_s, 1),
          s_learned)
    assert s_plain >= target and s_learned >= target
    assert s_learned >= s_plain - 0.1
    assert learned.evaluations <= plain.evaluations and learned.monitor_calls < plain.monitor_calls
    assert learned.samples_per_s > plain.samples_per_s


@pytest.mark.parametrize("kind", ["mlp", "cnn"])
def test_controller_beats_constant_predictor_on_heldout_networks(torch_gpu, kind):
    """The committed controller on networks it never saw (seeds 2000..2004,
    one per weight family): its mean absolute error on the monitor's
    per-gear SNR (gears 1-3, where networks differ) is below that of the
    constant predictor answering each gear's mean training label (the
    controller's stored label mean), by at least a third."""
    from cipherpu import gear
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu.gear_train import sweep_networks
    import time
    ck = CKKS(CKKSParams(), seed=8)
    sk = ck.keygen()
    ctl = gear.GearNet(seed=0)
    assert gear.load_trained(ctl, gear.TRAINED_PATHS[kind])
    F, T = sweep_networks(kind, ck, sk, range(2000, 2005), 64, 64, "held-out", time.perf_counter())
    pred = ctl.predict(F).cpu().numpy()
    const = ctl.y_mu.cpu().numpy()[None, :]
    mae = np.abs(pred - T)[:, 1:].mean()
    mae_c = np.abs(const - T)[:, 1:].mean()
    print(kind, "held-out MAE dB, controller vs constant:", round(float(mae), 2), round(float(mae_c), 2))
    assert mae < 0.67 * mae_c
