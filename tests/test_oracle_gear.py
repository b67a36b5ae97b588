"""CPU checks of the gear layer: the gear-network oracle (gradients by
finite differences, BF16 rounding against torch's conversion), the
block-pruning rule of the encrypted MLP against the oracle's restatement of
the reference rule, and the slot plans of the encrypted MLP and CNN
simulated in plaintext (rotations as cyclic shifts)."""
import math

import numpy as np
import pytest
import torch

import gear_oracle as go
from ckks_oracle import block_prune_mask


def _params(rng, dims=((16, 64), (64, 64), (64, 4))):
    return [(rng.normal(0, math.sqrt(2 / i), (o, i)).astype(np.float32), rng.normal(0, 0.1, o).astype(np.float32))
            for i, o in dims]


def test_bf16_round_matches_torch_and_known_values():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.normal(0, 10, 4096), [0.0, -0.0, 1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7,
                                                  -(1 + 2 ** -8), 65504.0, 1e-30]]).astype(np.float32)
    ours = go.bf16_round(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    # ties go to even: 1 + 2^-8 sits halfway between 1 and 1 + 2^-7
    assert go.bf16_round(np.float32(1 + 2 ** -8)) == np.float32(1.0)
    assert go.bf16_round(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    params = [(W.astype(np.float64), b.astype(np.float64)) for W, b in _params(rng)]
    X = rng.normal(0, 1, (32, 16))
    T = rng.normal(0, 1, (32, 4))

    def loss(ps):
        Y, _ = go.forward(ps, X, bf16_hidden=False)
        return go.mse(Y, T)

    Y, acts = go.forward(params, X, bf16_hidden=False)
    grads = go.backward(params, acts, Y, T)
    h = 1e-3
    for k in range(3):
        for which in (0, 1):
            arr = params[k][which]
            for idx in [tuple(rng.integers(0, s) for s in arr.shape) for _ in range(6)]:
                old = arr[idx]
                arr[idx] = old + h
                lp = loss(params)
                arr[idx] = old - h
                lm = loss(params)
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                an = grads[k][which][idx]
                assert abs(fd - an) <= 2e-3 * max(1.0, abs(fd)), (k, which, idx, fd, an)


def test_backward_uses_stored_bf16_activations():
    """With BF16 storage the weight gradient of layer k+1 is d^T @ (stored
    BF16 activation): a transposed or unrounded operand fails this."""
    rng = np.random.default_rng(2)
    params = _params(rng)
    X = rng.normal(0, 1, (8, 16)).astype(np.float32)
    T = rng.normal(0, 1, (8, 4)).astype(np.float32)
    Y, acts = go.forward(params, X, bf16_hidden=True)
    assert np.array_equal(acts[1], go.bf16_round(acts[1]))
    g = go.backward(params, acts, Y, T)
    d = 2 * (Y - T) / 8
    assert np.allclose(g[2][0], d.T @ acts[2], rtol=1e-5)
    assert np.allclose(g[2][1], d.sum(0), rtol=1e-5)
    assert g[0][0].shape == (64, 16)


@pytest.mark.parametrize("frac", [0.0, 0.25, 0.5, 0.75])
@pytest.mark.parametrize("shape", [(128, 784), (10, 128), (7, 30)])
def test_block_rule_matches_oracle(frac, shape):
    from cipherpu.nn import block_keep_mask
    rng = np.random.default_rng(3)
    W = rng.normal(0, 1, shape)
    assert np.array_equal(block_keep_mask(W, frac), block_prune_mask(W, frac))


def test_block_rule_closed_form():
    """Blocks with max |w| equal to 1..8 -> dropping 50% drops the blocks of
    max 1..4 exactly."""
    w = np.zeros(8 * 64)
    for b, m in enumerate([5, 1, 7, 2, 8, 3, 6, 4]):
        w[b * 64 + 10] = m
    from cipherpu.nn import block_keep_mask
    keep = block_keep_mask(w.reshape(8, 64), 0.5)
    assert keep.reshape(8, 64)[:, 0].tolist() == [True, False, True, False, True, False, True, False]


def _simulate_linear(x_slots, D, rows, cols, baby, giant, folds):
    """Plaintext restatement of the encrypted diagonal plan: rot(v, k)
    moves slot i + k to slot i inside the ciphertext (np.roll by -k)."""
    from cipherpu.nn import SLOTS_PER_SAMPLE
    S = x_slots.size // SLOTS_PER_SAMPLE
    rot = lambda v, k: np.roll(v, -k)
    z = np.zeros_like(x_slots)
    for g in range(giant):
        inner = np.zeros_like(x_slots)
        for b in range(baby):
            k = g * baby + b
            blk = np.zeros(SLOTS_PER_SAMPLE)
            if k < rows:
                blk[:cols] = D[k]
            d = np.roll(np.tile(blk, S), baby * g)
            inner += d * rot(x_slots, b)
        z += rot(inner, baby * g)
    for s in folds:
        z = z + rot(z, s)
    return z


def test_diagonal_plan_in_plaintext():
    from cipherpu import nn
    rng = np.random.default_rng(4)
    W1, W2 = nn.make_mlp(5)
    X = nn.synthetic_images(4, seed=6)
    v = np.zeros(4096)
    for j in range(4):
        blk = np.zeros(1024)
        blk[:784] = X[j]
        blk[896:] = blk[:128]
        v[j * 1024:(j + 1) * 1024] = blk
    z = _simulate_linear(v, nn.extended_diagonals(W1, 128, 896), 128, 896, 16, 8, (512, 256, 128))
    y1 = X @ W1.T
    for j in range(4):
        assert np.allclose(z[j * 1024:j * 1024 + 128], y1[j], atol=1e-12)
    # layer 2 on the masked, replicated hidden layer
    h = np.zeros(4096)
    for j in range(4):
        h[j * 1024:j * 1024 + 128] = y1[j] ** 2
        h[j * 1024 + 128:j * 1024 + 256] = y1[j] ** 2
    z2 = _simulate_linear(h, nn.extended_diagonals(W2, 16, 128), 16, 128, 4, 4, (64, 32, 16))
    ref = nn.plain_forward(W1, W2, X)
    for j in range(4):
        assert np.allclose(z2[j * 1024:j * 1024 + 10], ref[j], atol=1e-12)
    # pruned weights survive the diagonal round trip
    P1 = W1 * nn.block_keep_mask(W1, 0.5)
    assert np.array_equal(nn.weights_from_diagonals(nn.extended_diagonals(P1, 128, 896), 128, 896, P1.shape), P1)


def test_cnn_packing_in_plaintext():
    """The encrypted CNN's slot plan, run on plaintext slot vectors with
    rotations as cyclic shifts, equals the FP64 network: the pool folded
    into the FC weights (EncryptedCNN's fc_weight_blocks, the weight on all
    four slots of each 2x2 window), and the pool by (1 + rot_1)(1 + rot_16)
    with the weight on the window's first slot."""
    from cipherpu import cnn, nn
    K, F = cnn.make_cnn(1)
    X = nn.synthetic_images(4, seed=2)
    v = cnn.pack_images(X, 4)[0]
    rot = lambda z, k: np.roll(z, -k)
    acts = []
    for c in range(cnn.N_CH):
        y = sum(K[c, u, w] * rot(v, cnn.tap_step(u, w)) for u in range(5) for w in range(5))
        acts.append(y ** 2)
    pooled = [(a + rot(a, 1)) + rot(a + rot(a, 1), cnn.ROW) for a in acts]
    first = np.zeros((cnn.N_OUT * cnn.N_CH, 1024))
    for k in range(cnn.N_OUT):
        for c in range(cnn.N_CH):
            for i in range(6):
                for j in range(6):
                    first[cnn.N_CH * k + c, 32 * i + 2 * j] = F[k, 36 * c + 6 * i + j] / 4
    folded = cnn.fc_weight_blocks(F)
    assert np.count_nonzero(folded) == 4 * np.count_nonzero(first)
    for chans, W in ((acts, folded), (pooled, first)):
        outs = [sum(np.tile(W[cnn.N_CH * k + c], 4) * chans[c] for c in range(cnn.N_CH)) for k in range(cnn.N_OUT)]
        got = np.zeros((4, 10))
        for p in range(0, 10, 4):
            z = sum(rot(outs[p + m], -256 * m) for m in range(min(4, 10 - p)))
            for k in range(1, 9):
                z = z + rot(z, 256 >> k)
            for s in range(4):
                for m in range(min(4, 10 - p)):
                    got[s, p + m] = z[1024 * s + 256 * m]
        assert np.allclose(got, cnn.cnn_forward(K, F, X), atol=1e-10)


def test_cnn_forward_against_direct_convolution():
    """cnn_forward's strided slicing against a direct loop over outputs."""
    from cipherpu import cnn, nn
    K, F = cnn.make_cnn(3)
    X = nn.synthetic_images(2, seed=4)
    img = X.reshape(2, 28, 28)
    y = np.zeros((2, 8, 12, 12))
    for n in range(2):
        for c in range(8):
            for i in range(12):
                for j in range(12):
                    y[n, c, i, j] = np.sum(K[c] * img[n, 2 * i:2 * i + 5, 2 * j:2 * j + 5])
    a = y ** 2
    p = a.reshape(2, 8, 6, 2, 6, 2).mean(axis=(3, 5))
    assert np.allclose(cnn.cnn_forward(K, F, X), p.reshape(2, 288) @ F.T, atol=1e-12)


def test_weight_families_and_pruning_statistics():
    """The controller's training networks: every weight family has variance
    1 / fan_in; a gear's weight SNR is 10 log10(|W|^2 / |W - W_g|^2) of the
    block rule (checked against the oracle's mask), never rises with the
    gear, and reads the cap when the pruned blocks are exactly zero (the
    block-sparse family at gear 1)."""
    from cipherpu import cnn, nn
    for seed in range(len(nn.WEIGHT_FAMILIES)):
        W1, W2 = nn.make_mlp_family(seed)
        assert W1.shape == (128, 784) and W2.shape == (10, 128)
        assert abs(W1.var() * 784 - 1) < 1e-9 and abs(W2.var() * 128 - 1) < 1e-9
        K, F = cnn.make_cnn_family(seed)
        assert K.shape == (8, 5, 5) and F.shape == (10, 288)
        snrs = [nn.weight_snr_db(W1, g) for g in range(4)]
        assert snrs[0] == 60.0 and snrs[1] >= snrs[2] >= snrs[3]
        mask = block_prune_mask(W1, nn.GEAR_FRACTIONS[2], 64)
        want = 10 * math.log10(np.sum(W1 ** 2) / np.sum((W1 * ~mask) ** 2))
        assert abs(nn.weight_snr_db(W1, 2) - min(60.0, want)) < 1e-9
    W1, _ = nn.make_mlp_family(4)                       # block-sparse family
    zero_blocks = np.mean(np.abs(W1.reshape(-1, 64)).max(axis=1) == 0)
    assert zero_blocks > 0.4 and nn.weight_snr_db(W1, 1) == 60.0
    # families differ where the controller needs them to: gear-1 pruning
    # costs a Gaussian network more than a heavy-tailed one
    assert nn.weight_snr_db(nn.make_mlp_family(0)[0], 1) < nn.weight_snr_db(nn.make_mlp_family(2)[0], 1)
