"""Encrypted MLP 784-128-10 on the GPU against the FP64 plaintext network
(decrypted outputs within a stated error; output SNR above a bound)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mlp_env():
    import torch
    assert torch.cuda.is_available()
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu import nn
    ck = CKKS(CKKSParams(), seed=3)
    sk = ck.keygen()
    W1, W2 = nn.make_mlp(0)
    return ck, sk, W1, W2


def test_layers_match_plaintext(mlp_env):
    from cipherpu import nn
    ck, sk, W1, W2 = mlp_env
    X = nn.synthetic_images(8, seed=1)
    m = nn.EncryptedMLP(ck, sk, W1, W2)
    ct = m.encrypt_inputs(X)
    h = m.layer1(ct)
    z = ck.decode(ck.decrypt(h, sk), h.scale).cpu().numpy().real
    y1 = X @ W1.T
    for s in range(8):
        c, j = divmod(s, 4)
        blk = z[c, j * 1024:(j + 1) * 1024]
        assert np.max(np.abs(blk[:128] - y1[s])) < 1e-4   # ~26 key switches at scale >= 2^35
        assert np.max(np.abs(blk[128:256] - y1[s])) < 1e-4      # replicated copy
        assert np.max(np.abs(blk[256:])) < 1e-4                  # masked
    out = m.decrypt_outputs(m.layer2(m.activation(m.dare(h))), 8)
    ref = nn.plain_forward(W1, W2, X)
    assert np.max(np.abs(out - ref)) < 1e-3 * np.max(np.abs(ref))
    assert nn.output_snr_db(ref, out) > 60


@pytest.mark.parametrize("gear", [1, 3])
def test_gears_prune_and_match_pruned_network(mlp_env, gear):
    from cipherpu import nn
    ck, sk, W1, W2 = mlp_env
    X = nn.synthetic_images(4, seed=2)
    m = nn.EncryptedMLP(ck, sk, W1, W2, gear1=gear, gear2=gear)
    P1, P2 = m.pruned_weights()
    frac = nn.GEAR_FRACTIONS[gear]
    assert abs(np.mean(P1 == 0) - frac) < 0.05 + np.mean(W1 == 0)
    out = m.decrypt_outputs(m.forward(m.encrypt_inputs(X)), 4)
    ref_pruned = nn.plain_forward(P1, P2, X)
    assert np.max(np.abs(out - ref_pruned)) < 1e-3 * np.max(np.abs(ref_pruned))
    # pruning costs accuracy against the dense network, more for higher gears
    dense = nn.plain_forward(W1, W2, X)
    assert nn.output_snr_db(dense, out) < 60
