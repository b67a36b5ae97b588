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


@pytest.fixture(params=[True, False], ids=["fp64", "int64"])
def arith(request, mlp_env):
    """Runs a test under each arithmetic of the shared context, restoring FP64."""
    ring = mlp_env[0].ring
    ring.set_arith(request.param)
    yield request.param
    ring.set_arith(True)


def test_lincomb_scalar_bit_exact(mlp_env, arith):
    """out[g] = sum_b D[g][b] E[b] mod q_i with one integer per (g, b, limb),
    against Python integers."""
    import torch
    from cipherpu._lib import call
    from cipherpu.ring import to_numpy
    import ctypes
    ck = mlp_env[0]
    rng = np.random.default_rng(9)
    nb, ng, B, l = 3, 2, 2, 2
    E = np.stack([np.stack([np.stack([rng.integers(0, ck.primes[i], ck.N, dtype=np.uint64) for i in range(l)])
                            for _ in range(2)]) for _ in range(nb * B)]).reshape(nb, B, 2, l, ck.N)
    W = rng.integers(-2 ** 44, 2 ** 44, (ng, nb))
    D = np.zeros((ng, nb, l, 2), dtype=np.uint64)
    for g in range(ng):
        for b in range(nb):
            for i in range(l):
                q = ck.primes[i]
                r = int(W[g, b]) % q
                D[g, b, i] = ck.ring.prepared_pairs(r, q)
    Et, Dt = ck.ring.tensor(E), ck.ring.tensor(D)
    out = torch.empty(ng, B, 2, l, ck.N, dtype=torch.uint64, device=ck.device)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    call("ckks_lincomb_scalar", ck.ring.handle, p(Et), p(Dt), p(out), nb, ng, B, l, ck.ring.stream())
    got = to_numpy(out)
    for g in range(ng):
        for i in range(l):
            q = ck.primes[i]
            want = sum(E[b, :, :, i].astype(object) * (int(W[g, b]) % q) for b in range(nb)) % q
            assert np.array_equal(got[g, :, :, i].astype(object), want), (g, i)


@pytest.mark.parametrize("nb,ng,B,l,scalar", [(16, 8, 2, 4, 0), (17, 10, 3, 3, 0), (32, 5, 3, 1, 0),
                                              (4, 20, 1, 2, 0), (24, 8, 3, 2, 1), (5, 17, 2, 1, 1)])
def test_lincomb_shapes_bit_exact(mlp_env, arith, nb, ng, B, l, scalar):
    """out[g][row] = sum_b D[g][b] E[b][row] mod q_i, per coefficient, for
    giant-step counts across the kernel's blocks (<= 4, <= 8, <= 16 and
    chunks of 16), row counts that leave a partial row block, and up to 32
    baby steps; against Python integers on the first 64 coefficients (the
    combination is coefficient-wise)."""
    import ctypes
    import torch
    from cipherpu._lib import call
    from cipherpu.ring import to_numpy
    ck = mlp_env[0]
    rng = np.random.default_rng(nb * 100 + ng)
    N, K = ck.N, 64
    E = np.stack([rng.integers(0, ck.primes[i], (nb, B, 2, N), dtype=np.uint64) for i in range(l)], axis=3)
    dshape = (ng, nb, l) if scalar else (ng, nb, l, N)
    W = np.stack([rng.integers(0, ck.primes[i], dshape[:2] + dshape[3:], dtype=np.uint64) for i in range(l)], axis=2)
    q = np.array([ck.primes[i] for i in range(l)], dtype=object)
    qb = q.reshape((1, 1, l) + ((1,) if not scalar else ()))
    D = ck.ring.prepared_pairs(W, qb.astype(np.uint64))
    Et, Dt = ck.ring.tensor(E), ck.ring.tensor(D)
    out = torch.empty(ng, B, 2, l, N, dtype=torch.uint64, device=ck.device)
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    call("ckks_lincomb_scalar" if scalar else "ckks_lincomb", ck.ring.handle, p(Et), p(Dt), p(out), nb, ng, B, l,
         ck.ring.stream())
    got = to_numpy(out)[..., :K].astype(object)
    Eo = E[..., :K].astype(object)
    for g in range(ng):
        for i in range(l):
            w = [W[g, b, i] if scalar else W[g, b, i, :K] for b in range(nb)]
            want = sum(Eo[b, :, :, i] * (int(w[b]) if scalar else w[b].astype(object)) for b in range(nb))
            assert np.array_equal(got[g, :, :, i], want % int(q[i])), (g, i)


def test_cnn_matches_plaintext(mlp_env):
    from cipherpu import cnn, nn
    ck, sk = mlp_env[0], mlp_env[1]
    K, F = cnn.make_cnn(0)
    net = cnn.EncryptedCNN(ck, sk, K, F)
    X = nn.synthetic_images(8, seed=5)
    out = net.decrypt_outputs(net.forward(net.encrypt_inputs(X)), 8)
    ref = cnn.cnn_forward(K, F, X)
    assert np.max(np.abs(out - ref)) < 1e-3 * np.max(np.abs(ref))
    assert nn.output_snr_db(ref, out) > 60
    net.set_gears(2, 2)
    PK, PF = net.pruned_weights()
    out2 = net.decrypt_outputs(net.forward(net.encrypt_inputs(X)), 8)
    ref2 = cnn.cnn_forward(PK, PF, X)
    assert np.max(np.abs(out2 - ref2)) < 1e-3 * np.max(np.abs(ref2))
    assert np.mean(PF == 0) > 0.45


def test_networks_from_their_lowest_levels(mlp_env):
    """forward with drop_levels (each stage from the lowest level it needs:
    the input's upper limbs dropped, the DARE refresh at the second stage's
    level) and without (every stage from the top level) both match the FP64
    networks, and the outputs end at level 1 with the drop.  A layer plan's
    rows for a lower level equal the rows prepared at that level."""
    import torch
    from cipherpu import cnn, nn
    ck, sk, W1, W2 = mlp_env
    X = nn.synthetic_images(8, seed=21)
    K, F = cnn.make_cnn(0)
    for net in (nn.EncryptedMLP(ck, sk, W1, W2), cnn.EncryptedCNN(ck, sk, K, F)):
        ref = net.reference_outputs(X)
        for drop in (True, False):
            net.drop_levels = drop
            ct = net.encrypt_inputs(X)
            assert ct.level == ck.L
            y = net.forward(ct)
            ys = y if isinstance(y, list) else [y]
            assert all(o.level == (1 if drop else ck.L - 2) for o in ys), (type(net).__name__, drop)
            out = net.decrypt_outputs(y, 8)
            assert np.max(np.abs(out - ref)) < 1e-3 * np.max(np.abs(ref)), (type(net).__name__, drop)
            assert nn.output_snr_db(ref, out) > 60
    mlp = nn.EncryptedMLP(ck, sk, W1, W2)
    direct = mlp._plan(nn.extended_diagonals(mlp.P1, 128, 896), 128, 896, 16, 8, (512, 256, 128), 3)
    assert torch.equal(mlp.p1.rows_at(3), direct.D)
    assert mlp.p1.rows_at(ck.L) is mlp.p1.D


def test_device_samplers(mlp_env):
    """Uniform residues in range with the right mean; rounded-Gaussian errors
    the same integer in every limb, with the table's distribution; both
    reproducible from the seed; a device-sampled encryption decrypts."""
    import ctypes
    import torch
    from cipherpu._lib import call, load
    from cipherpu.ring import to_numpy
    ck, sk = mlp_env[0], mlp_env[1]
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    rows, l = 64, 3
    u = torch.empty(rows, l, ck.N, dtype=torch.uint64, device=ck.device)
    call("ckks_sample_uniform", ck.ring.handle, p(u), rows, l, 0, 12345, ck.ring.stream())
    U = to_numpy(u)
    for i in range(l):
        q = ck.primes[i]
        assert U[:, i].max() < q
        assert abs(U[:, i].astype(np.float64).mean() / q - 0.5) < 0.005
    u2 = torch.empty_like(u)
    call("ckks_sample_uniform", ck.ring.handle, p(u2), rows, l, 0, 12345, ck.ring.stream())
    assert torch.equal(u, u2)
    g = torch.empty(rows, l, ck.N, dtype=torch.uint64, device=ck.device)
    call("ckks_sample_gaussian", ck.ring.handle, p(g), rows, l, 0, ctypes.c_double(3.2), 777, ck.ring.stream())
    G = to_numpy(g).astype(np.int64)
    e = np.where(G[:, 0] > ck.primes[0] // 2, G[:, 0] - ck.primes[0], G[:, 0])
    for i in range(1, l):
        q = ck.primes[i]
        assert np.array_equal(np.where(G[:, i] > q // 2, G[:, i] - q, G[:, i]), e)
    assert abs(e.mean()) < 0.05 and abs(e.std() - 3.2) < 0.05
    cdf = (ctypes.c_uint64 * 48)()
    load().ckks_gaussian_cdf(ctypes.c_double(3.2), cdf, 48)
    probs = np.diff(np.concatenate([[0.0], np.array([float(c) for c in cdf]) / 2.0 ** 64]))
    for k in range(6):
        assert abs(np.mean(np.abs(e) == k) - probs[k]) < 0.01, k
    z = np.random.default_rng(3).uniform(-1, 1, (2, ck.N // 2))
    ct = ck.encrypt(ck.encode(z), sk)
    assert ck.sampler == "device"
    dec = to_numpy(ck.decode(ck.decrypt(ct, sk), ct.scale)).real
    assert np.max(np.abs(dec - z)) < 1e-7


def test_mixed_step_rotations_keep_every_word(mlp_env):
    """The MLP's giant steps and the CNN's output shifts rotate rows by
    different steps in one ckks_rotate_many call; both linear layers of the
    MLP and the CNN's output stage give the same words as with one rotation
    call per step (the stages are compared on the same inputs, since the
    DARE refresh between them draws fresh encryption randomness)."""
    import torch
    from cipherpu import cnn, nn
    ck, sk, W1, W2 = mlp_env
    X = nn.synthetic_images(8, seed=9)
    mlp = nn.EncryptedMLP(ck, sk, W1, W2)
    ct = mlp.encrypt_inputs(X)
    h = mlp.activation(mlp.dare(mlp.layer1(ct)))
    K, F = cnn.make_cnn(0)
    net = cnn.EncryptedCNN(ck, sk, K, F)
    y = net.activation(net.dare(net.conv(net.encrypt_inputs(X))))
    stages = [(mlp, lambda: [mlp.layer1(ct)]), (mlp, lambda: [mlp.layer2(h)]), (net, lambda: net.fc(y, ct.batch))]
    for model, stage in stages:
        got = []
        for mixed in (True, False):
            model.mixed_rotations = mixed
            model.ops["rot"] = 0
            got.append(([o.data.clone() for o in stage()], model.ops["rot"]))
        model.mixed_rotations = True
        assert got[0][1] == got[1][1] > 0
        assert all(torch.equal(a, b) for a, b in zip(got[0][0], got[1][0])), type(model).__name__
