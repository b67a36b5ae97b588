"""Gear selection for the encrypted network: the plain decrypt-based noise
monitor (reference semantics) and the learned gear controller.

Plain monitor (the reference's policy and the arbiter of predictions):
  * noise energy E = sum min(c, q - c)^2 / n, r_hat = sqrt(E) / q
    (verify_m4_h_multilayer.py:162-174; rtl/m4/noise_monitor.sv:44-78),
    computed on the GPU by ckks_noise_stats;
  * output SNR = 10 log10(sum s^2 / sum (s - s_hat)^2) against the FP64
    plaintext network (verify_m4h2_prod_calibration.py:64-69), after
    decrypting -- the "decrypt-based" monitor;
  * gear loop: try the cheapest gear first and step to safer gears until the
    decrypt-based SNR meets the target.

Learned controller: a 16 -> 64 -> 64 -> 4 network (FP32 weights, BF16
activations, forward/backward/Adam in HIP, include/ckks_core.h
ckks_gear_*) maps the network's per-gear pruning statistics, ciphertext
statistics, level and gear history to the predicted output SNR of each
gear.  It is trained across networks of several weight families, so the
pruning statistics are what tell one network's gears from another's (the
ciphertext coefficients are uniform modulo q and carry nothing about the
message).  It picks the cheapest gear whose
prediction clears the target by the margin (3 dB); when no gear does, the
plain monitor decides.  The predictor never touches the modular
arithmetic, and the plain monitor measures every reported SNR.
"""
from __future__ import annotations

import contextlib
import ctypes
import gc
import math
import os
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from ._lib import call

N_FEATURES = 16
HIDDEN = 64
N_GEARS = 4
MARGIN_DB = 3.0


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream(dev):
    return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


# ---------------------------------------------------------------- features

PROBE_IMAGES = 64     # synthetic images of the plaintext probe (fixed seed)


def network_features(net) -> np.ndarray:
    """The network's pruning statistics, 9 values, computed once per network
    in plaintext from its weights: the weight SNR in dB (nn.weight_snr_db,
    capped at 60) of each of its two weight matrices under gears 1, 2 and 3
    (6 values), then the output SNR of the pruned FP64 network against the
    dense one at gears 1, 2 and 3 on a fixed probe batch of synthetic
    images (3 values; capped at 60), all divided by 60.  No ciphertext and
    no input of the current batch enters them."""
    from .nn import synthetic_images, weight_snr_db
    cached = getattr(net, "_gear_features", None)
    if cached is None:
        mats = net.weight_matrices()
        w = [weight_snr_db(W, g) for W in mats for g in (1, 2, 3)]
        Xp = synthetic_images(PROBE_IMAGES, seed=777)
        ref = net.reference_outputs(Xp)
        o = [min(60.0, snr_db(ref, net.plain_outputs_at_gear(Xp, g))) for g in (1, 2, 3)]
        cached = np.array(w + o) / 60.0
        net._gear_features = cached
    return cached


def ciphertext_features(ck, ct, prev_gear: int, layer: int = 0, netf=None) -> torch.Tensor:
    """[B, 16] float32 features on the device: rms(c')/q of c0 and c1 on
    limb 0 (ckks_noise_stats moments; 2 values), the network's pruning
    statistics (network_features; 9), mean(c')/q of c0 on limb 0 (1), then
    level / L, log2(scale) / 64, the previous gear / 3 and the layer index."""
    return features_many(ck, [ct], prev_gear, layer, netf)


def features_many(ck, cts, prev_gear: int, layer: int = 0, netf=None) -> torch.Tensor:
    """ciphertext_features of several ciphertexts of one level and scale,
    rows concatenated, computed in one pass."""
    l, scale = cts[0].level, cts[0].scale
    if any(ct.level != l or ct.scale != scale for ct in cts):
        return torch.cat([features_many(ck, [ct], prev_gear, layer, netf) for ct in cts])
    B = sum(ct.batch for ct in cts)
    m = ck.ring.noise_moments_many([ct.data for ct in cts]).view(B, 2, l, 4)
    f = torch.zeros(B, N_FEATURES, dtype=torch.float64, device=m.device)
    f[:, 0] = m[:, 0, 0, 1]
    f[:, 1] = m[:, 1, 0, 1]
    if netf is not None:
        f[:, 2:11] = torch.as_tensor(np.asarray(netf, dtype=np.float64), device=m.device)
    f[:, 11] = m[:, 0, 0, 0]
    f[:, 12] = l / ck.L
    f[:, 13] = math.log2(scale) / 64.0
    f[:, 14] = prev_gear / 3.0
    f[:, 15] = float(layer)
    return f.float()


# ---------------------------------------------------------------- the network

class GearNet:
    """16 -> 64 -> 64 -> 4 dense ReLU network on the GPU."""

    def __init__(self, device=None, seed: int = 0, bf16: bool = True):
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        self.bf16 = bf16
        g = torch.Generator().manual_seed(seed)
        dims = [(N_FEATURES, HIDDEN), (HIDDEN, HIDDEN), (HIDDEN, N_GEARS)]
        self.W, self.b = [], []
        for i, o in dims:
            self.W.append((torch.randn(o, i, generator=g) * math.sqrt(2.0 / i)).float().to(self.dev).contiguous())
            self.b.append(torch.zeros(o, dtype=torch.float32, device=self.dev))
        self.m = [torch.zeros_like(p) for p in self.W + self.b]
        self.v = [torch.zeros_like(p) for p in self.W + self.b]
        self.t = 0
        self.mu = torch.zeros(N_FEATURES)
        self.sd = torch.ones(N_FEATURES)
        self.y_mu, self.y_sd = 0.0, 1.0

    def params(self):
        return [(W, b) for W, b in zip(self.W, self.b)]

    def _act_dtype(self):
        return torch.bfloat16 if self.bf16 else torch.float32

    def forward(self, X: torch.Tensor):
        """X [B, 16] float32 on device -> (Y [B, 4] float32, stored acts)."""
        acts = [X.contiguous()]
        h, h_bf = acts[0], 0
        s = _stream(self.dev)
        n = len(self.W)
        for k, (W, b) in enumerate(zip(self.W, self.b)):
            J, I = W.shape
            last = k == n - 1
            out = torch.empty(X.shape[0], J, dtype=torch.float32 if last else self._act_dtype(), device=self.dev)
            out_bf = 0 if last else int(self.bf16)
            call("ckks_gear_dense_fwd", _ptr(h), h_bf, _ptr(W), _ptr(b), _ptr(out), out_bf, X.shape[0], I, J,
                 0 if last else 1, s)
            if not last:
                acts.append(out)
            h, h_bf = out, out_bf
        return h, acts

    def backward(self, acts, dY: torch.Tensor):
        """Gradients (dW, db) of every layer for the loss gradient dY."""
        s = _stream(self.dev)
        grads = [None] * len(self.W)
        d = dY.contiguous()
        B = dY.shape[0]
        for k in range(len(self.W) - 1, -1, -1):
            W = self.W[k]
            J, I = W.shape
            out_act = acts[k + 1] if k < len(self.W) - 1 else d
            out_bf = int(self.bf16) if k < len(self.W) - 1 else 0
            in_act = acts[k]
            in_bf = 0 if k == 0 else int(self.bf16)
            dpre = torch.empty(B, J, dtype=torch.float32, device=self.dev)
            dW = torch.empty_like(W)
            db = torch.empty(J, dtype=torch.float32, device=self.dev)
            din = torch.empty(B, I, dtype=torch.float32, device=self.dev) if k > 0 else None
            call("ckks_gear_dense_bwd", _ptr(d), _ptr(out_act), out_bf, _ptr(in_act), in_bf, _ptr(W), _ptr(dpre),
                 _ptr(dW), _ptr(db), _ptr(din), B, I, J, 1 if k < len(self.W) - 1 else 0, s)
            grads[k] = (dW, db)
            d = din
        return grads

    def loss_grad(self, Y: torch.Tensor, T: torch.Tensor):
        B, J = Y.shape
        dY = torch.empty_like(Y)
        rl = torch.empty(B, dtype=torch.float32, device=self.dev)
        call("ckks_gear_mse_grad", _ptr(Y), _ptr(T.contiguous()), _ptr(dY), _ptr(rl), B, J, _stream(self.dev))
        return dY, rl

    def adam(self, grads, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        ps = self.W + self.b
        gs = [g[0] for g in grads] + [g[1] for g in grads]
        for p, g, m, v in zip(ps, gs, self.m, self.v):
            call("ckks_gear_adam", _ptr(p), _ptr(g), _ptr(m), _ptr(v), p.numel(), ctypes.c_float(lr),
                 ctypes.c_float(b1), ctypes.c_float(b2), ctypes.c_float(eps), self.t, _stream(self.dev))

    # -- training / inference on standardised data ------------------------------
    def _dev(self, a) -> torch.Tensor:
        return torch.as_tensor(a, dtype=torch.float32).to(self.dev)

    def fit(self, X, T, epochs: int = 30, batch: int = 1024, lr: float = 2e-3, seed: int = 0) -> List[float]:
        """Mean-squared-error training of the SNR predictor on the GPU with
        the forward, backward and Adam kernels; X [n, 16] features, T [n, 4]
        per-gear SNR labels (dB).  Returns the mean loss of every epoch (in
        standardised units)."""
        Xd, Td = self._dev(X), self._dev(T)
        self.mu, self.sd = Xd.mean(0), Xd.std(0) + 1e-6
        self.y_mu, self.y_sd = Td.mean(0), Td.std(0) + 1e-6
        Xd = ((Xd - self.mu) / self.sd).contiguous()
        Td = ((Td - self.y_mu) / self.y_sd).contiguous()
        g = torch.Generator(device="cpu").manual_seed(seed)
        n = Xd.shape[0]
        hist = []
        for _ in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.dev)
            tot = torch.zeros((), dtype=torch.float32, device=self.dev)
            for s in range(0, n, batch):
                idx = perm[s:s + batch]
                Y, acts = self.forward(Xd[idx])
                dY, rl = self.loss_grad(Y, Td[idx])
                self.adam(self.backward(acts, dY), lr=lr)
                tot += rl.sum()
            hist.append(float(tot) / n)
        return hist

    def predict(self, X) -> torch.Tensor:
        """Predicted per-gear SNR (dB), [n, 4] float32 on the device."""
        Xd = ((self._dev(X) - torch.as_tensor(self.mu, device=self.dev)) /
              torch.as_tensor(self.sd, device=self.dev)).contiguous()
        Y, _ = self.forward(Xd)
        return Y * torch.as_tensor(self.y_sd, device=self.dev) + torch.as_tensor(self.y_mu, device=self.dev)

    def state(self) -> dict:
        d = {f"W{k}": W.cpu().numpy() for k, W in enumerate(self.W)}
        d.update({f"b{k}": b.cpu().numpy() for k, b in enumerate(self.b)})
        for name in ("mu", "sd", "y_mu", "y_sd"):
            d[name] = torch.as_tensor(getattr(self, name)).cpu().numpy()
        return d

    def load_state(self, d: dict) -> None:
        for k in range(len(self.W)):
            self.W[k] = torch.as_tensor(d[f"W{k}"], dtype=torch.float32).to(self.dev).contiguous()
            self.b[k] = torch.as_tensor(d[f"b{k}"], dtype=torch.float32).to(self.dev).contiguous()
        for name in ("mu", "sd", "y_mu", "y_sd"):
            setattr(self, name, torch.as_tensor(d[name], dtype=torch.float32).to(self.dev))

    def save(self, path: str) -> None:
        np.savez(path, **self.state())

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as z:
            self.load_state({k: z[k] for k in z.files})


# Controller weights trained by cipherpu.gear_train on monitor sweeps of 40
# networks of each architecture (784-128-10 MLPs of nn.make_mlp_family and
# CNNs of cnn.make_cnn_family, synthetic images, every gear).
_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
TRAINED_PATH = os.path.join(_DATA, "gearnet_mlp784.npz")
TRAINED_PATHS = {"mlp": TRAINED_PATH, "cnn": os.path.join(_DATA, "gearnet_cnn.npz")}


def load_trained(net: GearNet, path: str = TRAINED_PATH) -> Optional[str]:
    """Load committed controller weights into net; the file name, or None
    when there are none."""
    if not os.path.exists(path):
        return None
    net.load(path)
    return os.path.basename(path)


def choose_gear(pred_snr, target_db: float, margin_db: float = MARGIN_DB) -> Optional[int]:
    """Cheapest (highest) gear whose predicted SNR clears target + margin;
    None when no gear does (the plain monitor then decides)."""
    for g in range(N_GEARS - 1, -1, -1):
        if float(pred_snr[g]) >= target_db + margin_db:
            return g
    return None


# ---------------------------------------------------------------- the monitor

SNR_CLIP_DB = 60.0          # labels are clipped here: far above every target


def snr_db(ref: np.ndarray, got: np.ndarray) -> float:
    """10 log10(sum ref^2 / sum (got - ref)^2)
    (verify_m4h2_prod_calibration.py:64-69)."""
    noise = float(np.sum((got - ref) ** 2))
    return math.inf if noise == 0 else 10.0 * math.log10(float(np.sum(ref ** 2)) / noise)


def monitor(mlp, ct_out, X: np.ndarray, ref: Optional[np.ndarray] = None) -> Tuple[float, np.ndarray]:
    """The plain decrypt-based monitor: decrypt and decode the network output,
    compare it with the dense FP64 plaintext network on the same inputs (ref,
    when the caller has it already).  Returns (SNR over all samples,
    per-ciphertext SNR) in dB."""
    n = X.shape[0]
    got = mlp.decrypt_outputs(ct_out, n)
    if ref is None:
        ref = mlp.reference_outputs(X)
    S = mlp.S
    per = np.array([snr_db(ref[i:i + S], got[i:i + S]) for i in range(0, n, S)])
    return snr_db(ref, got), per


def sweep(mlp, n_ct: int, seed: int = 0, chunk_cts: int = 64, gears: Sequence[int] = range(N_GEARS)):
    """Training data from the plain monitor: random images, every gear.
    Returns (features [n_ct, 16], per-gear SNR labels [n_ct, 4] clipped to
    +-SNR_CLIP_DB) -- 4 (input, gear) samples per ciphertext."""
    from .nn import synthetic_images
    rng = np.random.default_rng(seed)
    S = mlp.S
    F, T = [], []
    for c0 in range(0, n_ct, chunk_cts):
        nc = min(chunk_cts, n_ct - c0)
        X = synthetic_images(nc * S, seed=int(rng.integers(1 << 62)))
        ct = mlp.encrypt_inputs(X)
        prev = int(rng.integers(N_GEARS))
        F.append(ciphertext_features(mlp.ck, ct, prev, netf=network_features(mlp)).cpu().numpy())
        lab = np.zeros((nc, N_GEARS))
        for g in gears:
            mlp.set_gears(g, g)
            _, per = monitor(mlp, mlp.forward(ct), X)
            lab[:, g] = np.clip(per, -SNR_CLIP_DB, SNR_CLIP_DB)
        T.append(lab)
    return np.concatenate(F), np.concatenate(T)


# ---------------------------------------------------------------- gear loops

@dataclass
class LoopResult:
    outputs: np.ndarray             # [n, 10] decrypted network outputs
    gears: List[int]                # gear used for each chunk
    evaluations: int                # encrypted network evaluations (chunks)
    monitor_calls: int
    fallbacks: int
    seconds: float                  # timed region: evaluation and gear control
    samples: int

    @property
    def samples_per_s(self) -> float:
        return self.samples / self.seconds


def _setup(mlp, X: np.ndarray, chunk_cts: int):
    """Untimed set-up shared by both loops: every gear's plans and keys,
    the encrypted input chunks, and one warm-up evaluation."""
    mlp.prepare()
    S = mlp.S
    per = chunk_cts * S
    chunks = [(X[i:i + per], mlp.encrypt_inputs(X[i:i + per])) for i in range(0, X.shape[0], per)]
    mlp.forward(chunks[0][1])
    return chunks


def _sync(mlp):
    torch.cuda.synchronize(mlp.ck.device)


@contextlib.contextmanager
def _timed_region():
    """The loops' timed regions run with Python's cyclic collector paused
    (as timeit does), after one collection: a collection pause would fall
    on whichever loop happened to be running."""
    gc.collect()
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


def _plain_select(mlp, ct, Xc, target_db, start_gear):
    """Monitor-driven gear search for one chunk: evaluate at the start gear,
    step to the next safer gear until the decrypt-based SNR meets the
    target (gear 0 is always accepted).  Returns (output ct, gear, evals)."""
    g, evals = start_gear, 0
    ref = mlp.reference_outputs(Xc)   # once per chunk, whatever the number of trials
    while True:
        mlp.set_gears(g, g)
        out = mlp.forward(ct)
        evals += 1
        s, _ = monitor(mlp, out, Xc, ref)
        if s >= target_db or g == 0:
            return out, g, evals
        g -= 1


def plain_monitor_loop(mlp, X: np.ndarray, target_db: float, chunk_cts: int = 64, probe_every: int = 4) -> LoopResult:
    """The plain decrypt-based monitor gear loop: every chunk is measured by the
    monitor; a chunk that misses the target is evaluated again one gear
    safer.  The search starts at the last accepted gear and, every
    `probe_every` chunks, one gear looser (the reference controller starts
    at the loosest gear and steps on monitor readings;
    verify_project_b_mlp_reference.py:279-317)."""
    chunks = _setup(mlp, X, chunk_cts)
    outs, gears, evals, mon = [], [], 0, 0
    last = N_GEARS - 1
    with _timed_region():
        _sync(mlp)
        t0 = time.perf_counter()
        for i, (Xc, ct) in enumerate(chunks):
            start = min(N_GEARS - 1, last + 1) if (i % probe_every == 0 and i) else last
            out, g, e = _plain_select(mlp, ct, Xc, target_db, start)
            evals += e
            mon += e
            last = g
            gears.append(g)
            outs.append(out)
        _sync(mlp)
        dt = time.perf_counter() - t0
    res = np.concatenate([mlp.decrypt_outputs(o, Xc.shape[0]) for o, (Xc, _) in zip(outs, chunks)])
    return LoopResult(res, gears, evals, mon, 0, dt, X.shape[0])


def _gear_table(mlp, net: GearNet, chunks) -> np.ndarray:
    """[previous gear, chunk, gear] mean predicted SNR: one batched
    prediction for every chunk and every previous-gear value (the only
    feature the decisions feed back into), with a single device sync."""
    feats = features_many(mlp.ck, [ct for _, ct in chunks], 0, netf=network_features(mlp))
    n = feats.shape[0]
    f = feats.repeat(N_GEARS, 1)
    f[:, 14] = torch.arange(N_GEARS, device=f.device, dtype=f.dtype).repeat_interleave(n) / 3.0
    pred = net.predict(f).view(N_GEARS, n, N_GEARS)
    sizes = torch.tensor([ct.batch for _, ct in chunks], device=f.device)
    seg = torch.repeat_interleave(torch.arange(len(chunks), device=f.device), sizes)
    sums = torch.zeros(N_GEARS, len(chunks), N_GEARS, device=f.device, dtype=pred.dtype).index_add_(1, seg, pred)
    return (sums / sizes.view(1, -1, 1).to(pred.dtype)).cpu().numpy()


def learned_loop(mlp, net: GearNet, X: np.ndarray, target_db: float, chunk_cts: int = 64,
                 margin_db: float = MARGIN_DB) -> LoopResult:
    """The learned gear loop: per chunk, on-device features -> predicted SNR
    of every gear -> the cheapest gear that clears target + margin; when no
    gear clears it, the plain monitor search decides the chunk."""
    chunks = _setup(mlp, X, chunk_cts)
    _gear_table(mlp, net, chunks[:2])            # warm the feature and network kernels
    outs, gears, evals, mon, fb = [], [], 0, 0, 0
    prev = 0
    with _timed_region():
        _sync(mlp)
        t0 = time.perf_counter()
        table = _gear_table(mlp, net, chunks)
        for i, (Xc, ct) in enumerate(chunks):
            pred = table[prev, i]
            g = choose_gear(pred, target_db, margin_db)
            if g is None:
                fb += 1
                out, g, e = _plain_select(mlp, ct, Xc, target_db, N_GEARS - 1)
                evals += e
                mon += e
            else:
                mlp.set_gears(g, g)
                out = mlp.forward(ct)
                evals += 1
            prev = g
            gears.append(g)
            outs.append(out)
        _sync(mlp)
        dt = time.perf_counter() - t0
    res = np.concatenate([mlp.decrypt_outputs(o, Xc.shape[0]) for o, (Xc, _) in zip(outs, chunks)])
    return LoopResult(res, gears, evals, mon, fb, dt, X.shape[0])


def static_loop(mlp, X: np.ndarray, g: int, chunk_cts: int = 64) -> LoopResult:
    """Every chunk at one fixed gear, no monitor (the reference point)."""
    chunks = _setup(mlp, X, chunk_cts)
    mlp.set_gears(g, g)
    outs = []
    with _timed_region():
        _sync(mlp)
        t0 = time.perf_counter()
        for _, ct in chunks:
            outs.append(mlp.forward(ct))
        _sync(mlp)
        dt = time.perf_counter() - t0
    res = np.concatenate([mlp.decrypt_outputs(o, Xc.shape[0]) for o, (Xc, _) in zip(outs, chunks)])
    return LoopResult(res, [g] * len(chunks), len(chunks), 0, 0, dt, X.shape[0])
