"""CPU oracle for the learned gear controller's network -- TEST
INFRASTRUCTURE ONLY (tests/ import it to check the HIP forward/backward).

A numpy restatement of a dense ReLU network whose hidden activations are
stored in BF16 (round-to-nearest-even) and whose weights and arithmetic are
FP32, with the same backward rule as the kernels: the ReLU mask is taken
from the stored activation and weight gradients use the stored (BF16)
inputs.  The reference has no learned controller; this oracle pins the
kernels' gradients against finite differences of its own forward pass
(tests/test_oracle_gear.py)."""
from __future__ import annotations

import numpy as np


def bf16_round(x: np.ndarray) -> np.ndarray:
    """FP32 -> BF16 -> FP32, round to nearest even."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def forward(params, X, bf16_hidden=True):
    """params: list of (W [J][I], b [J]) float32.  Returns (Y, acts) with acts
    the stored input of every layer (acts[0] = X)."""
    acts = [np.asarray(X, dtype=np.float32)]
    h = acts[0]
    n = len(params)
    for k, (W, b) in enumerate(params):
        z = h @ W.T.astype(np.float32) + b.astype(np.float32)
        if k < n - 1:
            z = np.maximum(z, 0.0).astype(np.float32)
            if bf16_hidden:
                z = bf16_round(z)
            acts.append(z)
            h = z
        else:
            return z.astype(np.float32), acts
    raise ValueError("empty network")


def mse(Y, T):
    return float(np.mean(np.sum((Y - T) ** 2, axis=1)))


def backward(params, acts, Y, T):
    """Gradients of mean_b sum_j (Y - T)^2 w.r.t. every (W, b)."""
    B = Y.shape[0]
    d = (2.0 * (Y - T) / B).astype(np.float32)
    grads = [None] * len(params)
    for k in range(len(params) - 1, -1, -1):
        W, _ = params[k]
        if k < len(params) - 1:
            d = d * (acts[k + 1] > 0)
        grads[k] = (d.T @ acts[k], d.sum(axis=0))
        d = d @ W
    return grads
