"""Train the learned gear controller on plain-monitor sweeps, on the GPU.

    python -m cipherpu.gear_train [--net mlp|cnn] [--samples 200000] [--epochs 40] [--out PATH]

A sample is one (ciphertext of 4 synthetic images, gear) pair for the
encrypted MLP (nn.make_mlp(0)) or CNN (cnn.make_cnn(0)): its
features (per-limb coefficient moments of the input ciphertext, level,
scale, previous gear) and the output SNR that the plain decrypt-based
monitor measures for that gear against the FP64 network.  A held-out
sweep of fresh inputs gives the predictor's error, as measured by the monitor.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=200_000, help="(ciphertext, gear) samples")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--chunk", type=int, default=256, help="ciphertexts per sweep evaluation")
    ap.add_argument("--net", choices=("mlp", "cnn"), default="mlp")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    from . import gear, nn
    from .ckks import CKKS, CKKSParams
    ck = CKKS(CKKSParams(), seed=5)
    sk = ck.keygen()
    if args.net == "mlp":
        mlp = nn.EncryptedMLP(ck, sk, *nn.make_mlp(0))
    else:
        from . import cnn
        mlp = cnn.EncryptedCNN(ck, sk, *cnn.make_cnn(0))
    mlp.prepare()
    n_ct = max(1, args.samples // gear.N_GEARS)
    t0 = time.perf_counter()
    Fs, Ts = [], []
    block = max(args.chunk, 4096)
    for i, c0 in enumerate(range(0, n_ct, block)):
        f, t = gear.sweep(mlp, min(block, n_ct - c0), seed=17 + 1000 * i, chunk_cts=args.chunk)
        Fs.append(f)
        Ts.append(t)
        print(f"sweep {c0 + f.shape[0]}/{n_ct} ciphertexts, {time.perf_counter() - t0:.1f} s", flush=True)
    F, T = np.concatenate(Fs), np.concatenate(Ts)
    t_sweep = time.perf_counter() - t0
    Fh, Th = gear.sweep(mlp, max(64, n_ct // 20), seed=999_999, chunk_cts=args.chunk)
    net = gear.GearNet(seed=0)
    t0 = time.perf_counter()
    hist = net.fit(F, T, epochs=args.epochs, batch=args.batch, lr=args.lr)
    t_fit = time.perf_counter() - t0
    pred = net.predict(Fh).cpu().numpy()
    mae = np.abs(pred - Th).mean(axis=0)
    out = args.out or gear.TRAINED_PATHS[args.net]
    import os
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    net.save(out)
    print(json.dumps({"samples": int(F.shape[0] * gear.N_GEARS), "ciphertexts": int(F.shape[0]),
                      "sweep_s": round(t_sweep, 1), "fit_s": round(t_fit, 1), "loss_first": hist[0],
                      "loss_last": hist[-1], "heldout_mae_db_per_gear": [round(float(v), 3) for v in mae],
                      "label_mean_db": [round(float(v), 2) for v in T.mean(axis=0)],
                      "label_std_db": [round(float(v), 3) for v in T.std(axis=0)], "out": out}))


if __name__ == "__main__":
    main()
