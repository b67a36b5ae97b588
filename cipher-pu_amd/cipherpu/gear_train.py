"""Train the learned gear controller on plain-monitor sweeps, on the GPU.

    python -m cipherpu.gear_train [--net mlp|cnn] [--nets 40] [--samples 200000] [--epochs 60] [--out PATH]

A sample is one (ciphertext of 4 synthetic images, gear) pair of one
network: its features (the network's per-gear pruning statistics, per-limb
coefficient moments of the input ciphertext, level, scale, previous gear)
and the output SNR that the plain decrypt-based monitor measures for that
gear against the FP64 network.  The sweep covers `--nets` networks of the
architecture, drawn from the weight families of nn.WEIGHT_FAMILIES
(nn.make_mlp_family / cnn.make_cnn_family, seeds 0..nets-1).  Held-out
networks (seeds 1000..1004, one per family) give the predictor's error as
the monitor measures it, beside that of the constant predictor that always
answers each gear's mean training label.
"""
from __future__ import annotations

import argparse
import gc
import json
import time

import numpy as np


def _network(kind: str, ck, sk, seed: int):
    from . import nn
    if kind == "mlp":
        return nn.EncryptedMLP(ck, sk, *nn.make_mlp_family(seed))
    from . import cnn
    return cnn.EncryptedCNN(ck, sk, *cnn.make_cnn_family(seed))


def sweep_networks(kind: str, ck, sk, seeds, n_ct: int, chunk: int, tag: str, t0: float):
    """Monitor sweeps of n_ct ciphertexts over every gear of each network."""
    import torch
    from . import gear
    Fs, Ts = [], []
    for k, seed in enumerate(seeds):
        net = _network(kind, ck, sk, seed)
        net.prepare()
        f, t = gear.sweep(net, n_ct, seed=17 + 1000 * seed, chunk_cts=chunk)
        Fs.append(f)
        Ts.append(t)
        print(f"{tag}: network {k + 1}/{len(seeds)} (seed {seed}) mean SNR dB per gear "
              f"{np.round(t.mean(axis=0), 1).tolist()}, {time.perf_counter() - t0:.1f} s", flush=True)
        del net
        gc.collect()
        torch.cuda.empty_cache()
    return np.concatenate(Fs), np.concatenate(Ts)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=200_000, help="(ciphertext, gear) samples in all")
    ap.add_argument("--nets", type=int, default=40, help="training networks")
    ap.add_argument("--heldout-cts", type=int, default=256, help="ciphertexts per held-out network")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--chunk", type=int, default=256, help="ciphertexts per sweep evaluation")
    ap.add_argument("--net", choices=("mlp", "cnn"), default="mlp")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    from . import gear
    from .ckks import CKKS, CKKSParams
    ck = CKKS(CKKSParams(), seed=5)
    sk = ck.keygen()
    n_ct = max(1, args.samples // (gear.N_GEARS * args.nets))
    t0 = time.perf_counter()
    F, T = sweep_networks(args.net, ck, sk, range(args.nets), n_ct, args.chunk, "train", t0)
    t_sweep = time.perf_counter() - t0
    Fh, Th = sweep_networks(args.net, ck, sk, range(1000, 1005), args.heldout_cts, args.chunk, "held-out", t0)
    net = gear.GearNet(seed=0)
    t1 = time.perf_counter()
    hist = net.fit(F, T, epochs=args.epochs, batch=args.batch, lr=args.lr)
    t_fit = time.perf_counter() - t1
    pred = net.predict(Fh).cpu().numpy()
    mae = np.abs(pred - Th).mean(axis=0)
    mae_const = np.abs(T.mean(axis=0)[None, :] - Th).mean(axis=0)
    out = args.out or gear.TRAINED_PATHS[args.net]
    import os
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    net.save(out)
    print(json.dumps({"net": args.net, "networks": args.nets, "samples": int(F.shape[0] * gear.N_GEARS),
                      "ciphertexts": int(F.shape[0]), "sweep_s": round(t_sweep, 1), "fit_s": round(t_fit, 1),
                      "loss_first": hist[0], "loss_last": hist[-1],
                      "heldout_mae_db_per_gear": [round(float(v), 3) for v in mae],
                      "heldout_mae_db_constant_per_gear": [round(float(v), 3) for v in mae_const],
                      "label_mean_db": [round(float(v), 2) for v in T.mean(axis=0)],
                      "label_std_db": [round(float(v), 3) for v in T.std(axis=0)], "out": out}))


if __name__ == "__main__":
    main()
