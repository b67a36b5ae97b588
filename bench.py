"""Throughput of CKKS HMult + relinearise + rescale on one MI355X.

Workload (one step): the fused ckks_hmult call over a batch of ciphertext
pairs at N = 2^13, L = 4 limbs (the four chain primes), one special prime,
dnum = 4, scale 2^40, inputs already resident in HBM; 256 pairs encrypting
4096 uniform real slots in [-1, 1] under a ternary secret, sigma = 3.2.

Prints one JSON line.  ``value`` = HMult/s over the timed steps; the
roofline entry prices the step at its algorithmic HBM bytes (inputs
1,048,576 B + outputs 393,216 B + the relinearisation key 2,621,440 B read
once per batch, i.e. 1,452,032 B per HMult at batch 256) against 8 TB/s;
``cpu_baseline`` times the CPU oracle on a bounded sample.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--batch 256]
"""
from __future__ import annotations

import argparse
import json
import multiprocessing as mp
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cipher-pu_amd"))

HBM_PEAK_GBS = 8000.0
N, L = 8192, 4


def algorithmic_bytes(batch: int, l: int = L) -> int:
    limb = N * 8
    inputs = 2 * 2 * l * limb                 # two ciphertexts, 2 polys, l limbs
    outputs = 2 * (l - 1) * limb
    key = l * 2 * (l + 1) * limb               # dnum digits x (b, a) x (l + special) limbs
    return batch * (inputs + outputs) + key


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _c_oracle_worker(args):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_c as OC
    cta, ctb, key, pr, p, ps, pp = args
    t = time.perf_counter()
    OC.hmult(cta, ctb, key, pr, p, ps, pp)
    return time.perf_counter() - t


def cpu_baseline(sample_s: float, cta, ctb, key_coeff, primes, roots):
    """The oracle's pure-Python HMult (big-int restatement of the
    reference's golden path) on whole ciphertext pairs for about sample_s
    seconds on one core; also the C restatement over all cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ckks_oracle as O
    import numpy as np

    pr, p, ps, pp = primes[:L], primes[L], roots[:L], roots[L]
    keyl = [([list(map(int, key_coeff[j, 0, i])) for i in range(L + 1)],
             [list(map(int, key_coeff[j, 1, i])) for i in range(L + 1)]) for j in range(L)]
    done, t0 = 0, time.perf_counter()
    while True:
        a = [[list(map(int, cta[done % len(cta), c, i])) for i in range(L)] for c in range(2)]
        b = [[list(map(int, ctb[done % len(ctb), c, i])) for i in range(L)] for c in range(2)]
        O.hmult(a, b, keyl, pr, p, ps + [pp], N)
        done += 1
        el = time.perf_counter() - t0
        if el >= sample_s or done >= len(cta):
            break
    py = {"value": done / el, "unit": "HMult/s", "cores": 1, "kind": "port",
          "sample": f"{done} HMult (N=8192, L=4, dnum=4) in pure-Python big-int arithmetic, {el:.1f} s"}
    # C restatement, one process per core over independent pairs
    cores = min(os.cpu_count() or 1, 16)
    n = max(cores, 8)
    jobs = [(cta[i % len(cta)], ctb[i % len(ctb)], key_coeff, pr, p, ps, pp) for i in range(n)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        pool.map(_c_oracle_worker, jobs)
    el = time.perf_counter() - t0
    c = {"value": n / el, "unit": "HMult/s", "cores": cores, "kind": "port",
         "sample": f"{n} HMult in the C oracle (unsigned __int128), {cores} processes, {el:.1f} s"}
    return py, c


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--cpu-sample-s", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ks-threads", type=int, default=0)
    ap.add_argument("--ntt-threads", type=int, default=0)
    ap.add_argument("--tensor-threads", type=int, default=0)
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape tuner")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no "
              "multi-GPU path (the batch shards over ciphertexts with no exchange).", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu.ring import to_numpy

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    B = args.batch
    ck = CKKS(CKKSParams(), seed=2024)
    sk = ck.keygen()
    rlk = ck.relin_key(sk)
    rng = np.random.default_rng(7)
    za = rng.uniform(-1, 1, (B, N // 2))
    zb = rng.uniform(-1, 1, (B, N // 2))
    cta = ck.encrypt(ck.encode(za), sk)
    ctb = ck.encrypt(ck.encode(zb), sk)
    tuned = None
    if args.ntt_threads or args.ks_threads or args.tensor_threads or args.no_tune:
        ck.ring.set_threads(args.ntt_threads, args.ks_threads, args.tensor_threads)
    else:
        from cipherpu.tune import tune_hmult
        tuned = tune_hmult(ck, cta, ctb, rlk, reps=3)
    out = ck.hmult(cta, ctb, rlk)
    torch.cuda.synchronize()
    # decrypted product vs the FP64 slot-wise product (all pairs)
    dec = to_numpy(ck.decode(ck.decrypt(out, sk), out.scale)).real
    max_err = float(np.max(np.abs(dec - za * zb)))

    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        ck.hmult(cta, ctb, rlk)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        ck.hmult(cta, ctb, rlk)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    value = B / (ms / 1e3)
    nbytes = algorithmic_bytes(B)
    achieved = nbytes / (ms / 1e3) / 1e9

    res = {
        "metric": "HMult+relin+rescale/s at N=2^13 L=4",
        "value": round(value, 1),
        "unit": "HMult/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "ckks_hmult_relin_rescale", "N": N, "L": L, "dnum": L, "special_primes": 1,
                   "batch": B, "scale_bits": 40, "prime_bits": 45, "domain": "coefficient in/out",
                   "parallelism": "single_gpu"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "ckks_hmult step (5 launches: 2x ntt_fwd, tensor_intt, keyswitch, moddown_rescale)",
                     "algorithmic_bytes_per_step": nbytes},
        "max_abs_err_vs_fp64": max_err,
        "launch_shape": ({k: tuned[k] for k in ("ntt", "tensor", "ks")} if tuned else
                         {"ntt": args.ntt_threads, "tensor": args.tensor_threads, "ks": args.ks_threads}),
        "wall_s_timed": round(wall, 4),
        "device": torch.cuda.get_device_name(0),
    }
    if not args.no_cpu_baseline:
        py, c = cpu_baseline(args.cpu_sample_s, to_numpy(cta.data[:8]), to_numpy(ctb.data[:8]),
                             to_numpy(rlk.coeff), ck.primes, ck.ring.roots)
        res["cpu_baseline"] = py
        res["cpu_baseline_c"] = c
        res["cpu_model"] = cpu_model()
        res["cpu_count"] = os.cpu_count()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
