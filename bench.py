"""Throughput of CKKS HMult + relinearise + rescale on one MI355X, and of
the encrypted 784-128-10 MLP at a fixed output SNR.

Workload (one step): the fused ckks_hmult call over a batch of ciphertext
pairs at N = 2^13, L = 4 limbs (the four chain primes), one special prime,
dnum = 4, scale 2^40, inputs already resident in HBM; 256 pairs encrypting
4096 uniform real slots in [-1, 1] under a ternary secret, sigma = 3.2.

Prints one JSON line.  ``value`` = HMult/s over the timed steps.  The
roofline entry is for the dominant kernel of the step (the key switch),
timed per launch with HIP events on the call's stream over the timed
region.  Its ``achieved`` prices one launch at the algorithmic bytes of the
units it processes: 1,452,032 B per HMult (two input ciphertexts of 2 x 4
limb-polys, one output of 2 x 3, the 2.5 MiB relinearisation key once per
batch of 256) x the batch.  ``kernel_model`` prices the same launch at the
bytes the key-switch kernel itself must move (per ciphertext: the
coefficient digits d2, l limbs; d2 of the own limb in evaluation form, l;
d0/d1 in evaluation form, 2 l; the accumulators written, 2 (l + 1); the
key once per launch).  ``traffic`` and ``valu_issue`` come from rocprofv3
counter runs of this command for the same kernel, batch and launch shape
(the file is named in ``source``).  ``cpu_baseline`` runs the oracle's
pure-Python HMult (the reference's big-integer path) over every host core
with multiprocessing; the one-core figure and the C oracle are beside it.
``mlp`` runs the encrypted MLP on 4096 synthetic images under the plain
decrypt-based monitor gear loop and under the learned gear controller;
``cnn`` does the same for the encrypted CNN (conv 5x5x8 stride 2, x^2,
average pool, FC 288-10); ``batch_graph`` runs 2048 mixed operations (N = 4096 / 8192, levels 1..4,
HMult / rotation / plaintext product) as one captured HIP graph; ``plain_product`` times 4096 negacyclic
products at N = 4096 over one 45-bit prime.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--batch 256] [--no-mlp] [--no-cnn] [--no-graph]
    python bench.py --copy-probe      # plain copy of known size (counter scaling)
    python bench.py --modmul-probe    # Barrett vs special-form point-wise product
    python bench.py --mfma-probe [--mfma-probe-reps R]   # 64-point step: radix-2 VALU vs int8 MFMA
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


BYTES_PER_HMULT = 1452032   # (2 x 2 x 4 + 2 x 3) limb-polys of 65,536 B + the 2,621,440 B key / 256


def algorithmic_bytes(batch: int, l: int = L) -> int:
    """Whole HMult step: inputs, outputs, the relinearisation key once."""
    limb = N * 8
    inputs = 2 * 2 * l * limb                 # two ciphertexts, 2 polys, l limbs
    outputs = 2 * (l - 1) * limb
    key = l * 2 * (l + 1) * limb               # dnum digits x (b, a) x (l + special) limbs
    return batch * (inputs + outputs) + key


KERNEL_NAMES = {"front": "k_front<13,{}>", "tensor": "k_tensor_intt<13,{}>", "ntt": "k_ntt_fwd<13,{}>",
                "keyswitch": "k_keyswitch<13,{},false,false> (limb l-1 and the special prime)",
                "keyswitch_fused": "k_keyswitch<13,{},false,true> (limbs 0..l-2, ModDown and rescale fused)",
                "moddown": "k_moddown_rescale_l<4,true>"}


def kernel_model_bytes(kind: str, batch: int, l: int = L, fp64: bool = True):
    """Bytes one launch of an HMult kernel must move, from its inputs and
    outputs (limb-polys of N 8-byte words; key rows, read once per launch,
    are the 8-byte words as doubles on the FP64 arithmetic and 16-byte value
    and companion pairs on the integer one):
      front            per (pair, limb): a0, a1, b0, b1 in; p d0, p d1, d2
                       (evaluation form) and d2 (coefficient form) out;
      keyswitch        targets l-1 and p, per pair: the other limbs' digits
                       (l - 1 and l), the own digit and p d0, p d1 for limb
                       l-1, the 2 x 2 accumulators out; key rows of both;
      keyswitch_fused  targets 0..l-2, per (pair, target): l - 1 digits, the
                       own digit, p d0, p d1, the accumulators of limb l-1
                       and of p (4 rows), 2 output limbs; their key rows."""
    limb = N * 8
    key_row = limb * (1 if fp64 else 2)
    if kind == "front":
        return batch * l * 8 * limb
    if kind == "keyswitch":
        return batch * ((l - 1 + 1 + 2 + 2) + (l + 2)) * limb + 2 * l * 2 * key_row
    if kind == "keyswitch_fused":
        return batch * (l - 1) * ((l - 1) + 1 + 2 + 4 + 2) * limb + (l - 1) * l * 2 * key_row
    return None


# Counter figures per launch of the key switch, from rocprofv3 --pmc runs of
# this bench (commands and tables in PROFILE_SOURCE): HBM bytes from
# FETCH_SIZE / WRITE_SIZE scaled by the same counters on the --copy-probe
# copy, and wave-level VALU instructions (SQ_INSTS_VALU).  Keyed by (batch,
# key-switch threads); the bench only runs shapes listed here.
PROFILE_SOURCE = "profiles/hmult_kernels_n8192_l4_b256.md"
MEASURED_TRAFFIC = {("front", 256, 1024): 543272998, ("keyswitch_fused", 256, 1024): 475038519,
                    ("keyswitch", 256, 1024): 205836936}
MEASURED_VALU_INSTR = {("front", 256, 1024): 53895168, ("keyswitch_fused", 256, 1024): 54665216,
                       ("keyswitch", 256, 1024): 33030144}
PROFILED_THREADS = (1024,)   # key-switch and front/tensor threads of the profiled shapes
# The issue peak counts one wave64 VALU instruction per SIMD every 4 cycles:
# 1024 SIMDs x 2.4 GHz / 4.
VALU_ISSUE_PEAK_G = 1024 * 2.4 / 4


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


_PY_JOB = {}


def _py_oracle_worker(i):
    """One pure-Python HMult of pair i of the shared job (fork-inherited)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ckks_oracle as O
    cta, ctb, keyl, pr, p, psis = (_PY_JOB[k] for k in ("cta", "ctb", "key", "pr", "p", "psis"))
    a = [[list(map(int, cta[i % len(cta), c, j])) for j in range(L)] for c in range(2)]
    b = [[list(map(int, ctb[i % len(ctb), c, j])) for j in range(L)] for c in range(2)]
    t = time.perf_counter()
    O.hmult(a, b, keyl, pr, p, psis, N)
    return time.perf_counter() - t


def cpu_baseline(sample_s: float, cta, ctb, key_coeff, primes, roots):
    """The oracle's pure-Python HMult (big-int restatement of the
    reference's golden path) on whole ciphertext pairs: one process per host
    core (at most 16) over independent pairs, two rounds of pairs; the same
    on one core for about sample_s seconds; and the C restatement over all
    cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))

    pr, p, ps, pp = primes[:L], primes[L], roots[:L], roots[L]
    keyl = [([list(map(int, key_coeff[j, 0, i])) for i in range(L + 1)],
             [list(map(int, key_coeff[j, 1, i])) for i in range(L + 1)]) for j in range(L)]
    _PY_JOB.update(cta=cta, ctb=ctb, key=keyl, pr=pr, p=p, psis=ps + [pp])
    cores = min(os.cpu_count() or 1, 16)
    n_all = 2 * cores
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        pool.map(_py_oracle_worker, range(n_all), chunksize=1)
    el_all = time.perf_counter() - t0
    py_all = {"value": n_all / el_all, "unit": "HMult/s", "cores": cores, "kind": "port",
              "sample": f"{n_all} HMult (N=8192, L=4, dnum=4) in pure-Python big-int arithmetic, "
                        f"{cores} processes, {el_all:.1f} s"}
    done, busy = 0, 0.0
    while busy < sample_s and done < len(cta):
        busy += _py_oracle_worker(done)
        done += 1
    py = {"value": done / busy, "unit": "HMult/s", "cores": 1, "kind": "port",
          "sample": f"{done} HMult (N=8192, L=4, dnum=4) in pure-Python big-int arithmetic, {busy:.1f} s"}
    # C restatement, one process per core over independent pairs
    n = max(cores, 8)
    jobs = [(cta[i % len(cta)], ctb[i % len(ctb)], key_coeff, pr, p, ps, pp) for i in range(n)]
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(cores) as pool:
        pool.map(_c_oracle_worker, jobs)
    el = time.perf_counter() - t0
    c = {"value": n / el, "unit": "HMult/s", "cores": cores, "kind": "port",
         "sample": f"{n} HMult in the C oracle (unsigned __int128), {cores} processes, {el:.1f} s"}
    return py_all, py, c


def copy_probe(nbytes: int = 1 << 28, reps: int = 5) -> None:
    """A plain device copy of known size, for scaling HBM counters."""
    import torch
    src = torch.empty(nbytes // 8, dtype=torch.int64, device="cuda").fill_(1)
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "launches": reps}))


def modmul_probe(rows: int = 1024, reps: int = 10) -> None:
    """Point-wise products over the four chain primes: Barrett (op 0)
    against the special-form reduction for q = 2^44 + c (op 4)."""
    import numpy as np
    import torch
    from cipherpu.ckks import CKKSParams
    from cipherpu.ring import RingContext
    p = CKKSParams()
    ring = RingContext(p.N, list(p.primes))
    rng = np.random.default_rng(0)
    a = ring.tensor(np.stack([rng.integers(0, q, (rows, p.N), dtype=np.uint64) for q in p.primes], axis=1))
    b = ring.tensor(np.stack([rng.integers(0, q, (rows, p.N), dtype=np.uint64) for q in p.primes], axis=1))
    out = torch.empty_like(a)
    res = {"elements": a.numel()}
    for name, op in (("barrett", 0), ("special_form", 4)):
        ring._pw(a, b, op, 0, out)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            ring._pw(a, b, op, 0, out)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        res[name] = {"ms": round(ms, 4), "G_products_per_s": round(a.numel() / ms / 1e6, 1),
                     "GB_per_s": round(3 * 8 * a.numel() / ms / 1e6, 1)}
    print(json.dumps(res))


def mfma_probe(count: int = 1 << 18, reps_list=(1, 17), launches: int = 5) -> None:
    """The matrix-core question, measured: one 64-point transform step (six
    radix-2 stages) over `count` vectors of 45-bit residues, on the integer
    VALU (ckks_probe_dft64 method 0) and as an exact int8 product on the
    matrix cores (methods 1 and 2, the two candidate accumulator layouts).
    Each kernel applies the step `reps` times to data held in registers; the
    time per step is the difference between two repetition counts, so the
    loads and stores cancel.  Outputs are compared word for word."""
    import torch
    from cipherpu import dft64_probe as P
    from cipherpu.params import CHAIN4, default_root
    q = CHAIN4[0]
    T = P.build_tables(q, default_root(q, 128))
    g = torch.Generator(device="cpu").manual_seed(5)
    x = (torch.randint(0, 1 << 62, (64, count), generator=g, dtype=torch.int64) % q).cuda()
    tab = torch.from_numpy(T.blob.copy()).cuda()
    names = ("radix2", "mfma_layout0", "mfma_layout1", "mfma_layout0_interleaved")
    ys = {m: torch.empty_like(x) for m in range(4)}
    res = {"vectors": count, "elements": 64 * count, "q": q, "launches": launches, "by_reps": {}}
    for reps in reps_list:
        row = {}
        for m in range(4):
            P.launch(x, ys[m], tab, T, m, reps)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(launches):
                P.launch(x, ys[m], tab, T, m, reps)
            e1.record()
            torch.cuda.synchronize()
            row[names[m]] = round(e0.elapsed_time(e1) / launches, 4)
        row["equal_to_radix2"] = [bool(torch.equal(ys[m], ys[0])) for m in (1, 2, 3)]
        res["by_reps"][str(reps)] = row
    if len(reps_list) == 2:
        r0, r1 = (res["by_reps"][str(r)] for r in reps_list)
        span = reps_list[1] - reps_list[0]
        per = {k: (r1[k] - r0[k]) / span for k in names}
        res["us_per_step"] = {k: round(v * 1e3, 2) for k, v in per.items()}
        res["G_elements_per_s_per_step"] = {k: round(64 * count / v / 1e6, 1) for k, v in per.items()}
        res["over_radix2"] = {k: round(per["radix2"] / per[k], 3)
                              for k, ok in zip(names[1:], r1["equal_to_radix2"]) if ok}
    print(json.dumps({"mfma_probe": res}))


Q1, PSI1 = 35184371884033, 7388705119   # largest 45-bit prime = 1 mod 8192, smallest primitive 8192-th root
PRODUCT_BYTES = 3 * 4096 * 8             # two operands in, one product out


def product_leg(count: int = 4096, reps: int = 20) -> dict:
    """Plain negacyclic products at N = 4096 over the single 45-bit prime:
    `count` products (forward transforms, point-wise product, inverse
    transform, one fused kernel per product) with operands uniform in [0, q),
    resident in HBM.  Every built thread count is timed; the fastest is
    reported, priced at 98,304 algorithmic bytes a product."""
    import numpy as np
    import torch
    from cipherpu.ring import RingContext
    from cipherpu.tune import BUILT
    ring = RingContext(4096, [Q1], roots=[PSI1])
    rng = np.random.default_rng(11)
    a_h = rng.integers(0, Q1, (count, 1, 4096), dtype=np.uint64)
    b_h = rng.integers(0, Q1, (count, 1, 4096), dtype=np.uint64)
    a, b = ring.tensor(a_h), ring.tensor(b_h)
    out = torch.empty_like(a)
    shapes, words = {}, {}
    for fp64 in (False, True):
        ring.set_arith(fp64)
        for t in BUILT[12]:
            ring.set_threads(t, 0, 0)
            for _ in range(3):
                ring.poly_mul(a, b, out=out)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(reps):
                ring.poly_mul(a, b, out=out)
            e1.record()
            torch.cuda.synchronize()
            shapes[(fp64, t)] = e0.elapsed_time(e1) / reps
        words[fp64] = out.clone()
    if not torch.equal(words[False], words[True]):
        raise RuntimeError("the FP64 and integer products differ")
    best = min(shapes, key=shapes.get)
    ms = shapes[best]
    achieved = PRODUCT_BYTES * count / ms / 1e6
    kern = ("k_polymul_f<12,{}>" if best[0] else "k_polymul<12,{}>").format(best[1])
    return {"workload": "negacyclic product, N=4096, one 45-bit prime", "products": count,
            "products_per_s": round(count / ms * 1e3, 1), "ms": round(ms, 4), "threads": best[1],
            "arith": "fp64" if best[0] else "int64",
            "ms_by_threads": {("fp64/" if f else "int64/") + str(k): round(v, 4) for (f, k), v in shapes.items()},
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(achieved / HBM_PEAK_GBS, 4), "kernel": kern,
                         "algorithmic_bytes_per_product": PRODUCT_BYTES}}


def arith_leg(ck, cta, ctb, rlk, want, fp64: bool, steps: int, warmup: int) -> dict:
    """The same HMult batch on the other residue arithmetic: the key prepared
    again from its coefficients, the same launch shapes, K timed steps, and
    the output words checked equal to the headline run's."""
    import torch
    ck.ring.set_arith(fp64)
    key = ck.load_switch_key(rlk.coeff)
    for _ in range(warmup):
        ck.hmult(cta, ctb, key)
    stream = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(stream)
    for _ in range(steps):
        ck.hmult(cta, ctb, key)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / steps
    ck.ring.timing_start(8 * steps)
    for _ in range(steps):
        ck.hmult(cta, ctb, key)
    torch.cuda.synchronize()
    kern = ck.ring.timing_stop()
    same = bool(torch.equal(ck.hmult(cta, ctb, key).data, want.data))
    ck.ring.set_arith(not fp64)
    if not same:
        raise RuntimeError("the two residue arithmetics gave different HMult words")
    return {"arith": "fp64" if fp64 else "int64", "value": round(cta.batch / (ms / 1e3), 1), "unit": "HMult/s",
            "ms_per_step": round(ms, 4), "kernel_ms_per_launch": {k: round(v[0] / v[1], 4) for k, v in kern.items()},
            "same_words": same}


def cpu_product_baseline(products: int = 2048) -> dict:
    """The C oracle's transform product (N = 4096, the single 45-bit prime)
    on one host core, for the plain-product leg."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle_c as OC
    rng = np.random.default_rng(11)
    a = rng.integers(0, Q1, (products, 4096), dtype=np.uint64)
    b = rng.integers(0, Q1, (products, 4096), dtype=np.uint64)
    t0 = time.perf_counter()
    for i in range(products):
        OC.polymul_ntt(a[i], b[i], Q1, PSI1)
    el = time.perf_counter() - t0
    return {"value": products / el, "unit": "products/s", "cores": 1, "kind": "port",
            "sample": f"{products} products (N=4096, one 45-bit prime) in the C oracle, {el:.2f} s"}


def rotation_probe(ck, sk, batch: int = 64, steps=tuple(range(1, 16)), reps: int = 5) -> dict:
    """15 rotations of the same ciphertexts (a baby-step set): one at a time
    against hoisted (one shared digit decomposition)."""
    import numpy as np
    import torch
    rng = np.random.default_rng(5)
    ct = ck.encrypt(ck.encode(rng.uniform(-1, 1, (batch, ck.N // 2))), sk)
    gks = [ck.galois_key(sk, s) for s in steps]
    keys = torch.stack([ck.hoist_key(g) for g in gks]).contiguous()
    res = {"ciphertexts": batch, "rotations": len(steps), "level": ct.level}
    for name, fn in (("one_at_a_time", lambda: [ck.rotate(ct, g) for g in gks]),
                     ("hoisted", lambda: ck.rotate_hoisted(ct, gks, keys))):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        res[name + "_ms"] = round(e0.elapsed_time(e1) / reps, 3)
    res["speedup"] = round(res["one_at_a_time_ms"] / res["hoisted_ms"], 2)
    return res


def mlp_leg(samples: int, targets, chunk_cts: int, kind: str = "mlp"):
    """Encrypted MLP (or CNN) samples/s at a fixed output SNR, for each
    target: the plain decrypt-based monitor gear loop against the learned
    gear controller; static gear 0 without any monitor beside them (the
    headline comparison for the learned loop)."""
    import torch
    from cipherpu import gear, nn
    from cipherpu.ckks import CKKS, CKKSParams
    ck = CKKS(CKKSParams(), seed=99)
    ck.ring.set_front(True)   # HMult's fused front, as the HMult leg's tuner picks at N = 8192
    sk = ck.keygen()
    if kind == "mlp":
        mlp = nn.EncryptedMLP(ck, sk, *nn.make_mlp(0))
        desc = "784-128-10 x^2, DARE between layers"
    else:
        from cipherpu import cnn
        mlp = cnn.EncryptedCNN(ck, sk, *cnn.make_cnn(0))
        desc = "conv 5x5x8 s2 -> x^2 -> avgpool 2x2 -> FC 288-10, DARE after the convolution"
    net = gear.GearNet(seed=0)
    trained = gear.load_trained(net, gear.TRAINED_PATHS[kind])
    if not trained:
        raise RuntimeError(f"no committed controller weights for the {kind} leg")
    X = nn.synthetic_images(samples, seed=4096)
    ref = mlp.reference_outputs(X)
    out = {"metric": f"enc-{kind.upper()} samples/s at fixed output SNR", "network": desc,
           "samples": samples, "target_snr_db": list(targets), "chunk_ciphertexts": chunk_cts,
           "samples_per_ciphertext": mlp.S, "controller_weights": trained,
           "levels": {"encrypted": ck.L, "stage_entry": list(mlp.entry_levels())}, "hmult_front": "fused",
           "predicted_snr_db_per_gear": None}

    def summary(r):
        return {"samples_per_s": round(r.samples_per_s, 1), "snr_db": round(gear.snr_db(ref, r.outputs), 2),
                "evaluations": r.evaluations, "monitor_calls": r.monitor_calls, "fallbacks": r.fallbacks,
                "gears": {str(g): r.gears.count(g) for g in sorted(set(r.gears))}}
    # one untimed chunk through every loop first, so that first-launch costs
    # (code-object loading, allocator growth) fall on none of the timed loops
    Xw = X[:chunk_cts * mlp.S]
    for g in range(4):   # every gear's first evaluation (plans at the stage levels) untimed
        gear.static_loop(mlp, Xw, g, chunk_cts)
    gear.plain_monitor_loop(mlp, Xw, targets[0], chunk_cts)
    gear.learned_loop(mlp, net, Xw, targets[0], chunk_cts)
    torch.cuda.synchronize()
    chunks = gear._setup(mlp, X[:chunk_cts * mlp.S], chunk_cts)
    out["predicted_snr_db_per_gear"] = [round(float(v), 2) for v in gear._gear_table(mlp, net, chunks)[0, 0]]
    out["static_gear0"] = summary(gear.static_loop(mlp, X, 0, chunk_cts))
    # the same with every stage from the top level (inputs and refresh at L)
    mlp.drop_levels = False
    gear.static_loop(mlp, Xw, 0, chunk_cts)
    out["static_gear0_top_levels"] = summary(gear.static_loop(mlp, X, 0, chunk_cts))
    mlp.drop_levels = True
    by = {}
    for t in targets:
        plain = summary(gear.plain_monitor_loop(mlp, X, t, chunk_cts))
        learned = summary(gear.learned_loop(mlp, net, X, t, chunk_cts))
        by[str(t)] = {"plain_monitor": plain, "learned": learned,
                      "learned_over_plain": round(learned["samples_per_s"] / plain["samples_per_s"], 3),
                      "learned_over_static_gear0": round(learned["samples_per_s"] /
                                                         out["static_gear0"]["samples_per_s"], 3)}
    out["by_target"] = by
    return out


def graph_leg(n_ops: int, reps: int = 5):
    """Uneven batch: n_ops operations at N = 4096 / 8192, levels 1..4, mix
    60% HMult+rescale / 30% rotation / 10% plaintext product, bucketed,
    tuned per bucket and replayed as one captured HIP graph."""
    import torch
    from cipherpu.sched import BatchScheduler, make_workload
    ops = make_workload(n_ops, seed=0)
    s = BatchScheduler(ops, seed=1)
    s.fill_random(seed=2)
    t0 = time.perf_counter()
    s.tune(reps=2)
    t_tune = time.perf_counter() - t0
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.run_eager()
    e0.record()
    for _ in range(reps):
        s.run_eager()
    e1.record()
    torch.cuda.synchronize()
    eager_ms = e0.elapsed_time(e1) / reps
    s.capture()
    s.replay()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        s.replay()
    e1.record()
    torch.cuda.synchronize()
    graph_ms = e0.elapsed_time(e1) / reps
    kinds = [o.kind for o in ops]
    return {"workload": "mixed N 4096/8192, levels 1..4, 60/30/10 hmult/rotate/ptmul", "ops": n_ops,
            "ops_per_s": round(n_ops / (graph_ms / 1e3), 1), "ms_per_batch": round(graph_ms, 3),
            "eager_ms_per_batch": round(eager_ms, 3), "buckets": len(s.buckets), "streams": s.n_streams,
            "kinds": {k: kinds.count(k) for k in ("hmult", "rotate", "ptmul")}, "tune_s": round(t_tune, 1)}


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
    ap.add_argument("--front", action="store_true", help="fused HMult front kernel (with fixed shapes)")
    ap.add_argument("--no-tune", action="store_true", help="skip the launch-shape tuner")
    ap.add_argument("--no-mlp", action="store_true", help="skip the encrypted-MLP leg")
    ap.add_argument("--no-rotation-probe", action="store_true", help="skip the hoisted-rotation probe")
    ap.add_argument("--no-cnn", action="store_true", help="skip the encrypted-CNN leg")
    ap.add_argument("--mlp-samples", type=int, default=4096)
    ap.add_argument("--chunk-cts", type=int, default=64,
                    help="ciphertexts per gear decision in the encrypted-network legs")
    ap.add_argument("--target-snr", type=float, nargs="+", default=[3.0, 10.0, 15.0, 20.0],
                    help="output SNR targets (dB) of the encrypted-network legs: 3 dB and the reference's "
                         "gear thresholds 10 / 15 / 20 dB")
    ap.add_argument("--copy-probe", action="store_true")
    ap.add_argument("--modmul-probe", action="store_true", help="Barrett vs special-form point-wise product")
    ap.add_argument("--mfma-probe", action="store_true",
                    help="64-point transform step: radix-2 VALU stages vs the int8 matrix-core product")
    ap.add_argument("--mfma-probe-reps", type=int, default=0,
                    help="run the probe at this one repetition count only (for profiler runs)")
    ap.add_argument("--no-graph", action="store_true", help="skip the uneven-batch graph leg")
    ap.add_argument("--no-product", action="store_true", help="skip the N=4096 single-prime product leg")
    ap.add_argument("--graph-ops", type=int, default=2048)
    ap.add_argument("--arith", choices=("fp64", "int64"), default="fp64",
                    help="residue arithmetic of the timed HMult (ckks_ctx_set_arith)")
    ap.add_argument("--no-arith-leg", action="store_true",
                    help="skip timing the other residue arithmetic on the same batch")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the library runs on a single MI355X and has no "
              "multi-GPU path (the batch shards over ciphertexts with no exchange).", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe()
        return
    if args.modmul_probe:
        modmul_probe()
        return
    if args.mfma_probe:
        mfma_probe(reps_list=(args.mfma_probe_reps,) if args.mfma_probe_reps else (1, 17))
        return

    import numpy as np
    import torch
    from cipherpu.ckks import CKKS, CKKSParams
    from cipherpu.ring import to_numpy

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    B = args.batch
    ck = CKKS(CKKSParams(), seed=2024)
    ck.ring.set_arith(args.arith == "fp64")
    sk = ck.keygen()
    rlk = ck.relin_key(sk)
    rng = np.random.default_rng(7)
    za = rng.uniform(-1, 1, (B, N // 2))
    zb = rng.uniform(-1, 1, (B, N // 2))
    cta = ck.encrypt(ck.encode(za), sk)
    ctb = ck.encrypt(ck.encode(zb), sk)
    tuned = None
    if args.ntt_threads or args.ks_threads or args.tensor_threads or args.no_tune:
        ck.ring.set_front(args.front)
        ck.ring.set_threads(args.ntt_threads, args.ks_threads, args.tensor_threads)
    else:
        from cipherpu.tune import tune_hmult
        # at a profiled batch the tuner stays on the profiled key-switch and
        # front shapes, so the counter figures belong to the kernels timed
        prof = B in {b for _, b, _ in MEASURED_TRAFFIC}
        tuned = tune_hmult(ck, cta, ctb, rlk, reps=10, rounds=4, ks=PROFILED_THREADS if prof else None,
                           tensor=PROFILED_THREADS if prof else None)
    stream = torch.cuda.current_stream()
    for _ in range(args.warmup):
        ck.hmult(cta, ctb, rlk)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    # the timed region: K steps, nothing but the HMult launches
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        ck.hmult(cta, ctb, rlk)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1) / args.steps
    # the same K steps again with a HIP event pair around every kernel launch
    # (on the call's stream), for the per-kernel launch durations
    ck.ring.timing_start(8 * args.steps)
    ev0.record(stream)
    for _ in range(args.steps):
        ck.hmult(cta, ctb, rlk)
    ev1.record(stream)
    torch.cuda.synchronize()
    kern = ck.ring.timing_stop()
    ms_kernel_timed = ev0.elapsed_time(ev1) / args.steps
    # decrypted product vs the FP64 slot-wise product (all pairs), after the
    # timed steps so that the host-side check leaves no idle gap before them
    out = ck.hmult(cta, ctb, rlk)
    torch.cuda.synchronize()
    dec = to_numpy(ck.decode(ck.decrypt(out, sk), out.scale)).real
    max_err = float(np.max(np.abs(dec - za * zb)))
    value = B / (ms / 1e3)
    nbytes = algorithmic_bytes(B)
    shape = ({k: tuned[k] for k in ("front", "ntt", "tensor", "ks")} if tuned else
             {"front": int(args.front), "ntt": args.ntt_threads, "tensor": args.tensor_threads,
              "ks": args.ks_threads})
    if tuned:
        # the tuner's best HMult batch time for each front, as it measured them
        for f, name in ((0, "tuned_ms_separate_front"), (1, "tuned_ms_fused_front")):
            ts = [v for k, v in tuned["grid"].items() if k.startswith(f"{f}/")]
            if ts:
                shape[name] = round(min(ts), 4)
    # the dominant kernel: the longest launch of the step
    per_launch = {k: v[0] / v[1] for k, v in kern.items()}
    dom = max(per_launch, key=per_launch.get)
    dom_ms = per_launch[dom]
    threads = {"front": shape.get("tensor"), "tensor": shape.get("tensor"), "keyswitch": shape.get("ks"),
               "keyswitch_fused": shape.get("ks"), "ntt": shape.get("ntt"), "moddown": 256}.get(dom)
    kbytes = kernel_model_bytes(dom, B, fp64=args.arith == "fp64")
    k_achieved = nbytes / (dom_ms / 1e3) / 1e9
    prof_key = (dom, B, threads)
    if prof_key in MEASURED_TRAFFIC and args.arith == "fp64":
        traffic = MEASURED_TRAFFIC[prof_key]
        rate = MEASURED_VALU_INSTR[prof_key] / (dom_ms / 1e3) / 1e9
        valu = {"wave_instructions": MEASURED_VALU_INSTR[prof_key], "achieved": round(rate, 1),
                "peak": VALU_ISSUE_PEAK_G, "unit": "G wave-instr/s", "frac": round(rate / VALU_ISSUE_PEAK_G, 4),
                "source": PROFILE_SOURCE}
    elif B == 256 and args.arith == "fp64" and not (args.ntt_threads or args.ks_threads or args.tensor_threads
                                                    or args.no_tune):
        raise RuntimeError(f"no counter profile for {dom} at batch {B}, {threads} threads")
    else:   # an explicit shape or batch outside the profiled ones: say so
        traffic, valu = None, {"source": f"not profiled: {dom} at batch {B}, {threads} threads, {args.arith}"}

    res = {
        "metric": "HMult+relin+rescale/s at N=2^13 L=4",
        "value": round(value, 1),
        "unit": "HMult/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "ms_per_step_with_kernel_events": round(ms_kernel_timed, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.arith == "fp64" else "u64",
        "data": "synthetic",
        "config": {"workload": "ckks_hmult_relin_rescale", "N": N, "L": L, "dnum": L, "special_primes": 1,
                   "batch": B, "scale_bits": 40, "prime_bits": 45, "domain": "coefficient in/out",
                   "residue_arith": args.arith,
                   "parallelism": "single_gpu"},
        "roofline": {"bound": "hbm", "achieved": round(k_achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(k_achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": KERNEL_NAMES[dom].format(threads),
                     "avg_launch_ms": round(dom_ms, 4), "algorithmic_bytes_per_launch": nbytes,
                     "units_per_launch": B, "bytes_per_unit": BYTES_PER_HMULT,
                     "traffic_source": PROFILE_SOURCE if traffic else None,
                     "kernel_model": kbytes and {"bytes_per_launch": kbytes,
                                                 "achieved": round(kbytes / (dom_ms / 1e3) / 1e9, 1),
                                                 "frac": round(kbytes / (dom_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}},
        "valu_issue": valu,
        "kernel_ms_per_step": {k: round(v[0] / args.steps, 4) for k, v in kern.items()},
        "launches_per_step": {k: v[1] // args.steps for k, v in kern.items()},
        "step_roofline": {"achieved": round(nbytes / (ms / 1e3) / 1e9, 1), "unit": "GB/s",
                          "frac": round(nbytes / (ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
                          "algorithmic_bytes_per_step": nbytes},
        "max_abs_err_vs_fp64": max_err,
        "launch_shape": shape,
        "wall_s_timed": round(wall, 4),
        "device": torch.cuda.get_device_name(0),
    }
    if not args.no_arith_leg:
        res["other_arith"] = arith_leg(ck, cta, ctb, rlk, out, args.arith != "fp64", args.steps, args.warmup)
    if not args.no_rotation_probe:
        res["rotation_probe"] = rotation_probe(ck, sk)
    if not args.no_product:
        res["plain_product"] = product_leg()
    if not args.no_mlp:
        res["mlp"] = mlp_leg(args.mlp_samples, args.target_snr, args.chunk_cts)
    if not args.no_cnn:
        res["cnn"] = mlp_leg(args.mlp_samples, args.target_snr, args.chunk_cts, kind="cnn")
    if not args.no_graph:
        res["batch_graph"] = graph_leg(args.graph_ops)
    if not args.no_cpu_baseline:
        py_all, py, c = cpu_baseline(args.cpu_sample_s, to_numpy(cta.data[:32]), to_numpy(ctb.data[:32]),
                                     to_numpy(rlk.coeff), ck.primes, ck.ring.roots)
        res["cpu_baseline"] = py_all
        res["cpu_baseline_1core"] = py
        res["cpu_baseline_c"] = c
        if not args.no_product:
            res["cpu_baseline_product"] = cpu_product_baseline()
        res["cpu_model"] = cpu_model()
        res["cpu_count"] = os.cpu_count()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
