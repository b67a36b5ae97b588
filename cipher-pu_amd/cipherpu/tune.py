"""Launch-shape tuner for the LDS-resident kernels.

For a ring degree, level and batch it times the real operation (the fused
HMult, or a batched transform) over every built thread count of the
transform, tensor and key-switch kernels, in this process on this GPU, and
returns the fastest setting.  Results are cached per (device, N, l, batch).
Threads per workgroup decide coefficients per thread (N / T), hence the
radix of the register passes and the number of LDS exchanges."""
from __future__ import annotations

import itertools
import json
import os
from typing import Dict, Iterable, Tuple

import torch

BUILT = {12: (128, 256, 512), 13: (256, 512, 1024), 14: (512, 1024)}
BUILT_FUSED = {12: (256, 512), 13: (256, 512, 1024), 14: (512, 1024)}

_cache: Dict[Tuple, dict] = {}


def _time(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_hmult(ck, a, b, rlk, reps: int = 5, ntt: Iterable[int] = None, tensor: Iterable[int] = None,
               ks: Iterable[int] = None, fused: Iterable[int] = (0, 1), rounds: int = 3) -> dict:
    """Grid search for ck.hmult(a, b, rlk) over the HMult front (two batched
    transforms + tensor kernel, or the fused front kernel) and the thread
    counts of the transform, tensor/front and key-switch kernels.  The grid
    is walked `rounds` times and each setting keeps its fastest time, so a
    setting timed while the clocks were still rising is not penalised.
    Leaves the best setting applied on ck.ring and returns
    {"front": 0|1, "ntt": T, "tensor": T, "ks": T, "ms": best, "grid": {...}}."""
    logN = ck.N.bit_length() - 1
    key = (torch.cuda.get_device_name(ck.device), ck.N, a.level, a.batch)
    if key in _cache:
        r = _cache[key]
        ck.ring.set_front(bool(r["front"]))
        ck.ring.set_threads(r["ntt"], r["ks"], r["tensor"])
        return r
    ntt = tuple(ntt or BUILT[logN])
    tensor = tuple(tensor or BUILT_FUSED[logN])
    ks = tuple(ks or BUILT_FUSED[logN])
    # the fused front does not use the transform kernel's thread count
    settings = [(f, n, t, k) for f in fused
                for n, t, k in itertools.product(ntt if not f else ntt[:1], tensor, ks)]
    grid = {}
    for rnd in range(max(1, rounds)):
        # every other pass walks the grid backwards, so that a drift in the
        # clocks during a pass favours no end of it
        for f, n, t, k in (settings if rnd % 2 == 0 else settings[::-1]):
            ck.ring.set_front(bool(f))
            ck.ring.set_threads(n, k, t)
            ms = _time(lambda: ck.hmult(a, b, rlk), reps)
            name = f"{f}/{n}/{t}/{k}"
            grid[name] = min(ms, grid.get(name, ms))
    best = min(grid, key=grid.get)
    f, n, t, k = (int(x) for x in best.split("/"))
    ck.ring.set_front(bool(f))
    ck.ring.set_threads(n, k, t)
    r = {"front": f, "ntt": n, "tensor": t, "ks": k, "ms": grid[best], "grid": grid}
    _cache[key] = r
    return r


def tune_ntt(ring, rows: int, limbs: int, reps: int = 10) -> dict:
    """Fastest thread count for a batched forward transform of rows x limbs."""
    logN = ring.N.bit_length() - 1
    x = torch.zeros(rows, limbs, ring.N, dtype=torch.int64, device=ring.device).view(torch.uint64)
    grid = {}
    for T in BUILT.get(logN, (0,)):
        ring.set_threads(T, 0, 0)
        grid[T] = _time(lambda: ring.ntt_(x), reps)
    best = min(grid, key=grid.get)
    ring.set_threads(best, 0, 0)
    return {"ntt": best, "ms": grid[best], "grid": grid}


def save(path: str) -> None:
    with open(path, "w") as f:
        json.dump({"|".join(map(str, k)): v for k, v in _cache.items()}, f, indent=1)


def load(path: str) -> None:
    if not os.path.exists(path):
        return
    with open(path) as f:
        for k, v in json.load(f).items():
            dev, N, l, B = k.split("|")
            _cache[(dev, int(N), int(l), int(B))] = v
