"""Uneven batched CKKS evaluation in one captured HIP graph.

A workload is a list of operations on independent ciphertexts at mixed
ring degrees and levels: HMult + relinearise + rescale with a partner
ciphertext, rotation by a power-of-two slot step with its Galois key, or a
plaintext-ciphertext product.  The scheduler

  * buckets the operations by (N, level, kind) -- every bucket is one
    batched call of the C ABI on contiguous device buffers that the
    scheduler owns (inputs are written bucket-major, so no gather pass);
    rotations by different steps share a bucket, each ciphertext with its
    own Galois key (ckks_rotate_many);
  * chooses a launch shape per bucket with the tuner (threads per
    workgroup of the transform, tensor and key-switch kernels, i.e. the
    radix of the register passes);
  * spreads the buckets over a few streams (greedy by their tuned time),
    each with its own workspace, so that small buckets run side by side
    instead of each leaving most of the GPU idle;
  * captures every bucket's launches, each under its own launch shape, into
    one HIP graph (the streams fork from and join back into the capture
    stream), replayed as a single launch.

Level semantics: a ciphertext at level l has l rescales left and l + 1 RNS
limbs over the primes CHAIN5 (cipherpu.params), plus the special prime for
key switching.  The reference has no scheduler; its RTL time-multiplexes
one datapath (scripts/exp_e17_*:92-115 sketches that)."""
from __future__ import annotations

import ctypes
import itertools
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from ._lib import call
from .ckks import CKKS, CKKSParams, Ciphertext, SecretKey, SwitchKey
from .params import CHAIN5

KINDS = ("hmult", "rotate", "ptmul")


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@dataclass(frozen=True)
class Op:
    kind: str          # "hmult" | "rotate" | "ptmul"
    N: int
    level: int         # rescales left; limbs = level + 1
    step: int = 0      # rotation step (slots)


def make_workload(n: int = 2048, seed: int = 0, degrees=(4096, 8192), levels=(1, 2, 3, 4),
                  mix=(0.6, 0.3, 0.1)) -> List[Op]:
    """n operations: half at each ring degree, level uniform in `levels`,
    kind drawn with probabilities `mix` (HMult, rotation, plaintext
    product), rotation steps uniform over the powers of two below N/2."""
    rng = np.random.default_rng(seed)
    ops = []
    for i in range(n):
        N = degrees[i * len(degrees) // n]
        level = int(rng.choice(levels))
        kind = KINDS[int(rng.choice(3, p=mix))]
        step = 1 << int(rng.integers(0, int(np.log2(N // 2)))) if kind == "rotate" else 0
        ops.append(Op(kind, N, level, step))
    return ops


def bucket_ops(ops: List[Op]) -> List[Tuple[int, int, str, List[int], List[int]]]:
    """Group operations by (N, level, kind), sorted: (N, level, kind, op
    indices in workload order, rotation step of each op -- empty for the
    other kinds).  Rotations by different steps share a group."""
    groups: Dict[tuple, List[int]] = {}
    for i, o in enumerate(ops):
        groups.setdefault((o.N, o.level, o.kind), []).append(i)
    return [(N, lv, k, idx, [ops[i].step for i in idx] if k == "rotate" else [])
            for (N, lv, k), idx in sorted(groups.items())]


def place(costs: List[float], n_streams: int) -> List[int]:
    """Longest-first greedy placement of items with the given costs on
    n_streams streams: each item goes to the stream with the least load so
    far.  Returns the stream of each item."""
    load = [0.0] * n_streams
    where = [0] * len(costs)
    for i in sorted(range(len(costs)), key=lambda i: costs[i], reverse=True):
        k = min(range(n_streams), key=load.__getitem__)
        where[i] = k
        load[k] += costs[i]
    return where


@dataclass
class Bucket:
    N: int
    level: int
    kind: str
    ops: List[int]                          # indices into the workload
    steps: List[int] = field(default_factory=list)   # rotation step of each op (rotate buckets)
    tables: Tuple[torch.Tensor, torch.Tensor] = None  # per-ciphertext key addresses and Galois elements
    a: torch.Tensor = None                  # [B, 2, l+1, N] inputs
    b: torch.Tensor = None                  # partner ciphertexts (hmult) or plaintexts [B*2, l+1, N] (ptmul)
    out: torch.Tensor = None
    shape: Tuple[int, int, int] = (0, 0, 0)  # (ntt, ks, tensor) threads
    ms: float = 0.0
    stream: int = 0                         # index of the scheduler stream it runs on

    @property
    def limbs(self) -> int:
        return self.level + 1

    @property
    def batch(self) -> int:
        return len(self.ops)


class BatchScheduler:
    def __init__(self, ops: List[Op], seed: int = 0, device: int = 0, streams: int = 4):
        self.ops = ops
        self.n_streams = max(1, streams)
        self._streams: List[torch.cuda.Stream] = []
        self.ck: Dict[int, CKKS] = {}
        self.sk: Dict[int, SecretKey] = {}
        self.rlk: Dict[int, SwitchKey] = {}
        self.gk: Dict[Tuple[int, int], SwitchKey] = {}
        for N in sorted({o.N for o in ops}):
            ck = CKKS(CKKSParams(N=N, primes=CHAIN5), device=device, seed=seed + N)
            self.ck[N] = ck
            self.sk[N] = ck.keygen()
            self.rlk[N] = ck.relin_key(self.sk[N])
        for N, step in sorted({(o.N, o.step) for o in ops if o.kind == "rotate"}):
            self.gk[(N, step)] = self.ck[N].galois_key(self.sk[N], step)
        self.buckets = [Bucket(N, lv, k, idx, steps) for N, lv, k, idx, steps in bucket_ops(ops)]
        for bk in self.buckets:
            if bk.kind == "rotate":
                bk.tables = CKKS.rotation_tables([self.gk[(bk.N, st)] for st in bk.steps], self.ck[bk.N].device)
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self._ws: Dict[Tuple[int, int], torch.Tensor] = {}
        self.assign_streams()

    # -- data -------------------------------------------------------------
    def fill_random(self, seed: int = 1) -> None:
        """Encrypt uniform slots in [-1, 1] into every bucket's inputs."""
        rng = np.random.default_rng(seed)
        for bk in self.buckets:
            ck, sk, l, B = self.ck[bk.N], self.sk[bk.N], bk.limbs, bk.batch
            z = rng.uniform(-1, 1, (B, bk.N // 2))
            bk.a = ck.encrypt(ck.encode(z, level=l), sk).data
            if bk.kind == "hmult":
                bk.b = ck.encrypt(ck.encode(rng.uniform(-1, 1, (B, bk.N // 2)), level=l), sk).data
                bk.out = torch.empty(B, 2, l - 1, bk.N, dtype=torch.uint64, device=ck.device)
            elif bk.kind == "rotate":
                bk.out = torch.empty_like(bk.a)
            else:
                pt = ck.encode(rng.uniform(-1, 1, (B, bk.N // 2)), level=l)
                bk.b = pt.unsqueeze(1).expand(B, 2, l, bk.N).contiguous().view(B * 2, l, bk.N)
                bk.out = torch.empty_like(bk.a)

    def _workspace(self, N: int, stream: int = 0) -> torch.Tensor:
        """One workspace per (ring degree, stream): buckets that may run at the
        same time never share one."""
        if (N, stream) not in self._ws:
            ck = self.ck[N]
            need = max(call("ckks_workspace_bytes", ck.ring.handle, bk.batch, bk.limbs)
                       for bk in self.buckets if bk.N == N)
            self._ws[(N, stream)] = torch.empty((need + 7) // 8, dtype=torch.uint64, device=ck.device)
        return self._ws[(N, stream)]

    def assign_streams(self) -> None:
        """Longest-first greedy placement of the buckets on the streams, by
        tuned time (or, before tuning, by an estimate of the work)."""
        cost = lambda bk: bk.ms if bk.ms > 0 else bk.batch * bk.N * bk.limbs ** 2
        for bk, k in zip(self.buckets, place([cost(bk) for bk in self.buckets], self.n_streams)):
            bk.stream = k

    # -- launches ---------------------------------------------------------
    def _launch(self, bk: Bucket, stream: int = 0) -> None:
        ck = self.ck[bk.N]
        r = ck.ring
        n, k, t = bk.shape
        r.set_threads(n, k, t)
        l, B, K = bk.limbs, bk.batch, ck.L + 1
        ws = self._workspace(bk.N, stream)
        if bk.kind == "hmult":
            call("ckks_hmult", r.handle, _ptr(bk.a), _ptr(bk.b), _ptr(self.rlk[bk.N].prepared), K, _ptr(bk.out), B,
                 l, ck.sp, _ptr(ws), r.stream())
        elif bk.kind == "rotate":
            keys, gal = bk.tables
            call("ckks_rotate_many", r.handle, _ptr(bk.a), _ptr(keys), K, _ptr(gal), _ptr(bk.out), B, l, ck.sp,
                 _ptr(ws), r.stream())
        else:
            call("ckks_poly_mul", r.handle, _ptr(bk.a), _ptr(bk.b), _ptr(bk.out), B * 2, l, 0, r.stream())

    def _side_streams(self) -> List[torch.cuda.Stream]:
        if len(self._streams) != self.n_streams:
            dev = self.ck[next(iter(self.ck))].device
            self._streams = [torch.cuda.Stream(dev) for _ in range(self.n_streams)]   # default priority
        return self._streams

    def run_eager(self) -> None:
        """Every bucket on its stream: the streams fork from the current
        stream and join back into it."""
        if self.n_streams == 1:
            for bk in self.buckets:
                self._launch(bk)
            return
        cur = torch.cuda.current_stream()
        side = self._side_streams()
        for s in side:
            s.wait_stream(cur)
        for bk in self.buckets:
            with torch.cuda.stream(side[bk.stream]):
                self._launch(bk, bk.stream)
        for s in side:
            cur.wait_stream(s)

    # -- tuning -----------------------------------------------------------
    def tune(self, reps: int = 3) -> None:
        """Launch shape per bucket: the fastest built thread count of each
        kernel family the bucket's operation uses, timed on the bucket."""
        from .tune import BUILT, BUILT_FUSED, _time
        for bk in self.buckets:
            logN = bk.N.bit_length() - 1
            if bk.kind == "hmult":
                grid = itertools.product(BUILT[logN], BUILT_FUSED[logN], BUILT_FUSED[logN])
                cands = [(n, k, t) for n, t, k in grid]
            elif bk.kind == "rotate":
                cands = [(0, k, 0) for k in BUILT_FUSED[logN]]
            else:
                cands = [(n, 0, 0) for n in BUILT[logN]]
            best = None
            for c in cands:
                bk.shape = c
                ms = _time(lambda: self._launch(bk), reps)
                if best is None or ms < best[1]:
                    best = (c, ms)
            bk.shape, bk.ms = best
        self.assign_streams()

    # -- graph ------------------------------------------------------------
    def capture(self) -> torch.cuda.CUDAGraph:
        """Record every bucket's launches into one HIP graph."""
        for N in self.ck:
            for k in range(self.n_streams):
                self._workspace(N, k)
        self._side_streams()
        s = torch.cuda.Stream(self.ck[next(iter(self.ck))].device)
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):          # warm-up outside capture
            self.run_eager()
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.run_eager()
        self.graph = g
        return g

    def replay(self) -> None:
        self.graph.replay()

    def outputs(self) -> List[torch.Tensor]:
        return [bk.out for bk in self.buckets]

    def summary(self) -> List[dict]:
        return [{"N": bk.N, "level": bk.level, "kind": bk.kind, "steps": sorted(set(bk.steps)), "batch": bk.batch,
                 "threads": dict(zip(("ntt", "ks", "tensor"), bk.shape)), "tuned_ms": round(bk.ms, 4)}
                for bk in self.buckets]
