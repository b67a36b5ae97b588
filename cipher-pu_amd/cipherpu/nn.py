"""Encrypted MLP 784-128-10 on the CKKS core (x^2 activation, DARE refresh,
block-sparsity gears), with the diagonal matrix-vector method.

Packing.  N = 8192 gives 4096 slots; a ciphertext carries S = 4 samples in
1024-slot blocks.  In each block the 784 pixels sit in slots [0, 784),
zeros in [784, 896) and a copy of slots [0, 128) in [896, 1024), so that
every rotation by k < 128 used below stays inside the sample's block.

Layer 1 (W1: 128 x 784, padded to 128 x 896).  Hybrid diagonal method:
    z = sum_(k<128) d_k * rot(x, k),   d_k[i] = W1[i mod 128][(i + k) mod 896]
for i < 896 (0 on [896, 1024)), then y_j = sum_(t<8) z[j + 128 t] by
rotate-and-sum over 512, 256, 128.  k = 16 g + b is evaluated baby-step /
giant-step: z = sum_g rot(sum_b rot(d_(16g+b), -16g) * rot(x, b), 16 g), so
only 15 + 7 rotations are needed, the 15 baby steps as one hoisted call
(ckks_rotate_hoisted: one shared digit decomposition); the plaintext
products run in the evaluation domain (ckks_lincomb).  The reference computes the same layer
as coefficient-encoded dot products (scripts/exp_e11_ckks_e2e_real.py:126-166).

Then: mask slots [0, 128) of each block (the mask's scale is chosen so that
the rescale returns the ciphertext to the base scale 2^40), copy the 128
outputs to [128, 256) so that layer 2's rotations wrap, DARE refresh (decrypt, keep the scale-carrying integer message, encrypt
afresh at the top level; exp_e24_rotation_bootstrap.py:287-304), x^2 by
HMult + relinearise + rescale, and layer 2 (W2: 10 x 128 padded to 16 x 128)
with 4 x 4 baby/giant steps and folds over 64, 32, 16.

Gears (block sparsity, verify_m4_a_gear_ctrl.py:41-57, exp_e10:102-124):
gear g in {0, 1, 2, 3} zeroes {0, 25, 50, 75} % of the 64-weight blocks
(row-major runs of the flattened [out][in] weight matrix) with the smallest
max |w| -- the quantile form of the reference's max|w| < threshold rule.
The pruned matrix runs through the same diagonal plan; a diagonal that
pruning empties is skipped (no plaintext product; its baby rotation is
skipped when no giant step needs it).  Under diagonal packing the cost is
set by the rotations, which fine-grained pruning rarely removes, so the
gears change the output SNR far more than the cost.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from ._lib import call
from .ckks import CKKS, Ciphertext, SecretKey, SwitchKey
from .ring import to_numpy

SLOTS_PER_SAMPLE = 1024
GEAR_FRACTIONS = (0.0, 0.25, 0.50, 0.75)
BLOCK = 64


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


# ---------------------------------------------------------------- plaintext side

def synthetic_images(n: int, seed: int = 0) -> np.ndarray:
    """n synthetic 28x28 images, uint8 uniform in [0, 255], scaled to [0, 1]."""
    rng = np.random.default_rng(seed)
    return rng.integers(0, 256, size=(n, 784), dtype=np.uint8).astype(np.float64) / 255.0


def make_mlp(seed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """784-128-10 weights ~ N(0, 1/fan_in), FP64, no biases."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0.0, 1.0 / math.sqrt(784), (128, 784))
    W2 = rng.normal(0.0, 1.0 / math.sqrt(128), (10, 128))
    return W1, W2


# Weight families for training the gear controller across networks: the
# pruning error of a gear, hence the output SNR the monitor measures,
# depends on how a network's weight energy is spread over its 64-weight
# blocks.  Every family is rescaled to the variance 1 / fan_in of the
# bench's Gaussian networks.
WEIGHT_FAMILIES = ("gaussian", "laplace", "student3", "rowscaled", "blocksparse")


def family_weights(shape, fan_in: int, family: str, rng: np.random.Generator) -> np.ndarray:
    """Weights of `family`, variance 1 / fan_in:
      gaussian     N(0, 1);
      laplace      Laplace(0, 1) (heavier tails);
      student3     Student t with 3 degrees of freedom (heavy tails);
      rowscaled    N(0, 1) rows scaled by exp(N(0, 1)) (uneven output units);
      blocksparse  N(0, 1) with every other 64-weight block of the flattened
                   matrix zeroed at random (probability 1/2)."""
    shape = tuple(shape)
    if family == "gaussian":
        w = rng.normal(size=shape)
    elif family == "laplace":
        w = rng.laplace(size=shape)
    elif family == "student3":
        w = rng.standard_t(3, size=shape)
    elif family == "rowscaled":
        w = rng.normal(size=shape) * np.exp(rng.normal(size=(shape[0],) + (1,) * (len(shape) - 1)))
    elif family == "blocksparse":
        flat = rng.normal(size=int(np.prod(shape)))
        nb = -(-flat.size // BLOCK)
        keep = np.repeat(rng.random(nb) < 0.5, BLOCK)[: flat.size]
        keep[:BLOCK] = True                                   # never all zero
        w = (flat * keep).reshape(shape)
    else:
        raise ValueError(f"unknown weight family {family}")
    return w / (w.std() * math.sqrt(fan_in))


def make_mlp_family(seed: int) -> Tuple[np.ndarray, np.ndarray]:
    """784-128-10 weights of family WEIGHT_FAMILIES[seed % 5], seeded."""
    rng = np.random.default_rng(10_000 + seed)
    fam = WEIGHT_FAMILIES[seed % len(WEIGHT_FAMILIES)]
    return family_weights((128, 784), 784, fam, rng), family_weights((10, 128), 128, fam, rng)


def weight_snr_db(W: np.ndarray, gear: int, cap: float = 60.0) -> float:
    """10 log10(sum W^2 / sum (W - W_gear)^2) of one weight matrix under a
    gear's block pruning, capped (a lossless gear reads `cap`)."""
    W = np.asarray(W, dtype=np.float64)
    lost = float(np.sum((W * ~block_keep_mask(W, GEAR_FRACTIONS[gear])) ** 2))
    tot = float(np.sum(W ** 2))
    if lost <= tot * 10.0 ** (-cap / 10):
        return cap
    return min(cap, 10.0 * math.log10(tot / lost))


def plain_forward(W1, W2, X) -> np.ndarray:
    return (W2 @ ((W1 @ X.T) ** 2)).T


def extended_diagonals(W: np.ndarray, rows: int, cols: int) -> np.ndarray:
    """d_k[i] = W[i mod rows][(i + k) mod cols] for k < rows, i < cols
    (W zero-padded to rows x cols)."""
    P = np.zeros((rows, cols))
    P[: W.shape[0], : W.shape[1]] = W
    i = np.arange(cols)
    return np.stack([P[i % rows, (i + k) % cols] for k in range(rows)])


def block_keep_mask(W: np.ndarray, frac: float, block: int = BLOCK) -> np.ndarray:
    """Boolean keep-mask of W: the `frac` share of `block`-weight blocks
    (row-major runs of the flattened matrix, the last one short) with the
    smallest max |w| is dropped; ties keep the earlier block
    (scripts/exp_e10_sparsity_sweep_v2.py:102-124, quantile form)."""
    flat = np.abs(np.asarray(W, dtype=np.float64)).ravel()
    nb = -(-flat.size // block)
    padded = np.zeros(nb * block)
    padded[: flat.size] = flat
    bmax = padded.reshape(nb, block).max(axis=1)
    k = int(round(frac * nb))
    keep = np.ones(nb, dtype=bool)
    if k:
        keep[np.argsort(bmax, kind="stable")[:k]] = False
    return np.repeat(keep, block)[: flat.size].reshape(np.shape(W))


def apply_block_sparse_mask(weight_matrix, threshold_val, B):
    """The reference's threshold rule: every full run of B flattened weights
    whose largest |w| is below threshold_val is zeroed; a ragged tail run is
    never pruned.  Returns (masked weights, skipped blocks, full blocks,
    kept share) (scripts/exp_e10_sparsity_sweep_v2.py:102-124)."""
    flat = np.array(weight_matrix, dtype=np.float64).reshape(-1)
    nb = flat.size // B
    drop = np.abs(flat[: nb * B]).reshape(nb, B).max(axis=1) < threshold_val
    out = flat.copy()
    out[: nb * B].reshape(nb, B)[drop] = 0.0
    skipped = int(drop.sum())
    return out.reshape(np.shape(weight_matrix)), skipped, nb, (nb - skipped) / nb


def block_sparse_mask(weight_matrix, threshold_frac, B):
    """The same rule with the threshold threshold_frac times the largest
    |w|; returns (masked weights, skipped blocks, full blocks)
    (scripts/exp_e7_sparsity_sweep.py:72-90)."""
    t = np.abs(np.asarray(weight_matrix, dtype=np.float64)).max() * threshold_frac
    masked, skipped, nb, _ = apply_block_sparse_mask(weight_matrix, t, B)
    return masked, skipped, nb


def weights_from_diagonals(D: np.ndarray, rows: int, cols: int, shape) -> np.ndarray:
    """Inverse of extended_diagonals (to evaluate pruned networks in FP64)."""
    P = np.zeros((rows, cols))
    i = np.arange(cols)
    for k in range(rows):
        P[i % rows, (i + k) % cols] = D[k]
    return P[: shape[0], : shape[1]]


# ---------------------------------------------------------------- encrypted layers

@dataclass
class LayerPlan:
    rows: int            # output dimension (padded)
    cols: int            # input span of the diagonals (padded)
    baby: int
    giant: int
    folds: Tuple[int, ...]
    D: torch.Tensor = None          # [ng][nb][l][N][2] prepared plaintext rows
    used_b: List[int] = field(default_factory=list)
    used_g: List[int] = field(default_factory=list)
    nnz_diagonals: int = 0
    _lower: Dict[int, torch.Tensor] = field(default_factory=dict)

    def rows_at(self, l: int) -> torch.Tensor:
        """D for a ciphertext at level l <= D's: its first l limbs (an
        encoding's residues do not depend on the primes after them)."""
        if l == self.D.shape[2]:
            return self.D
        if l > self.D.shape[2]:
            raise ValueError(f"plan prepared at level {self.D.shape[2]}, ciphertext at {l}")
        if l not in self._lower:
            self._lower[l] = self.D[:, :, :l].contiguous()
        return self._lower[l]


class EncryptedNet:
    """What the encrypted networks share: keys, rotations, the DARE refresh,
    the slot pattern of the S samples a ciphertext carries, and the
    interface the gear loops drive (set_gears, prepare, encrypt_inputs,
    forward, decrypt_outputs, reference_outputs).

    Levels.  Inputs are encrypted at the top level L and the DARE refresh
    lifts to L.  With drop_levels, each of forward's two stages (before and
    after the refresh) starts at the lowest level it can run at, its
    rescales + 1 (stage_levels): the refresh and the final decryption read
    limb 0 alone, so a stage needs to end at level 1, no higher.  The input's
    unused limbs are dropped (modulus reduction, no division) and the
    refresh encrypts at the second stage's level; each is the top-level
    ciphertext's residues on fewer primes.  The gear features are taken on
    the input ciphertext before the drop."""

    stage_levels = (3, 3)   # the MLP: linear + mask, then x^2 + linear

    def __init__(self, ck: CKKS, sk: SecretKey, w_scale: float = 2.0 ** 40):
        self.ck, self.sk = ck, sk
        self.w_scale = w_scale
        self.S = (ck.N // 2) // SLOTS_PER_SAMPLE
        self.rlk = ck.relin_key(sk)
        self._gk: Dict[int, SwitchKey] = {}
        self._hkeys: Dict[tuple, torch.Tensor] = {}
        self.hoist = True          # baby steps / convolution taps as hoisted rotations
        self.eval_rotations = True  # baby steps leave the hoisted call in evaluation form (MLP)
        self.mixed_rotations = True  # rotations of different rows by different steps in one call
        self._rtab: Dict[tuple, tuple] = {}
        self.ops = {"rot": 0, "ptmul_rows": 0, "hmult": 0, "dare": 0}
        self.gears = (0, 0)
        self.fp64 = ck.ring.fp64    # prepared plaintexts and keys are in this arithmetic
        self.drop_levels = True     # each stage at the lowest level it needs
        self._ones: Dict[tuple, torch.Tensor] = {}   # rot_sum weights, per (blocks, level, arithmetic)

    def entry_levels(self) -> Tuple[int, int]:
        """Levels at which forward's two stages start (L, L without drop_levels)."""
        L = self.ck.L
        return tuple(min(L, l) for l in self.stage_levels) if self.drop_levels else (L, L)

    def enter(self, ct: Ciphertext, l: int) -> Ciphertext:
        """ct at level l: the limbs above l dropped (ct itself when at or below)."""
        return self.ck.drop_levels(ct, l) if ct.level > l else ct

    def check_arith(self):
        if self.ck.ring.fp64 != self.fp64:
            raise ValueError("the context's arithmetic changed since this network prepared its keys and "
                             "plaintexts: build the network again")

    # -- keys -------------------------------------------------------------
    def gk(self, step: int) -> SwitchKey:
        step %= self.ck.N // 2
        if step not in self._gk:
            self._gk[step] = self.ck.galois_key(self.sk, step)
        return self._gk[step]

    def rot(self, ct: Ciphertext, step: int) -> Ciphertext:
        if step % (self.ck.N // 2) == 0:
            return ct
        self.ops["rot"] += ct.batch
        return self.ck.rotate(ct, self.gk(step))

    def rot_add(self, ct: Ciphertext, step: int) -> Ciphertext:
        """ct + rot(ct, step), one call (ckks_rotate_add)."""
        if step % (self.ck.N // 2) == 0:
            return self.ck.add(ct, ct)
        self.ops["rot"] += ct.batch
        return self.ck.rotate_add(ct, self.gk(step))

    def rot_sum(self, ct: Ciphertext, B: int, steps: List[int]) -> Ciphertext:
        """sum_i rot(block i, steps[i]) over the len(steps) blocks of B rows
        of ct: the moving blocks rotated by one ckks_rotate_many call into a
        stacked buffer, summed by one ckks_lincomb_scalar call with weights
        1 (the words of rot_blocks followed by point-wise sums)."""
        ck = self.ck
        half = ck.N // 2
        k = len(steps)
        moving = [i for i, st in enumerate(steps) if st % half]
        m0 = moving[0] if moving else 0
        if k == 1 or not self.mixed_rotations or moving != list(range(m0, m0 + len(moving))):
            blocks = self.rot_blocks(ct, B, steps)
            z = blocks[0]
            for part in blocks[1:]:
                z = ck.add(z, part)
            return z
        l = ct.level
        buf = torch.empty_like(ct.data)
        for i in range(k):
            if i not in moving:
                buf[i * B:(i + 1) * B].copy_(ct.data[i * B:(i + 1) * B])
        key = tuple(steps[i] % half for i in moving for _ in range(B))
        gks = [self.gk(st) for st in key]
        if key not in self._rtab:
            self._rtab[key] = ck.rotation_tables(gks, ck.device)
        self.ops["rot"] += len(key)
        lo, hi = m0 * B, (m0 + len(moving)) * B
        ck.rotate_many(Ciphertext(ct.data[lo:hi], ct.scale), gks, tables=self._rtab[key], out=buf[lo:hi])
        ones = self._ones.get((k, l, ck.ring.fp64))
        if ones is None:
            W = np.zeros((1, k, l, 2), dtype=np.uint64)
            for i in range(l):
                W[0, :, i] = ck.ring.prepared_pairs(1, ck.primes[i])
            ones = self._ones[(k, l, ck.ring.fp64)] = ck.ring.tensor(W)
        out = torch.empty(B, 2, l, ck.N, dtype=torch.uint64, device=ck.device)
        call("ckks_lincomb_scalar", ck.ring.handle, _ptr(buf), _ptr(ones), _ptr(out), k, 1, B, l, ck.ring.stream())
        return Ciphertext(out, ct.scale)

    def rot_blocks(self, ct: Ciphertext, B: int, steps: List[int]) -> List[Ciphertext]:
        """ct holds len(steps) blocks of B rows; block i rotated by steps[i]
        (a multiple of N/2 leaves it as it is).  The blocks that move go
        through one ckks_rotate_many call, each row with its block's Galois
        key, or through one rotation call per distinct step when
        mixed_rotations is off; same words either way.  Returned blocks are
        views of ct or of the rotation output (nothing is copied back)."""
        ck = self.ck
        half = ck.N // 2
        blocks = [ct.data[i * B:(i + 1) * B] for i in range(len(steps))]
        moving = [i for i, st in enumerate(steps) if st % half]
        if moving:
            m0 = moving[0]
            if moving == list(range(m0, m0 + len(moving))):
                src = ct.data[m0 * B:(m0 + len(moving)) * B]
            else:
                src = torch.cat([blocks[i] for i in moving])
            if self.mixed_rotations:
                key = tuple(steps[i] % half for i in moving for _ in range(B))
                gks = [self.gk(st) for st in key]
                if key not in self._rtab:
                    self._rtab[key] = ck.rotation_tables(gks, ck.device)   # the keys stay alive in self._gk
                self.ops["rot"] += len(key)
                moved = ck.rotate_many(Ciphertext(src, ct.scale), gks, tables=self._rtab[key]).data
                for j, i in enumerate(moving):
                    blocks[i] = moved[j * B:(j + 1) * B]
            else:
                for j, i in enumerate(moving):
                    blocks[i] = self.rot(Ciphertext(src[j * B:(j + 1) * B], ct.scale), steps[i]).data
        return [Ciphertext(b, ct.scale) for b in blocks]

    def _slots(self, v: np.ndarray) -> np.ndarray:
        """A 1024-slot block pattern repeated for the S samples of a ciphertext."""
        return np.tile(v, self.S)

    def rot_stack(self, ct: Ciphertext, steps: List[int], eval_form: bool = False) -> torch.Tensor:
        """[len(steps), B, 2, l, N]: ct rotated by every step (0 = ct itself),
        the nonzero steps as one hoisted call sharing a single digit
        decomposition (up to 32 a call); in evaluation order when eval_form."""
        ck = self.ck
        out = torch.empty(len(steps), ct.batch, 2, ct.level, ck.N, dtype=torch.uint64, device=ck.device)
        nz = [i for i, s in enumerate(steps) if s % (ck.N // 2)]
        for i, s in enumerate(steps):
            if i not in nz:
                out[i].copy_(ct.data)
                if eval_form:
                    ck.ring.ntt_(out[i].view(-1, ct.level, ck.N))
        if not self.hoist:
            for i in nz:
                out[i].copy_(self.rot(ct, steps[i]).data)
                if eval_form:
                    ck.ring.ntt_(out[i].view(-1, ct.level, ck.N))
            return out
        for c0 in range(0, len(nz), 32):
            idx = nz[c0:c0 + 32]
            key = tuple(steps[i] for i in idx)
            if key not in self._hkeys:
                self._hkeys[key] = torch.stack([ck.hoist_key(self.gk(s)) for s in key]).contiguous()
            gks = [self.gk(s) for s in key]
            if idx == list(range(idx[0], idx[0] + len(idx))):      # contiguous rows: write in place
                ck.rotate_hoisted(ct, gks, self._hkeys[key], out=out[idx[0]:idx[0] + len(idx)], eval_form=eval_form)
            else:
                for i, r in zip(idx, ck.rotate_hoisted(ct, gks, self._hkeys[key], eval_form=eval_form)):
                    out[i].copy_(r.data)
            self.ops["rot"] += ct.batch * len(idx)
        return out

    def rotation_steps(self) -> List[int]:
        raise NotImplementedError

    def set_gears(self, gear1: int, gear2: int):
        raise NotImplementedError

    def prepare(self, gears=range(4)) -> None:
        """Build every gear's plans and every Galois key they need up front,
        so that gear switches inside a timed loop cost no set-up."""
        cur = self.gears
        for g in gears:
            self.set_gears(g, g)
            for s in self.rotation_steps():
                self.gk(s)
        self.set_gears(*cur)

    def dare(self, ct: Ciphertext, level: Optional[int] = None) -> Ciphertext:
        """Decrypt, keep the scale-carrying integer message (centred limb 0),
        encrypt afresh at the top level, or at `level`: the top-level
        ciphertext's first `level` limbs (exp_e24_rotation_bootstrap.py:287-304)."""
        ck = self.ck
        lo = ck.L if level is None else level
        pt = ck.decrypt(ct, self.sk)                                            # [B, l, N]
        B, l = pt.shape[0], pt.shape[1]
        top = torch.empty(B, lo, ck.N, dtype=torch.uint64, device=ck.device)
        call("ckks_lift_small", ck.ring.handle, _ptr(pt), _ptr(top), B, l, lo, ck.ring.stream())
        self.ops["dare"] += B
        return ck.encrypt(top, self.sk, scale=ct.scale)

    def prepared_pt(self, z: np.ndarray, level: int, scale: float) -> torch.Tensor:
        """Slot vectors [rows, N/2] -> evaluation-form prepared pairs
        (ring.prepared_pairs) [rows, level, N, 2] for ckks_lincomb."""
        ck = self.ck
        pt = ck.encode(np.asarray(z, dtype=np.complex128), level=level, scale=scale)
        prep = torch.empty(*pt.shape, 2, dtype=torch.uint64, device=ck.device)
        call("ckks_pt_prepare", ck.ring.handle, _ptr(pt), _ptr(prep), pt.shape[0], level, 0, ck.ring.stream())
        return prep


class EncryptedMLP(EncryptedNet):
    def __init__(self, ck: CKKS, sk: SecretKey, W1: np.ndarray, W2: np.ndarray, gear1: int = 0, gear2: int = 0,
                 w_scale: float = 2.0 ** 40):
        super().__init__(ck, sk, w_scale)
        self.W1, self.W2 = W1, W2
        self._masks = {}
        self._plan1: Dict[int, tuple] = {}
        self._plan2: Dict[int, tuple] = {}
        self.set_gears(gear1, gear2)

    def reference_outputs(self, X: np.ndarray) -> np.ndarray:
        """The dense FP64 network: the monitor's reference."""
        return plain_forward(self.W1, self.W2, X)

    def weight_matrices(self):
        """The two weight matrices as the gears prune them."""
        return self.W1, self.W2

    def plain_outputs_at_gear(self, X: np.ndarray, gear: int) -> np.ndarray:
        """The FP64 network with both layers pruned as `gear` prunes them."""
        return plain_forward(self.W1 * block_keep_mask(self.W1, GEAR_FRACTIONS[gear]),
                             self.W2 * block_keep_mask(self.W2, GEAR_FRACTIONS[gear]), X)

    # -- plans --------------------------------------------------------------
    def _plan(self, D: np.ndarray, rows, cols, baby, giant, folds, level) -> LayerPlan:
        ck = self.ck
        p = LayerPlan(rows, cols, baby, giant, folds)
        nz = [bool(np.any(D[k])) for k in range(rows)]
        p.nnz_diagonals = int(sum(nz))
        p.used_b = [b for b in range(baby) if any(nz[g * baby + b] for g in range(giant) if g * baby + b < rows)]
        p.used_g = [g for g in range(giant) if any(nz[g * baby + b] for b in range(baby) if g * baby + b < rows)]
        if not p.used_g:
            raise ValueError("every diagonal of the layer is pruned")
        rows_pt = []
        for g in p.used_g:
            for b in p.used_b:
                k = g * baby + b
                blk = np.zeros(SLOTS_PER_SAMPLE)
                if k < rows:
                    blk[:cols] = D[k]
                # pre-rotate by -baby*g so that the giant rotation aligns it
                v = np.roll(self._slots(blk), baby * g)
                rows_pt.append(v)
        z = np.asarray(rows_pt, dtype=np.complex128)
        pt = ck.encode(z, level=level, scale=self.w_scale)                  # [ng*nb, l, N]
        prepared = torch.empty(*pt.shape, 2, dtype=torch.uint64, device=ck.device)
        call("ckks_pt_prepare", ck.ring.handle, _ptr(pt), _ptr(prepared), pt.shape[0], level, 0, ck.ring.stream())
        p.D = prepared.view(len(p.used_g), len(p.used_b), level, ck.N, 2)
        return p

    def set_gears(self, gear1: int, gear2: int):
        """Select the layer gears; the pruned weights and prepared diagonal
        plans of each gear are built once and cached."""
        L = self.ck.L
        if gear1 not in self._plan1:
            P1 = self.W1 * block_keep_mask(self.W1, GEAR_FRACTIONS[gear1])
            self._plan1[gear1] = (P1, self._plan(extended_diagonals(P1, 128, 896), 128, 896, 16, 8,
                                                 (512, 256, 128), L))
        if gear2 not in self._plan2:
            P2 = self.W2 * block_keep_mask(self.W2, GEAR_FRACTIONS[gear2])
            self._plan2[gear2] = (P2, self._plan(extended_diagonals(P2, 16, 128), 16, 128, 4, 4, (64, 32, 16),
                                                 L - 1))
        self.gears = (gear1, gear2)
        (self.P1, self.p1), (self.P2, self.p2) = self._plan1[gear1], self._plan2[gear2]

    def pruned_weights(self):
        return self.P1, self.P2

    def rotation_steps(self) -> List[int]:
        """Rotation steps the current plans use (baby, giant, fold, replicate)."""
        steps = {-128}
        for p in (self.p1, self.p2):
            steps.update(b for b in p.used_b if b)
            steps.update(p.baby * g for g in p.used_g if g)
            steps.update(p.folds)
        return sorted(steps)

    # -- evaluation ---------------------------------------------------------
    def encrypt_inputs(self, X: np.ndarray) -> Ciphertext:
        ck = self.ck
        n = X.shape[0]
        nct = -(-n // self.S)
        z = np.zeros((nct, ck.N // 2))
        for s in range(n):
            c, j = divmod(s, self.S)
            blk = np.zeros(SLOTS_PER_SAMPLE)
            blk[:784] = X[s]
            blk[896:1024] = blk[:128]
            z[c, j * SLOTS_PER_SAMPLE:(j + 1) * SLOTS_PER_SAMPLE] = blk
        return ck.encrypt(ck.encode(z), self.sk)

    def _linear(self, ct: Ciphertext, p: LayerPlan) -> Ciphertext:
        ck = self.ck
        B, l = ct.batch, ct.level
        E = self.rot_stack(ct, p.used_b, eval_form=self.eval_rotations)         # [nb][B][2][l][N]
        if not self.eval_rotations:
            ck.ring.ntt_(E.view(-1, l, ck.N))
        nb, ng = len(p.used_b), len(p.used_g)
        acc = torch.empty(ng, B, 2, l, ck.N, dtype=torch.uint64, device=ck.device)
        call("ckks_lincomb", ck.ring.handle, _ptr(E), _ptr(p.rows_at(l)), _ptr(acc), nb, ng, B, l, ck.ring.stream())
        self.ops["ptmul_rows"] += nb * ng * B
        ck.ring.intt_(acc.view(-1, l, ck.N))
        # every giant step rescaled and rotated at once, each by its own step
        parts = ck.rescale(Ciphertext(acc.view(ng * B, 2, l, ck.N), ct.scale * self.w_scale))
        z = self.rot_sum(parts, B, [p.baby * g for g in p.used_g])
        for s in p.folds:
            z = self.rot_add(z, s)
        return z

    def _mask(self, ct: Ciphertext, keep: int, in_place: bool = False) -> Ciphertext:
        """Multiply by the 0/1 slot mask [0, keep) of each block, encoded at
        scale q_(l-1) * delta / scale(ct) so that the rescale brings the
        ciphertext back to the base scale delta (the DARE refresh keeps the
        scale, and the activation and layer 2 need its headroom: at a lower
        scale each later rotation's key-switch noise, about 700 per
        coefficient, would reach 1e-2 in the decoded slots).  in_place: ct's
        words are transformed in place (the caller drops ct)."""
        ck = self.ck
        l = ct.level
        m_scale = float(ck.primes[l - 1]) * ck.p.scale / ct.scale
        key = (keep, l, m_scale)
        if key not in self._masks:
            blk = np.zeros(SLOTS_PER_SAMPLE)
            blk[:keep] = 1.0
            pt = ck.encode(self._slots(blk)[None].astype(np.complex128), level=l, scale=m_scale)
            prep = torch.empty(*pt.shape, 2, dtype=torch.uint64, device=ck.device)
            call("ckks_pt_prepare", ck.ring.handle, _ptr(pt), _ptr(prep), 1, l, 0, ck.ring.stream())
            self._masks[key] = prep
        E = ct.data if in_place else ct.data.clone()
        ck.ring.ntt_(E.view(-1, l, ck.N))
        out = torch.empty_like(E)
        call("ckks_lincomb", ck.ring.handle, _ptr(E), _ptr(self._masks[key]), _ptr(out), 1, 1, ct.batch, l,
             ck.ring.stream())
        ck.ring.intt_(out.view(-1, l, ck.N))
        return ck.rescale(Ciphertext(out, ct.scale * m_scale))

    def layer1(self, ct: Ciphertext) -> Ciphertext:
        y = self._linear(ct, self.p1)                 # outputs in slots [0, 128) of each block
        y = self._mask(y, 128, in_place=True)
        return self.rot_add(y, -128)                  # copy to [128, 256) for layer 2's wrap

    def activation(self, ct: Ciphertext) -> Ciphertext:
        self.ops["hmult"] += ct.batch
        return self.ck.hmult(ct, ct, self.rlk)

    def layer2(self, ct: Ciphertext) -> Ciphertext:
        return self._linear(ct, self.p2)

    def forward(self, ct: Ciphertext) -> Ciphertext:
        self.check_arith()
        l1, l2 = self.entry_levels()
        h = self.dare(self.layer1(self.enter(ct, l1)), level=l2)
        return self.layer2(self.activation(h))

    def decrypt_outputs(self, ct: Ciphertext, n: int) -> np.ndarray:
        ck = self.ck
        z = to_numpy(ck.decode(ck.decrypt(ct, self.sk), ct.scale)).real
        out = np.zeros((n, 10))
        for s in range(n):
            c, j = divmod(s, self.S)
            out[s] = z[c, j * SLOTS_PER_SAMPLE: j * SLOTS_PER_SAMPLE + 10]
        return out


def output_snr_db(ref: np.ndarray, got: np.ndarray) -> float:
    """10 log10(sum ref^2 / sum (got - ref)^2)
    (verify_m4h2_prod_calibration.py:64-69)."""
    noise = float(np.sum((got - ref) ** 2))
    return math.inf if noise == 0 else 10.0 * math.log10(float(np.sum(ref ** 2)) / noise)
