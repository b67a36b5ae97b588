"""Encrypted CNN on the CKKS core: conv 5x5, 8 channels, stride 2 on 28x28
-> x^2 -> 2x2 average pool -> FC 288-10, with the DARE refresh after the
convolution and the same block-sparsity gears as the MLP.

Packing (4 samples a ciphertext, 1024 slots each).  The image is stored as
its four polyphase components X_(pr,pc)[r][s] = x[2r + pr][2s + pc]
(r, s < 14), component p = 2 pr + pc at slots 256 p + 16 r + s.  Then the
stride-2 correlation at output (i, j) needs X_(u%2, v%2)[i + u//2][j + v//2],
which rotation by 256 p + 16 (u//2) + v//2 brings to slot 16 i + j: one
rotation per kernel tap, shared by all 8 channels, all 24 of them one
hoisted call (ckks_rotate_hoisted).  Channel c is
    y_c = sum_(u,v) K[c][u][v] * rot(x, s_uv)
with integer weights, one ckks_lincomb_scalar call for all channels (in the
coefficient domain: a constant multiplies every coefficient alike).  The
reference computes its convolution as coefficient-encoded dot products, one
decryption per output (scripts/exp_e14_cnn_e2e.py:90-127).

After DARE and x^2 (one batched HMult over the 8 channel ciphertexts) come
the pool and the FC layer.  FC output k sums F[k][36 c + 6 a + b] / 4 times
the 2x2 window of channel c at rows 2a, 2a + 1 and columns 2b, 2b + 1 of
the output grid, slots 32 a + 2 b + {0, 1, 16, 17}.  The FC result is summed
over all slots of a 256-slot window in the end, so the pool needs no
rotation: W_(k,c) holds the weight on all four slots of each window, and
output k is sum_c W_(k,c) * A_c, one evaluation-domain ckks_lincomb call for
all 10 outputs.  (Pooling by (1 + rot_1)(1 + rot_16) first, with the weight
on slot 32 a + 2 b alone, gives the same slot sums after two more rotations
of all 8 channel ciphertexts.)  Outputs 4p..4p+3 are shifted into the four
256-slot windows of one ciphertext (all shifts, by three different steps, in
one call) and summed over each window by rotate-and-sum (128 ... 1), every
rotation batched over all the packed ciphertexts.
Output k ends on slot 256 (k mod 4) of packed ciphertext k // 4.
"""
from __future__ import annotations

import ctypes
import math
from typing import Dict, List, Tuple

import numpy as np
import torch

from ._lib import call
from .ckks import CKKS, Ciphertext, SecretKey
from .nn import GEAR_FRACTIONS, SLOTS_PER_SAMPLE, EncryptedNet, block_keep_mask
from .ring import to_numpy

ROW = 16          # row stride of the polyphase grids and of the output grid
PHASE = 256       # slots per polyphase component
WINDOW = 256      # FC summation window
N_CH, KS, N_OUT = 8, 5, 10
# Row order of the FC outputs: those shifted into windows 1..3 first, then
# 0, 4, 8 (window 0), so each set is one contiguous run of rows.
FC_ORDER = [k for k in range(N_OUT) if k % 4] + [k for k in range(N_OUT) if k % 4 == 0]


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def make_cnn(seed: int = 0) -> Tuple[np.ndarray, np.ndarray]:
    """Conv weights K [8][5][5] ~ N(0, 1/25), FC weights F [10][288] ~
    N(0, 1/288), FP64, no biases."""
    rng = np.random.default_rng(seed)
    K = rng.normal(0.0, 1.0 / math.sqrt(KS * KS), (N_CH, KS, KS))
    F = rng.normal(0.0, 1.0 / math.sqrt(288), (N_OUT, 288))
    return K, F


def make_cnn_family(seed: int) -> Tuple[np.ndarray, np.ndarray]:
    """Conv [8][5][5] and FC [10][288] weights of family
    WEIGHT_FAMILIES[seed % 5] (nn.family_weights), seeded."""
    from .nn import WEIGHT_FAMILIES, family_weights
    rng = np.random.default_rng(20_000 + seed)
    fam = WEIGHT_FAMILIES[seed % len(WEIGHT_FAMILIES)]
    return (family_weights((N_CH, KS, KS), KS * KS, fam, rng), family_weights((N_OUT, 288), 288, fam, rng))


def cnn_forward(K: np.ndarray, F: np.ndarray, X: np.ndarray) -> np.ndarray:
    """The FP64 network: conv 5x5 stride 2 (no padding) -> x^2 -> 2x2 average
    pool -> FC, FC inputs flattened as [channel][row][col]."""
    n = X.shape[0]
    img = X.reshape(n, 28, 28)
    y = np.zeros((n, N_CH, 12, 12))
    for u in range(KS):
        for v in range(KS):
            y += K[None, :, u, v, None, None] * img[:, None, u:u + 23:2, v:v + 23:2]
    a = y ** 2
    p = 0.25 * (a[:, :, 0::2, 0::2] + a[:, :, 1::2, 0::2] + a[:, :, 0::2, 1::2] + a[:, :, 1::2, 1::2])
    return p.reshape(n, 288) @ F.T


def tap_step(u: int, v: int) -> int:
    """Rotation that aligns kernel tap (u, v) with the output grid."""
    return PHASE * (2 * (u % 2) + (v % 2)) + ROW * (u // 2) + (v // 2)


def fc_weight_blocks(F: np.ndarray) -> np.ndarray:
    """[10 * 8, 1024] per-sample slot blocks W_(k,c), row 8 k + c: F[k][36 c +
    6 a + b] / 4 on slots 32 a + 2 b + {0, 1, 16, 17}, the 2x2 pooling window
    (a, b) of the 12 x 12 output grid (the pool and the FC weight in one)."""
    W = np.zeros((N_OUT * N_CH, SLOTS_PER_SAMPLE))
    for k in range(N_OUT):
        for c in range(N_CH):
            for a in range(6):
                for b in range(6):
                    w = F[k, 36 * c + 6 * a + b] / 4.0
                    for d in (0, 1, ROW, ROW + 1):
                        W[N_CH * k + c, 2 * ROW * a + 2 * b + d] = w
    return W


def pack_images(X: np.ndarray, S: int) -> np.ndarray:
    """[n, 784] images -> [ceil(n / S), S * 1024] polyphase slot vectors."""
    n = X.shape[0]
    nct = -(-n // S)
    z = np.zeros((nct, S * SLOTS_PER_SAMPLE))
    img = X.reshape(n, 28, 28)
    for s in range(n):
        c, j = divmod(s, S)
        base = j * SLOTS_PER_SAMPLE
        for pr in range(2):
            for pc in range(2):
                blk = img[s, pr::2, pc::2]                     # 14 x 14
                off = base + PHASE * (2 * pr + pc)
                for r in range(14):
                    z[c, off + ROW * r: off + ROW * r + 14] = blk[r]
    return z


class EncryptedCNN(EncryptedNet):
    stage_levels = (2, 3)   # conv, then x^2 + FC

    def __init__(self, ck: CKKS, sk: SecretKey, K: np.ndarray, F: np.ndarray, gear1: int = 0, gear2: int = 0,
                 w_scale: float = 2.0 ** 40):
        super().__init__(ck, sk, w_scale)
        self.K, self.F = K, F
        self._conv: Dict[Tuple[int, int], tuple] = {}
        self._fc: Dict[Tuple[int, float], tuple] = {}
        self.set_gears(gear1, gear2)

    def reference_outputs(self, X: np.ndarray) -> np.ndarray:
        return cnn_forward(self.K, self.F, X)

    def weight_matrices(self):
        """The conv weights as [channel][25] and the FC weights, as the gears
        prune them."""
        return self.K.reshape(N_CH, -1), self.F

    def plain_outputs_at_gear(self, X: np.ndarray, gear: int) -> np.ndarray:
        """The FP64 network with both layers pruned as `gear` prunes them."""
        K = self.K * block_keep_mask(self.K.reshape(N_CH, -1), GEAR_FRACTIONS[gear]).reshape(self.K.shape)
        return cnn_forward(K, self.F * block_keep_mask(self.F, GEAR_FRACTIONS[gear]), X)

    # -- gears ------------------------------------------------------------
    def set_gears(self, gear1: int, gear2: int):
        self.gears = (gear1, gear2)
        self.PK = self.K * block_keep_mask(self.K.reshape(N_CH, -1), GEAR_FRACTIONS[gear1]).reshape(self.K.shape)
        self.PF = self.F * block_keep_mask(self.F, GEAR_FRACTIONS[gear2])
        self.taps = [(u, v) for u in range(KS) for v in range(KS) if np.any(self.PK[:, u, v])]
        self.fc_rows = [(k, c) for k in range(N_OUT) for c in range(N_CH)]

    def pruned_weights(self):
        return self.PK, self.PF

    def rotation_steps(self) -> List[int]:
        steps = {tap_step(u, v) for u, v in self.taps} - {0}
        steps |= {-WINDOW * m for m in range(1, 4)}
        steps |= {WINDOW >> k for k in range(1, 9)}
        return sorted(steps)

    # -- plaintext material -------------------------------------------------
    def _conv_weights(self, level: int, scale: float) -> torch.Tensor:
        """[8][taps][level] prepared pairs (ring.prepared_pairs) of the
        residues r of round(K * scale)."""
        key = (self.gears[0], level, scale, self.ck.ring.fp64)
        if key not in self._conv:
            ck = self.ck
            D = np.zeros((N_CH, len(self.taps), level, 2), dtype=np.uint64)
            for c in range(N_CH):
                for b, (u, v) in enumerate(self.taps):
                    w = int(round(float(self.PK[c, u, v]) * scale))
                    for i in range(level):
                        q = ck.primes[i]
                        r = w % q
                        D[c, b, i] = ck.ring.prepared_pairs(r, q)
            self._conv[key] = ck.ring.tensor(D)
        return self._conv[key]

    def _fc_weights(self, level: int, scale: float) -> torch.Tensor:
        """[10][8] evaluation-form prepared plaintexts of fc_weight_blocks,
        outputs in FC_ORDER."""
        key = (self.gears[1], level, scale)
        if key not in self._fc:
            W = fc_weight_blocks(self.PF)
            rows = [self._slots(W[N_CH * k + c]) for k in FC_ORDER for c in range(N_CH)]
            prep = self.prepared_pt(np.asarray(rows), level, scale)
            self._fc[key] = prep.view(N_OUT, N_CH, level, self.ck.N, 2)
        return self._fc[key]

    # -- evaluation ---------------------------------------------------------
    def encrypt_inputs(self, X: np.ndarray) -> Ciphertext:
        ck = self.ck
        return ck.encrypt(ck.encode(pack_images(X, self.S)), self.sk)

    def conv(self, ct: Ciphertext) -> Ciphertext:
        """[B] input ciphertexts -> [8 B] channel ciphertexts (channel-major),
        rescaled back to the base scale."""
        ck = self.ck
        B, l = ct.batch, ct.level
        E = self.rot_stack(ct, [tap_step(u, v) for u, v in self.taps])       # [taps][B][2][l][N]
        scale = float(ck.primes[l - 1]) * ck.p.scale / ct.scale
        D = self._conv_weights(l, scale)
        out = torch.empty(N_CH, B, 2, l, ck.N, dtype=torch.uint64, device=ck.device)
        call("ckks_lincomb_scalar", ck.ring.handle, _ptr(E), _ptr(D), _ptr(out), len(self.taps), N_CH, B, l,
             ck.ring.stream())
        self.ops["ptmul_rows"] += len(self.taps) * N_CH * B
        return ck.rescale(Ciphertext(out.view(N_CH * B, 2, l, ck.N), ct.scale * scale))

    def activation(self, ct: Ciphertext) -> Ciphertext:
        self.ops["hmult"] += ct.batch
        return self.ck.hmult(ct, ct, self.rlk)

    def fc(self, ct: Ciphertext, B: int, in_place: bool = False) -> List[Ciphertext]:
        """[8 B] activated channel ciphertexts -> ceil(10 / 4) packed output
        ciphertexts [B] each (the 2x2 pool is in the FC weights).  in_place:
        ct's words are transformed in place (the caller drops ct)."""
        ck = self.ck
        l = ct.level
        E = ct.data if in_place else ct.data.clone()
        ck.ring.ntt_(E.view(-1, l, ck.N))
        scale = float(ck.primes[l - 1]) * ck.p.scale / ct.scale
        D = self._fc_weights(l, scale)
        acc = torch.empty(N_OUT, B, 2, l, ck.N, dtype=torch.uint64, device=ck.device)
        call("ckks_lincomb", ck.ring.handle, _ptr(E), _ptr(D), _ptr(acc), N_CH, N_OUT, B, l, ck.ring.stream())
        self.ops["ptmul_rows"] += N_CH * N_OUT * B
        ck.ring.intt_(acc.view(-1, l, ck.N))
        outs = ck.rescale(Ciphertext(acc.view(N_OUT * B, 2, l, ck.N), ct.scale * scale))
        # output k goes to window k mod 4 of packed ciphertext k // 4: every
        # output shifted by -256 (k mod 4) in one call (rows in FC_ORDER: the
        # shifted ones one contiguous run, outputs 0, 4, 8 the last rows)
        sh = self.rot_blocks(outs, B, [-WINDOW * (k % 4) for k in FC_ORDER])
        pos = {k: j for j, k in enumerate(FC_ORDER)}
        out = lambda k: sh[pos[k]].data
        groups = list(range(0, N_OUT, 4))
        first = pos[groups[0]]
        if [pos[p] for p in groups] == list(range(first, first + len(groups))):
            z = outs.data[first * B:(first + len(groups)) * B]    # a view: summed into in place
        else:
            z = torch.cat([out(p) for p in groups])
        for m in range(1, 4):
            for i, p in enumerate(groups):
                if p + m < N_OUT:
                    zi = z[i * B:(i + 1) * B]
                    ck.ring.add(zi, out(p + m), out=zi)
        # rotate-and-sum over each 256-slot window, all packed ciphertexts at once
        zc = Ciphertext(z, outs.scale)
        for k in range(1, 9):
            zc = self.rot_add(zc, WINDOW >> k)
        return [Ciphertext(zc.data[i * B:(i + 1) * B], zc.scale) for i in range(len(groups))]

    def forward(self, ct: Ciphertext) -> List[Ciphertext]:
        self.check_arith()
        B = ct.batch
        l1, l2 = self.entry_levels()
        y = self.conv(self.enter(ct, l1))
        y = self.dare(y, level=l2)
        return self.fc(self.activation(y), B, in_place=True)

    def decrypt_outputs(self, cts: List[Ciphertext], n: int) -> np.ndarray:
        ck = self.ck
        out = np.zeros((n, N_OUT))
        for p, ct in enumerate(cts):
            z = to_numpy(ck.decode(ck.decrypt(ct, self.sk), ct.scale)).real
            for s in range(n):
                c, j = divmod(s, self.S)
                for m in range(min(4, N_OUT - 4 * p)):
                    out[s, 4 * p + m] = z[c, j * SLOTS_PER_SAMPLE + WINDOW * m]
        return out
