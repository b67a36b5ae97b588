"""Host side of the matrix-core probe (ckks_probe_dft64): the tables the two
kernels read, and a driver that times both on one batch.

The probe computes y = G^reps x mod q for a batch of length-64 vectors, G
the 64-point negacyclic forward transform exactly as the radix-2 stages
compute it (natural order in, bit-reversed order out; the Longa-Naehrig
stage order with psi of order 128).  The matrix path needs G split into
balanced signed 8-bit digits, laid out in the operand order of
v_mfma_i32_16x16x64_i8, and one constant that carries the digit bias (the
kernel keeps each value as z = x + BIAS mod q, whose bytes u are fed as
u - 128) and the lift that keeps the recombined word positive.  The bounds the kernel's recombination relies on
are checked here for the digits built.  No counterpart in the reference."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

TABLE_BYTES = 51200          # CKKS_PROBE_DFT64_TABLE_BYTES
PLANES = 6                   # bytes of an operand below 2^48
BIAS = 128 * sum(1 << (8 * s) for s in range(PLANES))


def bitrev6(i: int) -> int:
    return int(f"{i:06b}"[::-1], 2)


def pi_index(cm: int, g: int, j: int) -> int:
    """Data index held at byte j of a lane of K group g (and, for the
    accumulators, the row of register e = j % 4 of tile j // 4)."""
    return 16 * (j >> 2) + 4 * g + (j & 3) if cm == 0 else 16 * (j >> 2) + g + 4 * (j & 3)


def radix2_forward(x: list[int], tw: list[int], q: int) -> list[int]:
    """The kernel's six stages with exact arithmetic mod q."""
    a = list(x)
    for m in (1, 2, 4, 8, 16, 32):
        t = 32 // m
        for i in range(m):
            w = tw[m + i]
            for j in range(2 * i * t, 2 * i * t + t):
                u, v = a[j], a[j + t] * w % q
                a[j], a[j + t] = (u + v) % q, (u - v) % q
    return a


def balanced_digits(v: int, n: int = PLANES) -> list[int]:
    out = []
    for _ in range(n - 1):
        d = ((v + 128) & 255) - 128
        out.append(d)
        v = (v - d) >> 8
    out.append(v)
    return out


@dataclass
class Dft64Tables:
    q: int
    psi: int
    G: list                  # 64 x 64 ints, y = G x
    blob: np.ndarray         # uint8[TABLE_BYTES], the device tables
    ws40: int
    kq: int                  # K q + BIAS / 2^40 mod q
    fdig: np.ndarray         # int64 [64, 64, PLANES] balanced digits of G


def build_tables(q: int, psi: int) -> Dft64Tables:
    if pow(psi, 64, q) != q - 1:
        raise ValueError("psi must have order 128 modulo q")
    if not (q & 1) or q >= 1 << 45 or q <= 1 << 40:
        raise ValueError("q must be odd with 2^40 < q < 2^45")
    tw = [pow(psi, bitrev6(k), q) for k in range(64)]
    cols = [radix2_forward([int(i == k) for i in range(64)], tw, q) for k in range(64)]
    G = [[cols[k][i] for k in range(64)] for i in range(64)]
    fdig = np.array([[balanced_digits(G[i][k]) for k in range(64)] for i in range(64)], dtype=np.int64)
    assert fdig[..., :PLANES - 1].min() >= -128 and fdig[..., :PLANES - 1].max() <= 127
    top = int(np.abs(fdig[..., PLANES - 1]).max())
    assert top <= 127
    # |A_k| bound: 64 products of digits, |d_s| <= 128 for every input byte
    dmax = [int(np.abs(fdig[..., t]).max()) for t in range(PLANES)]
    amax = [64 * 128 * sum(dmax[t] for t in range(PLANES) if 0 <= k - t < PLANES) for k in range(2 * PLANES - 1)]
    # the kernel's merged pairs p0 = A0 + 2^8 A1, p1 = A2 + 2^8 A3 (L) and
    # A5 + 2^8 A6, A7 + 2^8 A8, A9 + 2^8 A10 (M), each an int32
    pair = [amax[k] + 256 * amax[k + 1] for k in (0, 2, 5, 7, 9)]
    assert max(pair) < 1 << 31
    lmax = pair[0] + (pair[1] << 16) + (amax[4] << 32)
    mmax = pair[2] + (pair[3] << 16) + (pair[4] << 32)
    assert lmax < 1 << 55
    lift = mmax + (lmax >> 40) + 1
    kq = -(-lift // q) * q
    assert 2 * kq + q < 1 << 62
    inv40 = pow(1 << 40, -1, q)
    kq += BIAS * inv40 % q
    ws40 = ((1 << 40) << 63) // q
    blob = np.zeros(TABLE_BYTES, dtype=np.uint8)
    twp = np.zeros((64, 2), dtype=np.uint64)
    for k in range(1, 64):
        twp[k] = (tw[k], (tw[k] << 63) // q)
    blob[:1024] = twp.view(np.uint8).reshape(-1)
    cst = np.zeros(128, dtype=np.uint64)
    cst[:3] = (kq, (-BIAS) % q, BIAS)
    blob[1024:2048] = cst.view(np.uint8)
    fpk = np.zeros((2, 4, PLANES, 64, 16), dtype=np.int8)
    for cm in (0, 1):
        for rt in range(4):
            for t in range(PLANES):
                for lane in range(64):
                    row = 16 * rt + (lane & 15)
                    for j in range(16):
                        fpk[cm, rt, t, lane, j] = fdig[row, pi_index(cm, lane >> 4, j), t]
    blob[2048:] = fpk.view(np.uint8).reshape(-1)
    return Dft64Tables(q, psi, G, blob, ws40, kq, fdig)


def matmul_mod(G: list, x: np.ndarray, q: int, reps: int) -> np.ndarray:
    """G^reps x mod q for x [64][V] (exact, Python integers per column)."""
    y = [[int(v) for v in x[:, c]] for c in range(x.shape[1])]
    for _ in range(reps):
        y = [[sum(G[i][k] * col[k] for k in range(64)) % q for i in range(64)] for col in y]
    return np.array(y, dtype=np.uint64).T.copy()


def launch(x, y, tables, t: Dft64Tables, method: int, reps: int, stream: int = 0) -> None:
    """One launch through the C ABI on torch device tensors: x, y uint64
    [64][count] (count a multiple of 64), tables the blob on the device."""
    from . import _lib
    _lib.call("ckks_probe_dft64", x.data_ptr(), y.data_ptr(), tables.data_ptr(), t.q, t.ws40, x.shape[1],
              method, reps, stream)
