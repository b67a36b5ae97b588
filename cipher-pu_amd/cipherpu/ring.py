"""Ring context: device twiddle tables for one ring degree over a list of
primes, and thin wrappers of the C ABI over torch uint64 device tensors.

Tensors are [..., limbs, N] uint64 on the context's GPU; limb i of a call
with ``prime0`` is reduced modulo ``primes[prime0 + i]``.  All arithmetic
runs in the HIP library; torch only holds memory and streams."""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import call


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def as_u64(x, device) -> torch.Tensor:
    """Host data (list / numpy / tensor) -> contiguous uint64 tensor on device."""
    if isinstance(x, torch.Tensor):
        t = x
        if t.dtype != torch.uint64:
            t = t.to(torch.int64).view(torch.uint64) if t.dtype != torch.int64 else t.view(torch.uint64)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.uint64)))
    return t.to(device).contiguous()


def to_numpy(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy()


class RingContext:
    """Twiddle / Shoup tables for Z_q[X]/(X^N +- 1) over ``primes``
    (replaces the per-script constants of tb/ref/gen_poly_mult_vectors_n12.py:36-44
    and the table generator scripts/gen_omega_tables.py:96-106)."""

    def __init__(self, N: int, primes: Sequence[int], roots: Optional[Sequence[int]] = None,
                 cyclic: bool = False, device: int = 0):
        if not torch.cuda.is_available():
            raise RuntimeError("cipherpu needs a GPU: no HIP device is visible")
        self.N = int(N)
        self.primes = [int(q) for q in primes]
        self.cyclic = bool(cyclic)
        self.device = torch.device("cuda", device)
        self._h = ctypes.c_void_p()
        pr = (ctypes.c_uint64 * len(self.primes))(*self.primes)
        rt = (ctypes.c_uint64 * len(self.primes))(*[int(r) for r in roots]) if roots is not None else None
        with torch.cuda.device(self.device):
            call("ckks_ctx_create", ctypes.byref(self._h), self.N, len(self.primes), pr, rt,
                 _lib.CKKS_CYCLIC if cyclic else 0, device)
        r = ctypes.c_uint64()
        self.fp64 = True      # the context's arithmetic (set_arith)
        self.roots = []
        for i in range(len(self.primes)):
            call("ckks_ctx_root", self._h, i, ctypes.byref(r))
            self.roots.append(int(r.value))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().ckks_ctx_destroy(h)
            except Exception:
                pass
            self._h = None

    @property
    def handle(self):
        return self._h

    def set_threads(self, ntt_threads: int = 0, ks_threads: int = 0, tensor_threads: int = 0):
        call("ckks_ctx_set_threads", self._h, ntt_threads, ks_threads, tensor_threads)

    def set_front(self, fused: bool = True):
        call("ckks_ctx_set_front", self._h, 1 if fused else 0)

    def set_ks_order(self, pair_major: bool = True):
        """HMult key-switch placement: ciphertext-major per XCD (True) or
        target-major (False); the results are the same words."""
        call("ckks_ctx_set_ks_order", self._h, 1 if pair_major else 0)

    def set_arith(self, fp64: bool = True):
        """Kernels on the FP64 residue pipeline (True) or on 64-bit integer
        words with Shoup and Barrett products (False); the results are the
        same words.  Prepared keys and plaintexts are in the form of the
        arithmetic in force when they were prepared (``fp64``) and must be
        prepared again after a change."""
        call("ckks_ctx_set_arith", self._h, 1 if fp64 else 0)
        self.fp64 = bool(fp64)

    def prepared_pairs(self, w, q) -> np.ndarray:
        """prepared_pairs in this context's arithmetic."""
        return prepared_pairs(w, q, self.fp64)

    def set_hoist_order(self, btile: int = 0, xcd: bool = False):
        """Workgroup placement of the hoisted key switch (ckks_ctx_set_hoist_order)."""
        call("ckks_ctx_set_hoist_order", self._h, btile, 1 if xcd else 0)

    # -- helpers ---------------------------------------------------------
    def stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def tensor(self, x) -> torch.Tensor:
        return as_u64(x, self.device)

    def _rows(self, x: torch.Tensor, prime0: int):
        if x.dtype != torch.uint64 or not x.is_cuda or x.device != self.device:
            raise TypeError("expected a uint64 tensor on the context's GPU")
        if not x.is_contiguous():
            raise ValueError("tensor must be contiguous")
        if x.shape[-1] != self.N:
            raise ValueError(f"last dimension must be N={self.N}")
        limbs = x.shape[-2] if x.dim() >= 2 else 1
        polys = x.numel() // (limbs * self.N)
        if prime0 < 0 or prime0 + limbs > len(self.primes):
            raise ValueError("limbs outside the context's primes")
        return polys, limbs

    # -- transforms --------------------------------------------------------
    def ntt_(self, x: torch.Tensor, prime0: int = 0, natural: bool = False) -> torch.Tensor:
        p, l = self._rows(x, prime0)
        call("ckks_ntt_fwd", self._h, _ptr(x), p, l, prime0, _lib.CKKS_NATURAL_ORDER if natural else 0,
             self.stream())
        return x

    def intt_(self, x: torch.Tensor, prime0: int = 0, natural: bool = False) -> torch.Tensor:
        p, l = self._rows(x, prime0)
        call("ckks_ntt_inv", self._h, _ptr(x), p, l, prime0, _lib.CKKS_NATURAL_ORDER if natural else 0,
             self.stream())
        return x

    def ntt(self, x, prime0: int = 0, natural: bool = False) -> torch.Tensor:
        return self.ntt_(self.tensor(x).clone(), prime0, natural)

    def intt(self, x, prime0: int = 0, natural: bool = False) -> torch.Tensor:
        return self.intt_(self.tensor(x).clone(), prime0, natural)

    # -- ring arithmetic -------------------------------------------------------
    def poly_mul(self, a: torch.Tensor, b: torch.Tensor, prime0: int = 0, out=None) -> torch.Tensor:
        p, l = self._rows(a, prime0)
        if b.shape != a.shape:
            raise ValueError("operand shapes differ")
        self._rows(b, prime0)
        out = torch.empty_like(a) if out is None else out
        call("ckks_poly_mul", self._h, _ptr(a), _ptr(b), _ptr(out), p, l, prime0, self.stream())
        return out

    def _pw(self, a, b, op, prime0, out):
        p, l = self._rows(a, prime0)
        if b is not None and b.shape != a.shape:
            raise ValueError("operand shapes differ")
        out = torch.empty_like(a) if out is None else out
        call("ckks_pointwise", self._h, _ptr(a), _ptr(b) if b is not None else None, _ptr(out), p, l,
             prime0, op, self.stream())
        return out

    def mul(self, a, b, prime0=0, out=None):
        return self._pw(a, b, 0, prime0, out)

    def add(self, a, b, prime0=0, out=None):
        return self._pw(a, b, 1, prime0, out)

    def sub(self, a, b, prime0=0, out=None):
        return self._pw(a, b, 2, prime0, out)

    def neg(self, a, prime0=0, out=None):
        return self._pw(a, None, 3, prime0, out)

    def rescale(self, x: torch.Tensor, prime0: int = 0) -> torch.Tensor:
        """[..., l, N] -> [..., l-1, N]: exact rounding division by the last limb."""
        p, l = self._rows(x, prime0)
        out = torch.empty(*x.shape[:-2], l - 1, self.N, dtype=torch.uint64, device=self.device)
        call("ckks_rescale", self._h, _ptr(x), _ptr(out), p, l, prime0, self.stream())
        return out

    def automorphism(self, x: torch.Tensor, galois: int, prime0: int = 0) -> torch.Tensor:
        p, l = self._rows(x, prime0)
        out = torch.empty_like(x)
        call("ckks_automorphism", self._h, _ptr(x), _ptr(out), p, l, prime0, int(galois), self.stream())
        return out

    def noise_stats(self, x: torch.Tensor, prime0: int = 0) -> np.ndarray:
        """Per (poly, limb) row: [sum c', sum c'^2 (exact int), max c', sum centered],
        c' = min(c, q - c).  Returns an object array of Python ints."""
        p, l = self._rows(x, prime0)
        st = torch.empty(p * l, 5, dtype=torch.uint64, device=self.device)
        call("ckks_noise_stats", self._h, _ptr(x), _ptr(st), p, l, prime0, self.stream())
        a = to_numpy(st).astype(object)
        out = np.empty((p * l, 4), dtype=object)
        for r in range(p * l):
            sa, lo, hi, mx, sg = (int(v) for v in a[r])
            out[r] = [sa, lo + (hi << 64), mx, sg - (1 << 64) if sg >> 63 else sg]
        return out

    # -- kernel timing --------------------------------------------------------------
    def timing_start(self, max_launches: int = 4096) -> None:
        """Bracket every kernel the transform, product, HMult, relinearise and
        rotate entry points launch with HIP events on their stream."""
        call("ckks_timing_start", self._h, int(max_launches))

    def timing_stop(self) -> dict:
        """{kernel kind: (summed ms, launches)} since timing_start."""
        n = len(_lib.KERNEL_KINDS)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int * n)()
        call("ckks_timing_stop", self._h, ms, cnt, n)
        return {k: (ms[i], cnt[i]) for i, k in enumerate(_lib.KERNEL_KINDS) if cnt[i]}

    def noise_moments(self, x: torch.Tensor, prime0: int = 0) -> torch.Tensor:
        """Per (poly, limb) row, on the device in float64: [mean c' / q,
        rms c' / q, max c' / q, mean centred c / q], c' = min(c, q - c)
        (the monitor's energy E = sum c'^2 / N gives rms = sqrt(E);
        rtl/m4/noise_monitor.sv:44-78)."""
        return self.noise_moments_many([x], prime0)

    def noise_moments_many(self, xs, prime0: int = 0) -> torch.Tensor:
        """noise_moments of several tensors, rows concatenated in order: one
        statistics launch per tensor into a shared buffer and one conversion
        (the gear controller's features for many chunks at once)."""
        shapes = [self._rows(x, prime0) for x in xs]
        st = torch.empty(sum(p * l for p, l in shapes), 5, dtype=torch.uint64, device=self.device)
        qs, off = [], 0
        for x, (p, l) in zip(xs, shapes):
            call("ckks_noise_stats", self._h, _ptr(x), _ptr(st[off:off + p * l]), p, l, prime0, self.stream())
            qs += [float(self.primes[prime0 + i]) for i in range(l)] * p
            off += p * l
        v = st.view(torch.int64).to(torch.float64)
        v = torch.where(v < 0, v + 2.0 ** 64, v)          # unsigned words
        sg = st[:, 4].view(torch.int64).to(torch.float64)  # signed sum
        q = torch.tensor(qs, dtype=torch.float64, device=self.device)
        n = float(self.N)
        out = torch.stack([v[:, 0] / n / q, torch.sqrt((v[:, 1] + v[:, 2] * 2.0 ** 64) / n) / q, v[:, 3] / q,
                           sg / n / q], 1)
        return out

    # -- canonical embedding -------------------------------------------------------
    def embed(self, coeffs: torch.Tensor, order: int = 0) -> torch.Tensor:
        """Real coefficients [B, N] (float64) -> complex slots [B, N/2]."""
        c = coeffs.to(self.device, torch.float64).contiguous()
        B = c.numel() // self.N
        out = torch.empty(B, self.N // 2, dtype=torch.complex128, device=self.device)
        call("ckks_embed", self._h, _ptr(c), _ptr(out), B, order, self.stream())
        return out

    def embed_inv(self, slots: torch.Tensor, order: int = 0) -> torch.Tensor:
        """Complex slots [B, N/2] -> real coefficients [B, N] (unrounded)."""
        z = slots.to(self.device, torch.complex128).contiguous()
        B = z.numel() // (self.N // 2)
        out = torch.empty(B, self.N, dtype=torch.float64, device=self.device)
        call("ckks_embed_inv", self._h, _ptr(z), _ptr(out), B, order, self.stream())
        return out

    def encode(self, slots, scale: float, l: int, prime0: int = 0) -> torch.Tensor:
        """Complex (or real) slots [B, N/2] -> plaintext [B, l, N] uint64:
        round(scale * embedding^-1(z)) reduced into every limb."""
        z = torch.as_tensor(slots).to(self.device, torch.complex128).reshape(-1, self.N // 2).contiguous()
        B = z.shape[0]
        work = torch.empty(B, self.N, dtype=torch.float64, device=self.device)
        pt = torch.empty(B, l, self.N, dtype=torch.uint64, device=self.device)
        call("ckks_encode", self._h, _ptr(z), _ptr(pt), B, l, prime0, float(scale), _ptr(work), self.stream())
        return pt

    def decode(self, pt: torch.Tensor, scale: float, prime0: int = 0) -> torch.Tensor:
        """Plaintext [B, l, N] -> complex slots [B, N/2]: exact CRT, centered, / scale."""
        p, l = self._rows(pt, prime0)
        work = torch.empty(p, self.N, dtype=torch.float64, device=self.device)
        out = torch.empty(p, self.N // 2, dtype=torch.complex128, device=self.device)
        call("ckks_decode", self._h, _ptr(pt), _ptr(out), p, l, prime0, float(scale), _ptr(work), self.stream())
        return out


def prepared_pairs(w, q, fp64: bool = True) -> np.ndarray:
    """Reduced words w < q -> the prepared-pair form that the key switches
    and the plaintext products read, shape w.shape + (2,) uint64; q
    broadcasts against w.  FP64 arithmetic: the IEEE doubles w and
    RN(w / q), as bit patterns.  Integer arithmetic: w and its Shoup
    companion floor(w 2^63 / q)."""
    w = np.asarray(w, dtype=np.uint64)
    if fp64:
        wd = w.astype(np.float64)                             # exact below 2^53
        qd = np.asarray(q, dtype=np.float64)
        return np.ascontiguousarray(np.stack([wd, wd / qd], axis=-1)).view(np.uint64)
    wb, qb = np.broadcast_arrays(w, np.asarray(q, dtype=np.uint64))
    return np.ascontiguousarray(np.stack([wb, shoup_companion(wb, qb)], axis=-1))


def shoup_companion(w, q) -> np.ndarray:
    """floor(w 2^63 / q) for words w < q < 2^63, by long division in chunks
    of s bits with q 2^s < 2^64, so that no step overflows."""
    w, q = np.broadcast_arrays(np.asarray(w, dtype=np.uint64), np.asarray(q, dtype=np.uint64))
    if w.size == 0:
        return np.zeros(w.shape, dtype=np.uint64)
    step = 64 - int(q.max()).bit_length()
    r, quo, left = w.copy(), np.zeros(w.shape, dtype=np.uint64), 63
    while left:
        k = np.uint64(min(step, left))
        r = r << k
        d = r // q
        r = r - d * q
        quo = (quo << k) | d
        left -= int(k)
    return quo


def bitrev_(x: torch.Tensor) -> torch.Tensor:
    """In-place bit-reversal permutation of the rows of x (last dim N)."""
    N = x.shape[-1]
    call("ckks_bitrev", _ptr(x), x.numel() // N, N,
         ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream))
    return x


def mod_switch(x: torch.Tensor, q_in: int, q_out: int) -> torch.Tensor:
    out = torch.empty_like(x)
    call("ckks_mod_switch", _ptr(x), _ptr(out), x.numel(), int(q_in), int(q_out),
         ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream))
    return out
