"""CKKS scheme on the HIP core.

Conventions (frozen; the reference mixes several, see its
scripts/exp_e33_real_ciphertext.py:138-175 for the one kept here):
  ciphertext ct = (c0, c1) in coefficient form, [B, 2, l, N] uint64,
  decrypt m = c0 + c1 s;  fresh encryption c1 = -a, c0 = a s + e + m;
  key-switch key from s_from to s, one digit per prime (dnum = L) and one
  special prime p: digit j = (b_j, a_j) over q_0..q_(L-1), p with
  b_j = a_j s + e_j + g_j s_from, g_j = p on limb j and 0 elsewhere;
  relinearisation key: s_from = s^2; Galois key for g: s_from = s(X^g).
Sampling (host, numpy Generator): s ternary uniform in {-1, 0, 1}, errors
rounded Gaussian sigma = 3.2, a uniform per limb."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from ._lib import call
from .params import CHAIN4, SPECIAL_P
from .ring import RingContext, to_numpy


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


@dataclass
class CKKSParams:
    N: int = 8192
    primes: Sequence[int] = CHAIN4
    special: int = SPECIAL_P
    scale: float = 2.0 ** 40
    sigma: float = 3.2
    # roots of order 2N for primes + [special]; None derives them from the
    # smallest generator (which is what the reference's chain file holds)
    roots: Optional[Sequence[int]] = None

    @property
    def L(self) -> int:
        return len(self.primes)

    @classmethod
    def from_chain_json(cls, path: str, N: int = 8192, special: int = SPECIAL_P, **kw) -> "CKKSParams":
        """Parameters from a modulus chain in the reference's JSON schema
        (rtl/m2/modulus_chain_chain4.json): its primes, with the roots of
        order 2N taken from its omega column (params.chain_primes_and_roots);
        the special prime's root is the one the library derives."""
        from .params import chain_primes_and_roots, default_root, load_chain_json
        primes, roots = chain_primes_and_roots(load_chain_json(path), N)
        return cls(N=N, primes=tuple(primes), special=special,
                   roots=tuple(roots) + (default_root(special, 2 * N),), **kw)


@dataclass
class Ciphertext:
    data: torch.Tensor          # [B, 2, l, N] uint64, coefficient form
    scale: float

    @property
    def level(self) -> int:
        return self.data.shape[2]

    @property
    def batch(self) -> int:
        return self.data.shape[0]


@dataclass
class SecretKey:
    s: np.ndarray               # [N] int64 in {-1, 0, 1}
    rns: torch.Tensor           # [K, N] uint64 over all primes (incl. special)
    rows: Dict[object, torch.Tensor] = field(default_factory=dict, repr=False)   # CKKS.key_rows / key_eval caches


@dataclass
class SwitchKey:
    coeff: torch.Tensor         # [L, 2, L+1, N] uint64, coefficient form (b_j, a_j)
    prepared: torch.Tensor      # evaluation order: FP64 [L, 2, L+1, N] (the words as doubles), INT64
                                # [L, 2, L+1, N, 2] (word, Shoup companion) pairs
    galois: int = 0             # 0 for the relinearisation key
    fp64: bool = True           # the arithmetic the pairs were prepared for (RingContext.set_arith)
    hoisted: Optional[torch.Tensor] = None   # prepared, permuted for ckks_rotate_hoisted


class CKKS:
    def __init__(self, params: CKKSParams = CKKSParams(), device: int = 0, seed: int = 0,
                 sampler: str = "device"):
        """sampler: "device" draws the encryption mask and errors on the GPU
        (ckks_sample_uniform / ckks_sample_gaussian); "host" draws them with
        numpy.  Both are seeded from `seed`."""
        if sampler not in ("device", "host"):
            raise ValueError("sampler must be 'device' or 'host'")
        self.p = params
        self.N = params.N
        self.L = params.L
        self.primes = [int(q) for q in params.primes] + [int(params.special)]
        self.sp = self.L                      # ring index of the special prime
        self.ring = RingContext(self.N, self.primes, roots=params.roots, device=device)
        self.device = self.ring.device
        self.rng = np.random.default_rng(seed)
        self.sampler = sampler
        self._seed = int(seed) & ((1 << 64) - 1)
        self._draws = 0
        self._ws = {}      # stream handle -> workspace
        # device-sampled encryption and decryption in one kernel each
        # (ckks_encrypt_sk / ckks_decrypt_sk) where they apply: the words of
        # the separate samplers, product and sums
        self.fused_encrypt = True
        self.fused_decrypt = True

    # -- sampling (host) ------------------------------------------------------
    def ternary(self, *shape) -> np.ndarray:
        return self.rng.integers(-1, 2, size=shape + (self.N,), dtype=np.int64)

    def gaussian(self, *shape) -> np.ndarray:
        return np.rint(self.rng.normal(0.0, self.p.sigma, size=shape + (self.N,))).astype(np.int64)

    def uniform(self, *shape, limbs: Sequence[int]) -> torch.Tensor:
        """[..., len(limbs), N] uniform in [0, q_i) per limb."""
        out = np.empty(shape + (len(limbs), self.N), dtype=np.uint64)
        for k, i in enumerate(limbs):
            out[..., k, :] = self.rng.integers(0, self.primes[i], size=shape + (self.N,), dtype=np.uint64)
        return self.ring.tensor(out)

    def _next_seed(self) -> int:
        """A fresh 64-bit seed per device draw (splitmix64 of seed and count)."""
        self._draws += 1
        z = (self._seed + self._draws * 0x9E3779B97F4A7C15) & ((1 << 64) - 1)
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & ((1 << 64) - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & ((1 << 64) - 1)
        return z ^ (z >> 31)

    def uniform_device(self, rows: int, l: int) -> torch.Tensor:
        """[rows, l, N] uniform residues over primes 0..l-1, drawn on the GPU."""
        out = torch.empty(rows, l, self.N, dtype=torch.uint64, device=self.device)
        call("ckks_sample_uniform", self.ring.handle, _ptr(out), rows, l, 0, self._next_seed(), self.ring.stream())
        return out

    def gaussian_device(self, rows: int, l: int) -> torch.Tensor:
        """[rows, l, N] residues of rounded N(0, sigma^2) errors, one integer
        per (row, coefficient) across the limbs, drawn on the GPU."""
        out = torch.empty(rows, l, self.N, dtype=torch.uint64, device=self.device)
        call("ckks_sample_gaussian", self.ring.handle, _ptr(out), rows, l, 0, float(self.p.sigma),
             self._next_seed(), self.ring.stream())
        return out

    def small_to_rns(self, x: np.ndarray, limbs: Sequence[int]) -> torch.Tensor:
        """Small signed integers [..., N] -> [..., len(limbs), N] residues."""
        x = np.asarray(x, dtype=np.int64)
        out = np.empty(x.shape[:-1] + (len(limbs), self.N), dtype=np.uint64)
        for k, i in enumerate(limbs):
            q = self.primes[i]
            out[..., k, :] = np.where(x < 0, np.uint64(q) - (-x).astype(np.uint64), x.astype(np.uint64))
        return self.ring.tensor(out)

    # -- keys -----------------------------------------------------------------
    def keygen(self) -> SecretKey:
        s = self.ternary()
        return SecretKey(s=s, rns=self.small_to_rns(s, range(len(self.primes))))

    def switch_key(self, sk: SecretKey, s_from: torch.Tensor, galois: int = 0,
                   a: Optional[torch.Tensor] = None, e: Optional[np.ndarray] = None) -> SwitchKey:
        """Hybrid key from s_from ([K, N] RNS over all primes) to sk.
        a ([L, L+1, N]) and e ([L, N]) may be given (tests); else sampled."""
        L, K = self.L, self.L + 1
        if a is None:
            a = self.uniform(L, limbs=range(K))
        if e is None:
            e = self.gaussian(L)
        E = self.small_to_rns(e, range(K))
        S = sk.rns.unsqueeze(0).expand(L, K, self.N).contiguous()
        b = self.ring.add(self.ring.poly_mul(a, S), E)
        gs = torch.zeros_like(b)
        for j in range(L):
            gs[j, j] = s_from[j]
        g = np.zeros((L, K, self.N), dtype=np.uint64)
        for j in range(L):
            g[j, j, :] = self.p.special % self.primes[j]
        b = self.ring.add(b, self.ring.mul(gs, self.ring.tensor(g)))
        coeff = torch.stack([b, a], dim=1).contiguous()          # [L, 2, K, N]
        prepared = self._key_buffer()
        work = coeff.clone()
        call("ckks_key_prepare", self.ring.handle, _ptr(work), _ptr(prepared), L, K, self.sp, self.ring.stream())
        return SwitchKey(coeff=coeff, prepared=prepared, galois=galois, fp64=self.ring.fp64)

    def _key_buffer(self) -> torch.Tensor:
        """A prepared key in the context's arithmetic (ckks_key_prepare)."""
        shape = (self.L, 2, self.L + 1, self.N) + (() if self.ring.fp64 else (2,))
        return torch.empty(*shape, dtype=torch.uint64, device=self.device)

    def load_switch_key(self, coeff, galois: int = 0) -> SwitchKey:
        """Key from its coefficient form [L, 2, L+1, N] (b_j, a_j) -- e.g. read
        back from disk -- prepared for the kernels."""
        coeff = self.ring.tensor(coeff)
        L, K = self.L, self.L + 1
        if tuple(coeff.shape) != (L, 2, K, self.N):
            raise ValueError(f"key must be [{L}, 2, {K}, {self.N}]")
        prepared = self._key_buffer()
        work = coeff.clone()
        call("ckks_key_prepare", self.ring.handle, _ptr(work), _ptr(prepared), L, K, self.sp, self.ring.stream())
        return SwitchKey(coeff=coeff, prepared=prepared, galois=galois, fp64=self.ring.fp64)

    def relin_key(self, sk: SecretKey, **kw) -> SwitchKey:
        s2 = self.ring.poly_mul(sk.rns, sk.rns)
        return self.switch_key(sk, s2, 0, **kw)

    def galois_key(self, sk: SecretKey, step: int, **kw) -> SwitchKey:
        g = self.galois_elt(step)
        sg = self.ring.automorphism(sk.rns, g)
        return self.switch_key(sk, sg, g, **kw)

    def galois_elt(self, step: int) -> int:
        return pow(5, step % (self.N // 2), 2 * self.N)

    # -- encode / encrypt ---------------------------------------------------------
    def encode(self, z, level: Optional[int] = None, scale: Optional[float] = None) -> torch.Tensor:
        l = self.L if level is None else level
        return self.ring.encode(z, self.p.scale if scale is None else scale, l)

    def decode(self, pt: torch.Tensor, scale: float) -> torch.Tensor:
        return self.ring.decode(pt, scale)

    def encrypt(self, pt: torch.Tensor, sk: SecretKey, scale: Optional[float] = None,
                a: Optional[torch.Tensor] = None, e: Optional[np.ndarray] = None) -> Ciphertext:
        """pt [B, l, N] -> ct [B, 2, l, N]:  c1 = -a, c0 = a s + e + m
        (exp_e33_real_ciphertext.py:171-175)."""
        B, l = pt.shape[0], pt.shape[1]
        dev = self.sampler == "device"
        if (dev and a is None and e is None and self.fused_encrypt and self.ring.fp64
                and self.N in (4096, 8192, 16384) and not self.ring.cyclic):
            seed_a = self._next_seed()
            seed_e = self._next_seed()
            out = torch.empty(B, 2, l, self.N, dtype=torch.uint64, device=self.device)
            call("ckks_encrypt_sk", self.ring.handle, _ptr(pt.contiguous()), _ptr(self.key_eval(sk, l)), _ptr(out),
                 B, l, float(self.p.sigma), seed_a, seed_e, self.ring.stream())
            return Ciphertext(out, self.p.scale if scale is None else scale)
        if a is None:
            a = self.uniform_device(B, l) if dev else self.uniform(B, limbs=range(l))
        if e is None:
            E = self.gaussian_device(B, l) if dev else self.small_to_rns(self.gaussian(B), range(l))
        else:
            E = self.small_to_rns(e, range(l))
        S = self.key_rows(sk, B, l)
        c0 = self.ring.add(self.ring.add(self.ring.poly_mul(a, S), E), pt)
        c1 = self.ring.neg(a)
        return Ciphertext(torch.stack([c0, c1], dim=1).contiguous(),
                          self.p.scale if scale is None else scale)

    def decrypt(self, ct: Ciphertext, sk: SecretKey) -> torch.Tensor:
        d = ct.data
        B, l = d.shape[0], d.shape[2]
        if self.fused_decrypt and self.ring.fp64 and self.N in (4096, 8192, 16384) and not self.ring.cyclic:
            out = torch.empty(B, l, self.N, dtype=torch.uint64, device=self.device)
            call("ckks_decrypt_sk", self.ring.handle, _ptr(d.contiguous()), _ptr(self.key_eval(sk, l)), _ptr(out),
                 B, l, self.ring.stream())
            return out
        S = self.key_rows(sk, B, l)
        return self.ring.add(d[:, 0].contiguous(), self.ring.poly_mul(d[:, 1].contiguous(), S))

    def key_eval(self, sk: SecretKey, l: int) -> torch.Tensor:
        """[l, N] s on primes 0..l-1 in evaluation order (ckks_ntt_fwd),
        kept on the key."""
        t = sk.rows.get(("eval", l))
        if t is None or t.device != self.device:
            t = self.ring.ntt(sk.rns[:l].to(self.device).contiguous())
            sk.rows[("eval", l)] = t
        return t

    def key_rows(self, sk: SecretKey, B: int, l: int) -> torch.Tensor:
        """[B, l, N] contiguous copies of s on primes 0..l-1, the second
        operand of encryption's and decryption's products.  Kept on the key
        per level and grown to the largest batch asked for, so that a call
        takes a prefix instead of a fresh copy.  The rows are marked in use
        on the calling stream, so a copy replaced by a larger one is not
        reused while another stream's kernels may still read it."""
        t = sk.rows.get(l)
        if t is None or t.shape[0] < B or t.device != self.device:
            t = sk.rns[:l].to(self.device).unsqueeze(0).expand(B, l, self.N).contiguous()
            sk.rows[l] = t
        if t.is_cuda:
            t.record_stream(torch.cuda.current_stream(t.device))
        return t[:B]

    # -- evaluation -----------------------------------------------------------------
    def workspace(self, batch: int, l: int, nbytes: Optional[int] = None) -> torch.Tensor:
        """Scratch for one call on the current stream (ckks_workspace_bytes,
        or nbytes).  Every stream has its own buffer, allocated while that
        stream is current: calls on different streams never share scratch,
        and a buffer replaced by a larger one is released in the order of
        the stream whose kernels used it."""
        if nbytes is None:
            nbytes = call("ckks_workspace_bytes", self.ring.handle, batch, l)
        words = (nbytes + 7) // 8
        key = torch.cuda.current_stream(self.device).cuda_stream
        ws = self._ws.get(key)
        if ws is None or ws.numel() < words:
            ws = torch.empty(words, dtype=torch.uint64, device=self.device)
            self._ws[key] = ws
        return ws

    def check_keys(self, *keys: SwitchKey):
        """Prepared keys are read in the arithmetic they were prepared for."""
        for k in keys:
            if k.fp64 != self.ring.fp64:
                raise ValueError("key prepared for the %s arithmetic, context set to %s: prepare it again"
                                 % ("FP64" if k.fp64 else "INT64", "FP64" if self.ring.fp64 else "INT64"))

    def hmult(self, a: Ciphertext, b: Ciphertext, rlk: SwitchKey) -> Ciphertext:
        """HMult + relinearise + rescale (fused; verify_e2e_mult.py:78-102)."""
        self.check_keys(rlk)
        if a.data.shape != b.data.shape:
            raise ValueError("ciphertext shapes differ")
        B, l = a.batch, a.level
        out = torch.empty(B, 2, l - 1, self.N, dtype=torch.uint64, device=self.device)
        ws = self.workspace(B, l)
        call("ckks_hmult", self.ring.handle, _ptr(a.data), _ptr(b.data), _ptr(rlk.prepared), self.L + 1,
             _ptr(out), B, l, self.sp, _ptr(ws), self.ring.stream())
        return Ciphertext(out, a.scale * b.scale / self.primes[l - 1])

    def relin(self, d0: torch.Tensor, d1: torch.Tensor, d2: torch.Tensor, rlk: SwitchKey) -> torch.Tensor:
        """(d0, d1, d2) [B, l, N] each -> [B, 2, l, N] = (d0, d1) + keyswitch(d2)
        (verify_keyswitch.py:62-68 generalised to the hybrid key)."""
        self.check_keys(rlk)
        B, l = d2.shape[0], d2.shape[1]
        d01 = torch.stack([d0, d1], dim=1).contiguous()
        out = torch.empty(B, 2, l, self.N, dtype=torch.uint64, device=self.device)
        ws = self.workspace(B, l)
        call("ckks_relin", self.ring.handle, _ptr(d01), _ptr(d2.contiguous()), _ptr(rlk.prepared), self.L + 1,
             _ptr(out), B, l, self.sp, _ptr(ws), self.ring.stream())
        return out

    def keyswitch(self, d2: torch.Tensor, key: SwitchKey) -> torch.Tensor:
        self.check_keys(key)
        B, l = d2.shape[0], d2.shape[1]
        out = torch.empty(B, 2, l, self.N, dtype=torch.uint64, device=self.device)
        ws = self.workspace(B, l)
        call("ckks_relin", self.ring.handle, None, _ptr(d2.contiguous()), _ptr(key.prepared), self.L + 1,
             _ptr(out), B, l, self.sp, _ptr(ws), self.ring.stream())
        return out

    def tensor(self, a: Ciphertext, b: Ciphertext):
        """(d0, d1, d2) in coefficient form."""
        A0, A1 = a.data[:, 0].contiguous(), a.data[:, 1].contiguous()
        B0, B1 = b.data[:, 0].contiguous(), b.data[:, 1].contiguous()
        r = self.ring
        return (r.poly_mul(A0, B0), r.add(r.poly_mul(A0, B1), r.poly_mul(A1, B0)), r.poly_mul(A1, B1))

    def rescale(self, ct: Ciphertext) -> Ciphertext:
        l = ct.level
        return Ciphertext(self.ring.rescale(ct.data), ct.scale / self.primes[l - 1])

    def rotate(self, ct: Ciphertext, gk: SwitchKey) -> Ciphertext:
        """Slot rotation by the step gk was made for."""
        self.check_keys(gk)
        B, l = ct.batch, ct.level
        out = torch.empty_like(ct.data)
        ws = self.workspace(B, l)
        call("ckks_rotate", self.ring.handle, _ptr(ct.data), _ptr(gk.prepared), self.L + 1, _ptr(out), B, l,
             self.sp, gk.galois, _ptr(ws), self.ring.stream())
        return Ciphertext(out, ct.scale)

    def rotate_add(self, ct: Ciphertext, gk: SwitchKey) -> Ciphertext:
        """ct + rotate(ct, gk) in one call (ckks_rotate_add): the same words
        as add(ct, rotate(ct, gk))."""
        self.check_keys(gk)
        B, l = ct.batch, ct.level
        out = torch.empty_like(ct.data)
        ws = self.workspace(B, l)
        call("ckks_rotate_add", self.ring.handle, _ptr(ct.data), _ptr(gk.prepared), self.L + 1, _ptr(out), B, l,
             self.sp, gk.galois, _ptr(ws), self.ring.stream())
        return Ciphertext(out, ct.scale)

    @staticmethod
    def rotation_tables(gks: Sequence[SwitchKey], device) -> Tuple[torch.Tensor, torch.Tensor]:
        """Device tables for rotate_many: the prepared key address and the
        Galois element of each ciphertext's rotation.  The keys must outlive
        every call that uses the tables."""
        keys = torch.tensor([g.prepared.data_ptr() for g in gks], dtype=torch.int64).to(device)
        gal = torch.tensor([g.galois for g in gks], dtype=torch.int64).to(device)
        return keys, gal

    def rotate_many(self, ct: Ciphertext, gks: Sequence[SwitchKey],
                    tables: Optional[Tuple[torch.Tensor, torch.Tensor]] = None,
                    out: Optional[torch.Tensor] = None) -> Ciphertext:
        """Ciphertext b rotated by the step gks[b] was made for, all in one
        call (ckks_rotate_many); tables may hold rotation_tables(gks) for
        reuse; out, if given, is a contiguous tensor of ct's shape to write."""
        B, l = ct.batch, ct.level
        if len(gks) != B:
            raise ValueError("one Galois key per ciphertext")
        self.check_keys(*gks)
        keys, gal = tables if tables is not None else self.rotation_tables(gks, self.device)
        if out is None:
            out = torch.empty_like(ct.data)
        elif out.shape != ct.data.shape or not out.is_contiguous():
            raise ValueError("out must be a contiguous tensor of the input's shape")
        ws = self.workspace(B, l)
        call("ckks_rotate_many", self.ring.handle, _ptr(ct.data), _ptr(keys), self.L + 1, _ptr(gal), _ptr(out), B,
             l, self.sp, _ptr(ws), self.ring.stream())
        return Ciphertext(out, ct.scale)

    def hoist_key(self, gk: SwitchKey) -> torch.Tensor:
        """The Galois key permuted by its inverse evaluation-domain
        automorphism, as ckks_rotate_hoisted reads it (cached on the key)."""
        if gk.hoisted is None:
            out = torch.empty_like(gk.prepared)
            rows = self.L * 2 * (self.L + 1)
            call("ckks_key_permute", self.ring.handle, _ptr(gk.prepared), _ptr(out), rows, gk.galois,
                 self.ring.stream())
            gk.hoisted = out
        return gk.hoisted

    def rotate_hoisted(self, ct: Ciphertext, gks: Sequence[SwitchKey], keys: Optional[torch.Tensor] = None,
                       out: Optional[torch.Tensor] = None, eval_form: bool = False) -> List[Ciphertext]:
        """Rotations of ct by every key of gks (at most 32) sharing one digit
        decomposition.  keys may hold the stacked hoisted keys (reused across
        calls); else they are stacked here.  out, if given, is a contiguous
        [R, B, 2, l, N] tensor to write into.  eval_form=True returns the
        rotations in evaluation order (ModDown done in that domain), ready
        for plaintext products."""
        B, l, R = ct.batch, ct.level, len(gks)
        if R == 0:
            return []
        self.check_keys(*gks)
        if keys is None:
            keys = torch.stack([self.hoist_key(g) for g in gks]).contiguous()
        if out is None:
            out = torch.empty(R, B, 2, l, self.N, dtype=torch.uint64, device=self.device)
        elif tuple(out.shape) != (R, B, 2, l, self.N) or not out.is_contiguous():
            raise ValueError("out must be a contiguous [R, B, 2, l, N] tensor")
        ws = self.workspace(B, l, call("ckks_workspace_hoisted_bytes", self.ring.handle, B, l, R))
        gal = (ctypes.c_uint64 * R)(*[int(g.galois) for g in gks])
        call("ckks_rotate_hoisted_eval" if eval_form else "ckks_rotate_hoisted", self.ring.handle, _ptr(ct.data),
             _ptr(keys), self.L + 1, gal, R, _ptr(out), B, l, self.sp, _ptr(ws), self.ring.stream())
        return [Ciphertext(out[r], ct.scale) for r in range(R)]

    def add(self, a: Ciphertext, b: Ciphertext) -> Ciphertext:
        return Ciphertext(self.ring.add(a.data, b.data), a.scale)

    def sub(self, a: Ciphertext, b: Ciphertext) -> Ciphertext:
        return Ciphertext(self.ring.sub(a.data, b.data), a.scale)

    def mul_plain(self, ct: Ciphertext, pt: torch.Tensor, pt_scale: float) -> Ciphertext:
        """ct * pt (pt [B, l, N] or [1, l, N] coefficient form), no rescale."""
        B, l = ct.batch, ct.level
        P = pt.expand(B, l, self.N).contiguous() if pt.shape[0] != B else pt
        c0 = self.ring.poly_mul(ct.data[:, 0].contiguous(), P)
        c1 = self.ring.poly_mul(ct.data[:, 1].contiguous(), P)
        return Ciphertext(torch.stack([c0, c1], 1).contiguous(), ct.scale * pt_scale)

    def add_plain(self, ct: Ciphertext, pt: torch.Tensor) -> Ciphertext:
        B, l = ct.batch, ct.level
        P = pt.expand(B, l, self.N).contiguous() if pt.shape[0] != B else pt
        return Ciphertext(torch.stack([self.ring.add(ct.data[:, 0].contiguous(), P), ct.data[:, 1]], 1).contiguous(),
                          ct.scale)

    def drop_levels(self, ct: Ciphertext, l: int) -> Ciphertext:
        """Keep the first l limbs (modulus reduction without division)."""
        return Ciphertext(ct.data[:, :, :l].contiguous(), ct.scale)


# -- serialisation (keys, ciphertexts) ------------------------------------------

def save(path: str, **objs) -> None:
    """Write ciphertexts / switch keys / secret keys to one .npz (raw uint64)."""
    out = {}
    for name, o in objs.items():
        if isinstance(o, Ciphertext):
            out[f"{name}.ct"] = to_numpy(o.data)
            out[f"{name}.scale"] = np.array([o.scale])
        elif isinstance(o, SwitchKey):
            out[f"{name}.ksk"] = to_numpy(o.coeff)
            out[f"{name}.galois"] = np.array([o.galois], dtype=np.uint64)
        elif isinstance(o, SecretKey):
            out[f"{name}.sk"] = o.s
        else:
            raise TypeError(f"cannot save {type(o).__name__}")
    np.savez(path, **out)


def load(path: str, ck: "CKKS") -> dict:
    """Inverse of ``save``; keys are prepared for ck's kernels."""
    d = np.load(path, allow_pickle=False)
    res = {}
    for k in d.files:
        name, kind = k.rsplit(".", 1)
        if kind == "ct":
            res[name] = Ciphertext(ck.ring.tensor(d[k]), float(d[name + ".scale"][0]))
        elif kind == "ksk":
            res[name] = ck.load_switch_key(d[k], int(d[name + ".galois"][0]))
        elif kind == "sk":
            s = d[k].astype(np.int64)
            res[name] = SecretKey(s=s, rns=ck.small_to_rns(s, range(len(ck.primes))))
    return res
