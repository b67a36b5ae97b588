"""ctypes binding of the C oracle (oracle/ckks_oracle_c.c) -- TEST
INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker, never by the product."""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liboracle_c.so")
_lib = None

_P = ctypes.c_void_p
_U = ctypes.c_uint64
_I = ctypes.c_int


def build(force: bool = False) -> str:
    if force or not os.path.exists(_SO) or (
        os.path.getmtime(_SO) < os.path.getmtime(os.path.join(_HERE, "ckks_oracle_c.c"))
    ):
        subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle_c.so"], check=True)
    return _SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = ctypes.CDLL(_SO)
        sig = {
            "oracle_ntt_cyclic": [_P, _U, _U, _U],
            "oracle_intt_cyclic": [_P, _U, _U, _U],
            "oracle_ntt_negacyclic": [_P, _U, _U, _U],
            "oracle_intt_negacyclic": [_P, _U, _U, _U],
            "oracle_polymul_schoolbook": [_P, _P, _P, _U, _U],
            "oracle_polymul_ntt": [_P, _P, _P, _U, _U, _U],
            "oracle_rescale": [_P, _P, _U, _I, _P],
            "oracle_mod_down": [_P, _P, _U, _I, _P, _U],
            "oracle_keyswitch": [_P, _P, _P, _P, _U, _I, _I, _P, _U, _P, _P],
            "oracle_hmult": [_P, _P, _P, _P, _U, _I, _I, _P, _U, _P, _P],
            "oracle_automorphism": [_P, _P, _U, _U, _U],
            "oracle_rotate": [_P, _P, _P, _U, _I, _I, _P, _U, _P, _P, _U],
            "oracle_rotate_hoisted": [_P, _P, _P, _U, _I, _I, _P, _U, _P, _P, _U],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        _lib = L
    return _lib


def _u64(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def ntt_cyclic(a, N, q, omega):
    x = _u64(a).copy()
    lib().oracle_ntt_cyclic(_ptr(x), N, q, omega)
    return x


def intt_cyclic(a, N, q, omega_inv):
    x = _u64(a).copy()
    lib().oracle_intt_cyclic(_ptr(x), N, q, omega_inv)
    return x


def ntt_negacyclic(a, N, q, psi):
    x = _u64(a).copy()
    lib().oracle_ntt_negacyclic(_ptr(x), N, q, psi)
    return x


def intt_negacyclic(a, N, q, psi):
    x = _u64(a).copy()
    lib().oracle_intt_negacyclic(_ptr(x), N, q, psi)
    return x


def polymul_schoolbook(a, b, q):
    a, b = _u64(a), _u64(b)
    out = np.empty_like(a)
    lib().oracle_polymul_schoolbook(_ptr(a), _ptr(b), _ptr(out), a.size, q)
    return out


def polymul_ntt(a, b, q, psi):
    a, b = _u64(a), _u64(b)
    out = np.empty_like(a)
    lib().oracle_polymul_ntt(_ptr(a), _ptr(b), _ptr(out), a.size, q, psi)
    return out


def rescale(limbs, primes):
    x = _u64(limbs)
    l, N = x.shape
    pr = _u64(primes)
    out = np.empty((l - 1, N), dtype=np.uint64)
    lib().oracle_rescale(_ptr(x), _ptr(out), N, l, _ptr(pr))
    return out


def mod_down(limbs_qp, primes, p):
    x = _u64(limbs_qp)
    nb, N = x.shape
    pr = _u64(primes)
    out = np.empty((nb - 1, N), dtype=np.uint64)
    lib().oracle_mod_down(_ptr(x), _ptr(out), N, nb - 1, _ptr(pr), p)
    return out


def keyswitch(d2, key, primes, p, psis, psi_p):
    """d2 [l][N]; key [l_key][2][key_limbs][N] (coefficient form)."""
    d2, key = _u64(d2), _u64(key)
    l, N = d2.shape
    key_limbs = key.shape[2]
    pr, ps, pp = _u64(primes[:l]), _u64(psis[:l]), _u64([psi_p])
    o0 = np.empty((l, N), dtype=np.uint64)
    o1 = np.empty((l, N), dtype=np.uint64)
    lib().oracle_keyswitch(_ptr(d2), _ptr(key), _ptr(o0), _ptr(o1), N, l, key_limbs,
                           _ptr(pr), p, _ptr(ps), _ptr(pp))
    return o0, o1


def hmult(cta, ctb, key, primes, p, psis, psi_p):
    """cta, ctb [2][l][N]; key [>=l][2][key_limbs][N]; returns [2][l-1][N]."""
    cta, ctb, key = _u64(cta), _u64(ctb), _u64(key)
    _, l, N = cta.shape
    key_limbs = key.shape[2]
    pr, ps, pp = _u64(primes[:l]), _u64(psis[:l]), _u64([psi_p])
    out = np.empty((2, l - 1, N), dtype=np.uint64)
    lib().oracle_hmult(_ptr(cta), _ptr(ctb), _ptr(key), _ptr(out), N, l, key_limbs,
                       _ptr(pr), p, _ptr(ps), _ptr(pp))
    return out


def automorphism(a, q, g):
    a = _u64(a)
    out = np.empty_like(a)
    lib().oracle_automorphism(_ptr(a), _ptr(out), a.size, q, g)
    return out


def rotate(ct, key, primes, p, psis, psi_p, g):
    ct, key = _u64(ct), _u64(key)
    _, l, N = ct.shape
    key_limbs = key.shape[2]
    pr, ps, pp = _u64(primes[:l]), _u64(psis[:l]), _u64([psi_p])
    out = np.empty((2, l, N), dtype=np.uint64)
    lib().oracle_rotate(_ptr(ct), _ptr(key), _ptr(out), N, l, key_limbs, _ptr(pr), p,
                        _ptr(ps), _ptr(pp), g)
    return out


def rotate_hoisted(ct, key, primes, p, psis, psi_p, g):
    """The hoisted rotation's reference: digits = automorphism over Z of the
    canonical lifts of c1's limbs."""
    ct, key = _u64(ct), _u64(key)
    _, l, N = ct.shape
    key_limbs = key.shape[2]
    pr, ps, pp = _u64(primes[:l]), _u64(psis[:l]), _u64([psi_p])
    out = np.empty((2, l, N), dtype=np.uint64)
    lib().oracle_rotate_hoisted(_ptr(ct), _ptr(key), _ptr(out), N, l, key_limbs, _ptr(pr), p,
                                _ptr(ps), _ptr(pp), g)
    return out
