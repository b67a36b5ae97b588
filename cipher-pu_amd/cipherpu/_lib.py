"""ctypes binding of libcipherpu_hip.so (include/ckks_core.h).

There is no CPU fallback: if the library is missing, or no GPU is visible,
every entry point raises.  Build with ``python -m cipherpu.build``."""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcipherpu_hip.so")

CKKS_CYCLIC = 1
CKKS_NATURAL_ORDER = 2
CKKS_ARITH_INT64 = 0
CKKS_ARITH_FP64 = 1

_ERRORS = {
    -1: "bad ring degree N",
    -2: "bad modulus (not an odd prime < 2^45, or no root of the needed order)",
    -3: "device pointer not 8-byte aligned",
    -4: "bad shape",
    -5: "HIP runtime error",
    -6: "NULL pointer",
    -7: "bad argument",
    -8: "no usable GPU device",
}


class CkksError(RuntimeError):
    def __init__(self, fn: str, code: int):
        super().__init__(f"{fn}: {_ERRORS.get(code, 'error')} ({code})")
        self.code = code


_P = ctypes.c_void_p
_U = ctypes.c_uint64
_I = ctypes.c_int
_D = ctypes.c_double

_SIGS = {
    "ckks_ctx_create": (_I, [ctypes.POINTER(_P), _U, _I, _P, _P, _I, _I]),
    "ckks_ctx_destroy": (_I, [_P]),
    "ckks_ctx_root": (_I, [_P, _I, ctypes.POINTER(_U)]),
    "ckks_ctx_set_threads": (_I, [_P, _I, _I, _I]),
    "ckks_ctx_set_front": (_I, [_P, _I]),
    "ckks_ctx_set_arith": (_I, [_P, _I]),
    "ckks_ctx_set_ks_order": (_I, [_P, _I]),
    "ckks_ctx_set_hoist_order": (_I, [_P, _I, _I]),
    "ckks_version": (ctypes.c_char_p, []),
    "ckks_ntt_fwd": (_I, [_P, _P, _I, _I, _I, _I, _P]),
    "ckks_ntt_inv": (_I, [_P, _P, _I, _I, _I, _I, _P]),
    "ckks_poly_mul": (_I, [_P, _P, _P, _P, _I, _I, _I, _P]),
    "ckks_pointwise": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _P]),
    "ckks_rescale": (_I, [_P, _P, _P, _I, _I, _I, _P]),
    "ckks_workspace_bytes": (_U, [_P, _I, _I]),
    "ckks_key_prepare": (_I, [_P, _P, _P, _I, _I, _I, _P]),
    "ckks_relin": (_I, [_P, _P, _P, _P, _I, _P, _I, _I, _I, _P, _P]),
    "ckks_hmult": (_I, [_P, _P, _P, _P, _I, _P, _I, _I, _I, _P, _P]),
    "ckks_hmult_noalloc": (_I, [_P, _P, _P, _P, _I, _P, _I, _I, _I, _P]),
    "ckks_automorphism": (_I, [_P, _P, _P, _I, _I, _I, _U, _P]),
    "ckks_rotate": (_I, [_P, _P, _P, _I, _P, _I, _I, _I, _U, _P, _P]),
    "ckks_rotate_add": (_I, [_P, _P, _P, _I, _P, _I, _I, _I, _U, _P, _P]),
    "ckks_rotate_many": (_I, [_P, _P, _P, _I, _P, _P, _I, _I, _I, _P, _P]),
    "ckks_encode": (_I, [_P, _P, _P, _I, _I, _I, _D, _P, _P]),
    "ckks_decode": (_I, [_P, _P, _P, _I, _I, _I, _D, _P, _P]),
    "ckks_bitrev": (_I, [_P, _I, _U, _P]),
    "ckks_mod_switch": (_I, [_P, _P, _U, _U, _U, _P]),
    "ckks_mod_switch_scaled": (_I, [_P, _P, _U, _U, _U, _I, _P]),
    "ckks_poly_mul_modq": (_I, [_P, _P, _P, _P, _I, _U, _P, _P]),
    "ckks_noise_stats_q": (_I, [_P, _P, _I, _U, _U, _P]),
    "ckks_probe_dft64": (_I, [_P, _P, _P, _U, _U, _U, _I, _I, _P]),
    "ckks_embed": (_I, [_P, _P, _P, _I, _I, _P]),
    "ckks_embed_inv": (_I, [_P, _P, _P, _I, _I, _P]),
    "ckks_noise_stats": (_I, [_P, _P, _P, _I, _I, _I, _P]),
    "ckks_pt_prepare": (_I, [_P, _P, _P, _I, _I, _I, _P]),
    "ckks_lincomb": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _P]),
    "ckks_lincomb_scalar": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _P]),
    "ckks_lift_small": (_I, [_P, _P, _P, _I, _I, _I, _P]),
    "ckks_gear_dense_fwd": (_I, [_P, _I, _P, _P, _P, _I, _I, _I, _I, _I, _P]),
    "ckks_gear_dense_bwd": (_I, [_P, _P, _I, _P, _I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _P]),
    "ckks_gear_adam": (_I, [_P, _P, _P, _P, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                            _I, _P]),
    "ckks_gear_mse_grad": (_I, [_P, _P, _P, _P, _I, _I, _P]),
    "ckks_sample_uniform": (_I, [_P, _P, _I, _I, _I, _U, _P]),
    "ckks_sample_gaussian": (_I, [_P, _P, _I, _I, _I, ctypes.c_double, _U, _P]),
    "ckks_encrypt_sk": (_I, [_P, _P, _P, _P, _I, _I, ctypes.c_double, _U, _U, _P]),
    "ckks_decrypt_sk": (_I, [_P, _P, _P, _P, _I, _I, _P]),
    "ckks_gaussian_cdf": (_I, [ctypes.c_double, _P, _I]),
    "ckks_workspace_hoisted_bytes": (_U, [_P, _I, _I, _I]),
    "ckks_rotate_hoisted": (_I, [_P, _P, _P, _I, _P, _I, _P, _I, _I, _I, _P, _P]),
    "ckks_rotate_hoisted_eval": (_I, [_P, _P, _P, _I, _P, _I, _P, _I, _I, _I, _P, _P]),
    "ckks_key_permute": (_I, [_P, _P, _P, _I, _U, _P]),
    "ckks_timing_start": (_I, [_P, _I]),
    "ckks_timing_stop": (_I, [_P, _P, _P, _I]),
}

# Kernel kinds of ckks_timing_stop (include/ckks_core.h CKKS_KIND_*)
KERNEL_KINDS = ("ntt", "tensor", "front", "keyswitch", "moddown", "automorphism", "polymul", "keyswitch_fused")

EXPORTED = tuple(_SIGS)

_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the HIP library (no GPU call is made).  torch is imported first so
    that the library binds to the HIP runtime torch already carries (same
    soname) instead of loading a second runtime into the process."""
    global _lib
    if _lib is not None:
        return _lib
    import torch  # noqa: F401  (HIP runtime first)
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP library first (python -m cipherpu.build); "
            "there is no CPU fallback")
    L = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def call(name: str, *args) -> int:
    rc = getattr(load(), name)(*args)
    if isinstance(rc, int) and rc < 0 and name != "ckks_workspace_bytes":
        raise CkksError(name, rc)
    return rc
