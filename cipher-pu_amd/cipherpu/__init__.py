"""cipherpu: MI355X (gfx950) CKKS evaluation core.

Hand-written HIP kernels behind the C ABI in include/ckks_core.h, bound
with ctypes; ciphertexts are device-resident uint64 tensors [B, 2, l, N].

  cipherpu.golden_api  drop-ins for the reference's golden-model functions
  cipherpu.ring        RingContext: transforms and RNS arithmetic
  cipherpu.ckks        CKKS scheme: keys, encryption, HMult, rotation
  cipherpu.params      parameter sets and the reference's file formats
"""
from . import _lib
from ._lib import CkksError, LIB_PATH

__all__ = ["CkksError", "LIB_PATH", "lib_loaded"]


def lib_loaded() -> bool:
    return _lib._lib is not None
