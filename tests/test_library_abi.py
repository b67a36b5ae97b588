"""CPU-side checks of the boundary: the HIP library loads without a GPU and
exports every entry point include/ckks_core.h declares."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ckks_core.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|uint64_t|const char \*)\s*\*?\s*(ckks_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "ckks_hmult" in names and "ckks_ntt_fwd" in names and len(names) >= 20


def test_library_exports_every_declared_symbol():
    from cipherpu import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in _declared():
        assert hasattr(L, name), name
    assert set(_declared()) == set(_lib.EXPORTED)
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    for name in _declared():
        assert re.search(rf"\bT {name}$", out, re.M), name


def test_version_string_without_gpu():
    from cipherpu import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    assert b"gfx950" in _lib.load().ckks_version()


def test_errors_without_gpu_are_reported_not_faked():
    """With no GPU the context cannot be created: the call returns an error
    code (never a CPU fallback)."""
    import torch
    from cipherpu import _lib
    if torch.cuda.is_available() or not os.path.exists(_lib.LIB_PATH):
        pytest.skip("only meaningful without a GPU")
    h = ctypes.c_void_p()
    pr = (ctypes.c_uint64 * 1)(17592186634241)
    rc = _lib.load().ckks_ctx_create(ctypes.byref(h), 8192, 1, pr, None, 0, 0)
    assert rc == -8
    from cipherpu.ring import RingContext
    with pytest.raises(RuntimeError):
        RingContext(8192, [17592186634241])


def test_argument_validation_without_gpu():
    from cipherpu import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.load()
    h = ctypes.c_void_p()
    bad_n = (ctypes.c_uint64 * 1)(17592186634241)
    assert L.ckks_ctx_create(ctypes.byref(h), 3000, 1, bad_n, None, 0, 0) == -1
    not_prime = (ctypes.c_uint64 * 1)(17592186634243)
    assert L.ckks_ctx_create(ctypes.byref(h), 8192, 1, not_prime, None, 0, 0) == -2
    too_big = (ctypes.c_uint64 * 1)((1 << 45) + 32769)
    assert L.ckks_ctx_create(ctypes.byref(h), 8192, 1, too_big, None, 0, 0) == -2
    assert L.ckks_hmult(None, None, None, None, 5, None, 1, 4, 4, None, None) == -6
