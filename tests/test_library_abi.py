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


def test_gaussian_table_against_closed_form():
    """The sampler's table: cdf[k] = 2^64 P(|round(Z)| <= k), Z ~ N(0, 3.2^2),
    against scipy's normal CDF (host-only entry point, no GPU)."""
    import ctypes
    from scipy.special import erfc
    from cipherpu import _lib
    L = _lib.load()
    n = 48
    buf = (ctypes.c_uint64 * n)()
    assert L.ckks_gaussian_cdf(ctypes.c_double(3.2), buf, n) == 0
    cdf = [int(v) for v in buf]
    for k in range(n):
        want = 2.0 ** 64 * (1.0 - erfc((k + 0.5) / (3.2 * 2 ** 0.5)))
        assert abs(cdf[k] - want) <= 2.0 ** 64 * 1e-13 + 1, k
    assert cdf == sorted(cdf) and cdf[-1] == 2 ** 64 - 1
    # P(e = 0) for sigma = 3.2 is erf(0.5 / (3.2 sqrt 2)) = 0.124164...
    assert abs(cdf[0] / 2.0 ** 64 - 0.124164) < 1e-5
    assert L.ckks_gaussian_cdf(ctypes.c_double(0.0), buf, n) == -7
    assert L.ckks_gaussian_cdf(ctypes.c_double(3.2), buf, 49) == -7
