"""The C ABI's host code under AddressSanitizer (CPU; no GPU needed).

tests/capi_host_check.cpp includes cipher-pu_amd/csrc/capi.cpp and checks
the argument checking of the entry points, the context's host tables (built
for every ring degree before the device is looked for) and the ModDown /
CRT constants against 128-bit arithmetic.  It is compiled by hipcc with
-fsanitize=address on the host code only (-Xarch_host) and linked with the
library's kernel objects, built without a sanitizer."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "cipher-pu_amd")
OUT = os.path.join(PKG, "_build", "asan")


def _kernel_objects():
    """The library's kernel objects from the last build (compiled again when
    missing or older than their sources)."""
    sys.path.insert(0, PKG)
    from cipherpu import build as B
    objs = []
    for src in B.SOURCES:
        if src == "capi.cpp":
            continue
        obj = os.path.join(B.BUILD, os.path.splitext(src)[0] + ".o")
        stale = not os.path.exists(obj) or any(
            os.path.getmtime(os.path.join(B.CSRC, f)) > os.path.getmtime(obj)
            for f in os.listdir(B.CSRC) if f.endswith((".h", ".hip")))
        if stale:
            os.makedirs(B.BUILD, exist_ok=True)
            subprocess.run([B.hipcc(), *B.flags(src), "-x", "hip", "-c", os.path.join(B.CSRC, src), "-o", obj],
                           check=True, capture_output=True)
        objs.append(obj)
    return B, objs


@pytest.mark.skipif(not (shutil.which("hipcc") or os.path.exists("/opt/rocm/bin/hipcc")), reason="needs hipcc")
def test_capi_host_code_under_asan():
    B, objs = _kernel_objects()
    os.makedirs(OUT, exist_ok=True)
    exe = os.path.join(OUT, "capi_host_check")
    src = os.path.join(ROOT, "tests", "capi_host_check.cpp")
    flags = [f for f in B.flags() if f not in ("-O3", "-fPIC")]
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer", "-g", "-O1"]
    r = subprocess.run([B.hipcc(), *flags, *san, "-x", "hip", src, "-x", "none", *objs, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", HIP_VISIBLE_DEVICES="")
    run = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    assert run.returncode == 0, run.stdout[-3000:] + run.stderr[-5000:]
    assert "capi host checks passed" in run.stdout
