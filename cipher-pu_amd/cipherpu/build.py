"""Builds the HIP library libcipherpu_hip.so in-tree for gfx950.

    python -m cipherpu.build        (from cipher-pu_amd/)

Each translation unit is compiled by hipcc in parallel, then linked into a
shared library next to this file.  Compiles every time it is called."""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
CSRC = os.path.join(PKG_ROOT, "csrc")
INCLUDE = os.path.join(os.path.dirname(PKG_ROOT), "include")
BUILD = os.path.join(PKG_ROOT, "_build")
LIB = os.path.join(HERE, "libcipherpu_hip.so")
SOURCES = ["k_ntt.hip", "k_ckks.hip", "k_ckks_int.hip", "k_fft.hip", "k_gear.hip", "k_sample.hip", "k_hoist.hip",
           "k_mfma_probe.hip", "capi.cpp"]
ARCH = os.environ.get("CIPHERPU_ARCH", "gfx950")


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


# The matrix-core probe keeps its int32 accumulators in VGPRs (the VGPR form
# of the MFMA instructions): in AGPRs every recombined sum would first need
# a v_accvgpr_read.  No other source uses the matrix cores.
SOURCE_FLAGS = {"k_mfma_probe.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def flags(src: str | None = None):
    return [f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I", CSRC, "-I", INCLUDE, *SOURCE_FLAGS.get(src, [])]


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    cc = hipcc()
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmd = [cc, *flags(src), "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError(f"HIP build failed:\n{msg}")
    tmp = LIB + ".tmp"
    subprocess.run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
