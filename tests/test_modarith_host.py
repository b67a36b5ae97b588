"""Device integer code compiled for the host.  The modular arithmetic (modarith.h)
checked against exact 128-bit arithmetic: Shoup products (exact and lazy
quotient), the lazy high half, Barrett, 64-bit reduction, the special-form
product and both lazy butterflies on their documented bounds, for the chain
primes, the special primes and the single-prime product's 45-bit prime
(tests/modarith_check.cpp).
The checker itself was seen to fail on a planted error in the lazy high half."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "cipher-pu_amd", "csrc")


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    pytest.skip("hipcc not found")


def test_modarith_against_128_bit_arithmetic(tmp_path):
    exe = str(tmp_path / "modarith_check")
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "modarith_check.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe, "200000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith(", 0 failures"), r.stdout


def test_hoist_placement_is_a_bijection(tmp_path):
    """hoist_slot (launch.h) maps every workgroup of the hoisted key switch to
    a distinct (target, rotation, ciphertext) for every tile size and XCD
    setting (tests/hoist_order_check.cpp).  The checker was seen to fail when
    the rotation index of a ragged last tile was divided by the tile size
    instead of the tile's length."""
    exe = str(tmp_path / "hoist_order_check")
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "hoist_order_check.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith(", 0 failures"), r.stdout


def test_fp64_residue_arithmetic_against_128_bit_arithmetic(tmp_path):
    """The FP64 residue arithmetic (modf64.h: the fixed-operand and data
    products, the two reductions, the word conversions and both butterflies)
    on its documented input bounds, against exact 128-bit integer arithmetic
    (tests/modf64_check.cpp).  The checker was seen to fail when the product's
    low word was dropped, when canon's offset was raised to 2^-42, and on
    fixed-operand inputs beyond 2^51."""
    exe = str(tmp_path / "modf64_check")
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "modf64_check.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe, "200000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith(", 0 failures"), r.stdout


def test_lds_exchange_schedules_race_free(tmp_path):
    """Every LDS exchange schedule of the FP64 transforms (single, paired and
    software-pipelined, double-buffered, as the kernels chain them) for every
    built launch shape: between two workgroup barriers no LDS slot is touched
    by two waves with a write among them (tests/lds_schedule_check.cpp, with
    the headers' own layouts, paddings and barrier predicates).  The checker
    was seen to report the race of a pipelined pair without the barrier
    before the second polynomial's write at N = 4096 with 512 threads (a
    wave-local exchange that changes the padding), and a double-buffered
    transform without its barrier after a wave-local exchange."""
    exe = str(tmp_path / "lds_schedule_check")
    subprocess.run([_hipcc(), "-O2", "-std=c++17", "-I", CSRC, os.path.join(HERE, "lds_schedule_check.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith(", 0 failures"), r.stdout
