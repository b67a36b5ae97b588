"""Writes tests/golden/golden_vectors.npz from the CPU oracle.

Inputs are drawn from numpy PCG64 with the seeds below; expected outputs
come from the oracle (oracle/ckks_oracle_c.c, itself pinned by
tests/test_oracle_pinning.py).  Large outputs are stored as SHA-256 digests
of their little-endian uint64 bytes.  Run from the repository root:

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import ckks_oracle as O  # noqa: E402
import oracle_c as OC  # noqa: E402

CHAIN4 = [17592186634241, 17592186372097, 17592186273793, 17592186175489]
SPECIAL_P = 17592186765313
Q1, PSI1 = 35184371884033, 7388705119


def digest(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<u8").tobytes()).hexdigest()


def hmult_case(N: int, l: int, seed: int):
    """Seeded RLWE-shaped inputs and a hybrid key in coefficient form, all
    drawn here with PCG64 so the GPU test can rebuild them bit-for-bit."""
    rng = np.random.default_rng(seed)
    primes = CHAIN4[:l]
    basis = CHAIN4 + [SPECIAL_P]
    roots = [O.root_of_unity(q, 2 * N) for q in basis]
    s = rng.integers(-1, 2, N)
    key = np.zeros((4, 2, 5, N), dtype=np.uint64)
    for j in range(4):
        e = rng.integers(-5, 6, N)
        for i, q in enumerate(basis):
            a = rng.integers(0, q, N, dtype=np.uint64)
            sr = np.where(s < 0, q + s, s).astype(np.uint64)
            b = OC.polymul_ntt(a, sr, q, roots[i])
            b = (b + np.where(e < 0, q + e, e).astype(np.uint64)) % np.uint64(q)
            if i == j:
                s2 = OC.polymul_ntt(sr, sr, q, roots[i])
                b = np.array([(int(x) + (SPECIAL_P % q) * int(y)) % q for x, y in zip(b, s2)], dtype=np.uint64)
            key[j, 0, i], key[j, 1, i] = b, a
    cta = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes]) for _ in range(2)])
    ctb = np.stack([np.stack([rng.integers(0, q, N, dtype=np.uint64) for q in primes]) for _ in range(2)])
    out = OC.hmult(cta, ctb, key, basis[:4], SPECIAL_P, roots[:4], roots[4])
    return s, key, cta, ctb, out


def main():
    g = {}
    rng = np.random.default_rng(20261203)
    a = rng.integers(0, Q1, 4096, dtype=np.uint64)
    b = rng.integers(0, Q1, 4096, dtype=np.uint64)
    g["cfg1_a"], g["cfg1_b"] = a, b
    g["cfg1_c"] = OC.polymul_ntt(a, b, Q1, PSI1)
    assert np.array_equal(g["cfg1_c"], OC.polymul_schoolbook(a, b, Q1))
    x = rng.integers(0, CHAIN4[0], 8192, dtype=np.uint64)
    psi0 = O.root_of_unity(CHAIN4[0], 16384)
    g["ntt8192_q0_in"] = x
    g["ntt8192_q0_out"] = OC.ntt_negacyclic(x, 8192, CHAIN4[0], psi0)
    limbs = np.stack([rng.integers(0, q, 64, dtype=np.uint64) for q in CHAIN4])
    g["rescale_in"] = limbs
    g["rescale_out"] = np.array(O.rns_rescale([list(map(int, r)) for r in limbs], CHAIN4), dtype=np.uint64)
    s, key, cta, ctb, out = hmult_case(8192, 4, 777)
    g["hmult8192_seed"] = np.array([777])
    g["hmult8192_out_sha256"] = np.frombuffer(bytes.fromhex(digest(out)), dtype=np.uint8)
    path = os.path.join(ROOT, "tests", "golden", "golden_vectors.npz")
    np.savez_compressed(path, **g)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
