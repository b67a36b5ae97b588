"""Parameter sets and the reference's on-disk parameter formats.

CHAIN4: the four 45-bit primes of rtl/m2/modulus_chain_chain4.json (all
1 mod 2^15).  SPECIAL_P: the next prime of the same form
2^44 + 32768 j + 1 (j = 22) above the chain, used as the hybrid key-switch
special prime.  CONFIG1_Q / CONFIG1_PSI: the largest 45-bit prime that is
1 mod 8192 and its smallest primitive 8192-th root (N = 4096 product).
CHAIN5: CHAIN4 and the next prime of the form 2^44 + 32768 j + 1 above the
special prime (j = 48), for ciphertexts with four rescales left."""
from __future__ import annotations

import json
from typing import List

CHAIN4 = (17592186634241, 17592186372097, 17592186273793, 17592186175489)
SPECIAL_P = 17592186765313
CHAIN5 = CHAIN4 + (17592187617281,)
CONFIG1_Q = 35184371884033
CONFIG1_PSI = 7388705119


def load_chain_json(path: str) -> List[dict]:
    """Modulus chain in the reference's JSON schema
    (rtl/m2/modulus_chain_chain4.json: list of {q, primitive_root, omega,
    omega_inv, N, N_INV, ...})."""
    with open(path, encoding="utf-8") as f:
        chain = json.load(f)
    for e in chain:
        q, w, N = int(e["q"]), int(e["omega"]), int(e["N"])
        if pow(w, N, q) != 1 or w * int(e["omega_inv"]) % q != 1 or N * int(e["N_INV"]) % q != 1:
            raise ValueError(f"inconsistent chain entry for q={q}")
    return chain


def chain_primes_and_roots(chain: List[dict], N: int):
    """(primes, roots) of a loaded chain for the negacyclic ring of degree
    N: each entry's omega is a primitive entry["N"]-th root, and the ring
    needs one of order 2N, omega^(entry["N"] / 2N) (the chain file's
    N = 16384 gives psi itself at N = 8192)."""
    primes, roots = [], []
    for e in chain:
        q, w, n = int(e["q"]), int(e["omega"]), int(e["N"])
        if n % (2 * N):
            raise ValueError(f"chain root of order {n} has no power of order {2 * N}")
        primes.append(q)
        roots.append(pow(w, n // (2 * N), q))
    return primes, roots


def default_root(q: int, order: int) -> int:
    """g^((q - 1) / order) for the smallest generator g of Z_q^*: the root
    the library derives when given none (scripts/gen_modulus_chain.py:128-130
    builds the chain's roots the same way)."""
    n, fs, d = q - 1, [], 2
    while d * d <= n:
        if n % d == 0:
            fs.append(d)
            while n % d == 0:
                n //= d
        d += 1 if d == 2 else 2
    if n > 1:
        fs.append(n)
    g = 2
    while any(pow(g, (q - 1) // f, q) == 1 for f in fs):
        g += 1
    return pow(g, (q - 1) // order, q)


def root_from_rom(values: List[int], q: int) -> int:
    """The root a twiddle ROM tabulates (values[k] = w^k,
    scripts/gen_omega_tables.py:96-106), checked entry by entry."""
    w = int(values[1]) % q
    acc = 1
    for k, v in enumerate(values):
        if int(v) != acc:
            raise ValueError(f"ROM entry {k} is not the {k}-th power of {w} mod {q}")
        acc = acc * w % q
    return w


def save_chain_json(path: str, entries: List[dict]) -> None:
    with open(path, "w", encoding="utf-8") as f:
        json.dump(entries, f, indent=2)


def read_hex_table(path: str) -> List[int]:
    """$readmemh twiddle ROM: one hex word per line
    (scripts/gen_omega_tables.py:244-247)."""
    with open(path) as f:
        return [int(l.strip(), 16) for l in f if l.strip() and not l.lstrip().startswith("//")]


def write_hex_table(path: str, values, width: int = 12) -> None:
    with open(path, "w") as f:
        for v in values:
            f.write(f"{int(v):0{width}x}\n")
