"""Extract the reference's known-answer data into a small committed fixture.

Reads (as data, never as code) from the reference checkout:
  tb/tb_ntt_n8.sv, tb/tb_pm_n8.sv, tb/tb_pm_n16.sv   known-answer vectors
  rtl/m2/modulus_chain_chain4.json                    prime chain
  rtl/m2/omega*_table_*.hex, rtl/m5/tw/*.hex          $readmemh twiddle ROMs

Writes tests/golden/reference_fixtures.json.  Tables of at most 64 entries
are stored whole; larger ones as (count, head, tail, sha256 of the values
packed little-endian uint64), so a test can rebuild a table and compare its
digest.  Run:  python tests/golden/make_reference_fixtures.py [REF_ROOT]
"""
from __future__ import annotations

import glob
import hashlib
import json
import os
import re
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fixtures.json")


def ints_in(text: str):
    return [int(x) for x in re.findall(r"\d+", text)]


def read_hex(path: str):
    with open(path) as f:
        return [int(l.strip(), 16) for l in f if l.strip() and not l.startswith("//")]


def digest(vals):
    return hashlib.sha256(np.asarray(vals, dtype="<u8").tobytes()).hexdigest()


def table_entry(vals):
    e = {"count": len(vals), "sha256": digest(vals)}
    if len(vals) <= 64:
        e["values"] = vals
    else:
        e["head"] = vals[:4]
        e["tail"] = vals[-4:]
    return e


def main():
    fx = {"source": "YI-HAIBO/cipher-pu", "kats": {}, "tables": {}}

    # tb_ntt_n8.sv: expected DFT in the header comment and the gold8 array
    t = open(os.path.join(REF, "tb/tb_ntt_n8.sv"), encoding="utf-8").read()
    m = re.search(r"gold8\s*\[0:7\]\s*=\s*'\{([^}]*)\}", t)
    fx["kats"]["ntt_n8"] = {
        "file": "tb/tb_ntt_n8.sv", "q": 113, "N": 8,
        "input": list(range(8)), "expected": ints_in(m.group(1)),
        "bitrev_input": ints_in(re.search(r"bit-reversal[^\[]*\[([^\]]*)\]", t).group(1)),
    }
    # tb_pm_n8.sv: a = 0..7, b = 1,3,..,15, cyclic product
    t = open(os.path.join(REF, "tb/tb_pm_n8.sv"), encoding="utf-8").read()
    exp = ints_in(re.search(r"expected\s+\[([^\]]*)\]", t).group(1))
    fx["kats"]["pm_n8"] = {
        "file": "tb/tb_pm_n8.sv", "q": 113, "N": 8, "ring": "cyclic",
        "a": list(range(8)), "b": [2 * i + 1 for i in range(8)], "expected": exp,
    }
    # tb_pm_n16.sv: a = 2i+1, b = 4i+3 (mod 113), expected[] assignments
    t = open(os.path.join(REF, "tb/tb_pm_n16.sv"), encoding="utf-8").read()
    exp16 = [0] * 16
    for k, v in re.findall(r"expected\[(\d+)\]\s*=\s*(\d+);", t):
        exp16[int(k)] = int(v)
    fx["kats"]["pm_n16"] = {
        "file": "tb/tb_pm_n16.sv", "q": 113, "N": 16, "ring": "cyclic",
        "a": [(2 * i + 1) % 113 for i in range(16)], "b": [(4 * i + 3) % 113 for i in range(16)],
        "expected": exp16,
        "n_inv": int(re.search(r"N_INV\s*=\s*7'd(\d+)", t).group(1)),
    }

    with open(os.path.join(REF, "rtl/m2/modulus_chain_chain4.json"), encoding="utf-8") as f:
        fx["chain4"] = json.load(f)

    for path in sorted(glob.glob(os.path.join(REF, "rtl/m2/*.hex")) +
                       glob.glob(os.path.join(REF, "rtl/m5/tw/*.hex"))):
        rel = os.path.relpath(path, REF)
        fx["tables"][rel] = table_entry(read_hex(path))

    with open(OUT, "w") as f:
        json.dump(fx, f, indent=1)
    print(f"wrote {OUT}: {len(fx['kats'])} KATs, {len(fx['tables'])} tables")


if __name__ == "__main__":
    main()
