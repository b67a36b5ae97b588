"""The reference's on-disk parameter formats through cipherpu.params, against
the committed fixture data (tests/golden/reference_fixtures.json holds the
chain JSON's entries and the twiddle ROMs' digests as data): the chain
JSON schema of rtl/m2/modulus_chain_chain4.json and the one-hex-word-per-line
$readmemh ROMs of scripts/gen_omega_tables.py:244-247.  CPU only."""
import hashlib
import json

import numpy as np
import pytest

import ckks_oracle as O


def _digest(vals):
    return hashlib.sha256(np.asarray(vals, dtype="<u8").tobytes()).hexdigest()


def test_chain_json_round_trip_and_parameters(fixtures, tmp_path):
    from cipherpu.ckks import CKKSParams
    from cipherpu.params import CHAIN4, SPECIAL_P, chain_primes_and_roots, load_chain_json, save_chain_json
    path = tmp_path / "modulus_chain_chain4.json"
    path.write_text(json.dumps(fixtures["chain4"], indent=2))
    chain = load_chain_json(str(path))
    assert [int(e["q"]) for e in chain] == list(CHAIN4)
    save_chain_json(str(tmp_path / "again.json"), chain)
    assert load_chain_json(str(tmp_path / "again.json")) == chain
    primes, roots = chain_primes_and_roots(chain, 8192)
    for q, r, e in zip(primes, roots, chain):
        # the chain's omega (order 16384) is psi for N = 8192: order exactly 2N
        assert r == int(e["omega"]) and pow(r, 8192, q) == q - 1
    _, roots4096 = chain_primes_and_roots(chain, 4096)
    assert all(r == pow(int(e["omega"]), 2, int(e["q"])) for r, e in zip(roots4096, chain))
    p = CKKSParams.from_chain_json(str(path))
    assert tuple(p.primes) == CHAIN4 and p.special == SPECIAL_P and len(p.roots) == 5
    # the derived roots are the oracle's (and the library's default) ones
    assert list(p.roots) == [O.root_of_unity(q, 16384) for q in list(CHAIN4) + [SPECIAL_P]]
    with pytest.raises(ValueError):
        chain_primes_and_roots(chain, 16384)
    bad = [dict(e) for e in fixtures["chain4"]]
    bad[2]["omega_inv"] = int(bad[2]["omega_inv"]) + 1
    (tmp_path / "bad.json").write_text(json.dumps(bad))
    with pytest.raises(ValueError):
        load_chain_json(str(tmp_path / "bad.json"))


@pytest.mark.parametrize("i", range(4))
def test_hex_rom_round_trip_matches_reference_digest(fixtures, tmp_path, i):
    """A ROM written in the reference's format from the chain's omega reads
    back to the values whose digest the reference's committed ROM has, and
    root_from_rom recovers omega."""
    from cipherpu.params import read_hex_table, root_from_rom, write_hex_table
    e = fixtures["chain4"][i]
    q, w = int(e["q"]), int(e["omega"])
    t = fixtures["tables"][f"rtl/m2/omega_table_chain4_q{i}.hex"]
    vals = O.power_table(w, t["count"], q)
    path = tmp_path / "omega.hex"
    write_hex_table(str(path), vals)
    got = read_hex_table(str(path))
    assert _digest(got) == t["sha256"]
    assert got[:len(t["head"])] == t["head"] and got[-len(t["tail"]):] == t["tail"]
    assert root_from_rom(got[:64], q) == w
    with pytest.raises(ValueError):
        root_from_rom([1, w, w * w % q + 1], q)


def test_hex_rom_reader_on_small_reference_tables(fixtures, tmp_path):
    """The small ROMs whose every value the fixture holds, written one hex
    word per line with comment and blank lines, read back exactly."""
    from cipherpu.params import read_hex_table
    small = {k: v for k, v in fixtures["tables"].items() if "values" in v}
    assert small
    for name, t in small.items():
        path = tmp_path / "rom.hex"
        path.write_text("// " + name + "\n" + "\n".join(f"{v:012x}" for v in t["values"]) + "\n\n")
        assert read_hex_table(str(path)) == t["values"], name
        assert _digest(t["values"]) == t["sha256"], name
