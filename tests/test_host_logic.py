"""Host-side logic that needs no GPU: the uneven-batch workload generator,
the learned controller's gear choice, the monitor's SNR and the trained
controller file."""
import math

import numpy as np


def test_workload_generator_mix_and_levels():
    from cipherpu.sched import make_workload
    ops = make_workload(2048, seed=0)
    kinds = [o.kind for o in ops]
    assert len(ops) == 2048
    assert abs(kinds.count("hmult") / 2048 - 0.6) < 0.05
    assert abs(kinds.count("rotate") / 2048 - 0.3) < 0.05
    assert abs(kinds.count("ptmul") / 2048 - 0.1) < 0.03
    assert sum(o.N == 4096 for o in ops) == 1024 and sum(o.N == 8192 for o in ops) == 1024
    assert {o.level for o in ops} == {1, 2, 3, 4}
    for o in ops:
        if o.kind == "rotate":
            assert 0 < o.step < o.N // 2 and o.step & (o.step - 1) == 0
        else:
            assert o.step == 0
    assert make_workload(64, seed=3) == make_workload(64, seed=3)


def test_buckets_group_by_degree_level_kind():
    """Every operation lands in exactly one bucket; a bucket holds one
    (N, level, kind) in workload order; rotations by different steps share
    a bucket and keep their own steps."""
    from cipherpu.sched import bucket_ops, make_workload
    ops = make_workload(2048, seed=0)
    groups = bucket_ops(ops)
    assert len(groups) == 24                     # 2 degrees x 4 levels x 3 kinds
    seen = sorted(i for *_, idx, _ in groups for i in idx)
    assert seen == list(range(len(ops)))
    for N, lv, kind, idx, steps in groups:
        assert idx == sorted(idx)
        assert all((ops[i].N, ops[i].level, ops[i].kind) == (N, lv, kind) for i in idx)
        assert steps == ([ops[i].step for i in idx] if kind == "rotate" else [])
    assert any(len(set(steps)) > 1 for *_, steps in groups)


def test_longest_first_placement():
    """Greedy placement: the largest items go to distinct streams first and
    no stream ends more than the largest item above the lightest."""
    from cipherpu.sched import place
    costs = [5.0, 1.0, 4.0, 3.0, 3.0, 2.0, 2.0, 9.0]
    where = place(costs, 3)
    load = [sum(c for c, w in zip(costs, where) if w == k) for k in range(3)]
    assert sorted(load) == [9.0, 10.0, 10.0]
    assert len({where[7], where[0], where[2]}) == 3
    assert max(load) - min(load) <= max(costs)
    assert place(costs, 1) == [0] * len(costs)
    rng = np.random.default_rng(1)
    c = rng.uniform(0, 1, 200).tolist()
    w = place(c, 4)
    loads = [sum(x for x, k in zip(c, w) if k == j) for j in range(4)]
    assert max(loads) - min(loads) <= max(c)


def test_choose_gear_margin_and_fallback():
    from cipherpu.gear import choose_gear
    # cheapest gear first: gear 3 clears 10 + 3 dB
    assert choose_gear([60, 20, 15, 13.5], 10.0) == 3
    assert choose_gear([60, 20, 15, 12.9], 10.0) == 2
    assert choose_gear([60, 12, 11, 10.5], 10.0) == 0
    # no gear clears target + margin: the plain monitor decides
    assert choose_gear([12, 11, 10, 9], 10.0) is None
    assert choose_gear([12, 11, 10, 9], 10.0, margin_db=0.0) == 2


def test_snr_closed_form():
    from cipherpu.gear import snr_db
    ref = np.array([[3.0, 4.0]])
    assert math.isinf(snr_db(ref, ref))
    # signal energy 25, noise energy 0.25 -> 20 dB
    assert abs(snr_db(ref, ref + np.array([[0.3, 0.4]])) - 20.0) < 1e-12


def test_trained_controller_file_loads_safely():
    from cipherpu import gear
    with np.load(gear.TRAINED_PATH, allow_pickle=False) as z:
        shapes = {k: z[k].shape for k in z.files}
    assert shapes["W0"] == (64, 16) and shapes["W1"] == (64, 64) and shapes["W2"] == (4, 64)
    assert shapes["mu"] == (16,) and shapes["y_mu"] == (4,)


def _ntt_layout(LOGN, T):
    """The LDS-resident transform's pass plan (csrc/ntt_core.h Cfg)."""
    E = (1 << LOGN) // T
    R = E.bit_length() - 1
    REST = LOGN - R
    NREST = (REST + R - 1) // R
    NP = 1 + NREST
    rsize = lambda p: R if p == 0 else REST // NREST + (1 if (p - 1) < REST % NREST else 0)

    def hi(p):
        return LOGN - 1 - sum(rsize(k) for k in range(p))
    return E, NP, rsize, hi, lambda p: hi(p) - rsize(p) + 1


def test_wave_local_exchanges_table():
    """Exhaustive enumeration of which transform exchanges stay inside each
    64-lane wave (the table the kernels' static_asserts pin)."""
    want = {(13, 512): [False, True, True], (13, 1024): [False, False, False, True], (13, 256): [False, True],
            (14, 1024): [False, False, True], (14, 512): [False, False], (12, 128): [False, False],
            (12, 256): [False, True], (12, 512): [False, True, True]}
    for (LOGN, T), table in want.items():
        E, NP, rsize, hi, lo = _ntt_layout(LOGN, T)

        def slots(P, w):
            L, H, M, G = lo(P), hi(P), 1 << rsize(P), E >> rsize(P)
            return {((f >> L) << (H + 1)) | (e << L) | (f & ((1 << L) - 1))
                    for tid in range(64 * w, 64 * w + 64) for g in range(G) for e in range(M) for f in [g * T + tid]}
        got = [all(slots(P - 1, w) == slots(P, w) for w in range(T // 64)) for P in range(1, NP)]
        assert got == table, (LOGN, T, got)


def test_reference_gear_policies_match_the_oracle():
    """The drop-in monitor/gear policy functions (host logic) against the
    oracle's restatements, which are pinned to the reference's thresholds
    and hand-traced register sequences, on random reading sequences."""
    import ckks_oracle as O
    from cipherpu import golden_api as G
    rng = np.random.default_rng(7)
    for x in list(rng.uniform(-5, 30, 200)) + [20, 15, 10, 10.0000001]:
        assert G.snr_to_gear(x) == f"Gear {O.snr_to_gear(x)}"
        assert G.snr_to_gear(x, 12.0) == f"Gear {O.snr_to_gear(x, 12.0)}"
    for r in list(rng.uniform(0, 0.7, 200)) + [0.10, 0.20, 0.35, 0.55, 0.551]:
        assert G.select_gear(r) == O.AdaptiveGearCtrl.best_gear(r)
    assert [G.select_gear(r, safety_first=False) for r in (0.05, 0.15, 0.3, 0.5, 0.9)] == [3, 3, 3, 3, 0]
    for _ in range(20):
        rs = rng.choice([0.05, 0.15, 0.3, 0.5, 0.9], size=60)
        g, o = G.AdaptiveGearCtrl(), O.AdaptiveGearCtrl()
        assert [g.step(r) for r in rs] == [o.step(r) for r in rs]
        es = rng.choice([600, 300, 100, 10], size=80)
        g, o = G.RtlGearCtrl(500, 200, 50), O.RtlGearCtrl(500, 200, 50)
        assert [g.step(e) for e in es] == [o.step(e) for e in es]
    s, n = rng.normal(size=64), rng.normal(size=64) * 1e-3
    assert G.compute_snr(s, n) == O.compute_snr(s, n)
    assert math.isinf(G.compute_snr(s, np.zeros(4)))
    assert G.estimate_r_hat(1.1, 1.0) == min(abs(1.1 - 1.0) / (1.0 + 1e-8), 1.0)
    assert G.estimate_r_hat(5.0, 0.0) == 1.0


def test_reference_block_sparsity_rules():
    """nn.apply_block_sparse_mask / block_sparse_mask (the reference's
    threshold rules) against a hand example and the oracle's loops."""
    import ckks_oracle as O
    from cipherpu import nn
    w = np.array([[0.1, -0.2, 0.05, 0.0, 3.0, -0.1], [0.01, 0.02, -0.03, 0.9, 0.5, 0.4]])
    # B = 4 over 12 weights: blocks [0.1 -0.2 0.05 0] (max 0.2), [3 -0.1 0.01 0.02] (3),
    # [-0.03 0.9 0.5 0.4] (0.9); threshold 0.5 drops the first only
    m, skip, nb, kept = nn.apply_block_sparse_mask(w, 0.5, 4)
    assert (skip, nb) == (1, 3) and abs(kept - 2 / 3) < 1e-15
    assert not m.ravel()[:4].any() and np.array_equal(m.ravel()[4:], w.ravel()[4:])
    # B = 5: two full blocks and a ragged tail of 2 that is never pruned
    m, skip, nb, _ = nn.apply_block_sparse_mask(w, 10.0, 5)
    assert (skip, nb) == (2, 2) and not m.ravel()[:10].any() and np.array_equal(m.ravel()[10:], w.ravel()[10:])
    m, skip, nb = nn.block_sparse_mask(w, 0.25, 4)          # threshold 0.75: blocks 0 drops, 1 and 2 stay
    assert (skip, nb) == (1, 3)
    rng = np.random.default_rng(3)
    for shape, B in (((784, 128), 64), ((128, 10), 64), ((30, 7), 16)):
        W = rng.normal(0, 1 / np.sqrt(shape[0]), shape)
        for t in (0.0, 0.05, 0.1, 0.2, 1.0):
            got, want = nn.apply_block_sparse_mask(W, t, B), O.apply_block_sparse_mask(W, t, B)
            assert np.array_equal(got[0], want[0]) and got[1:] == want[1:]
            got, want = nn.block_sparse_mask(W, t, B), O.block_sparse_mask(W, t, B)
            assert np.array_equal(got[0], want[0]) and got[1:] == want[1:]


def test_prepared_pair_forms():
    """The host's prepared pairs: FP64 form (w, RN(w / q)) as IEEE division
    gives it; integer form (w, floor(w 2^63 / q)) against Python integers,
    for 45-bit primes and the widest q < 2^63, at 0, 1 and q - 1."""
    from cipherpu.ring import prepared_pairs, shoup_companion
    rng = np.random.default_rng(5)
    q45 = np.array([35184372744193, 35184371138561, 17592186028033], dtype=np.uint64)
    w = rng.integers(0, 2 ** 45, (50, 3)).astype(np.uint64) % q45
    w[0], w[1], w[2] = 0, 1, q45 - np.uint64(1)
    f = prepared_pairs(w, q45).view(np.float64)
    assert np.array_equal(f[..., 0], w.astype(np.float64))
    assert np.array_equal(f[..., 1], w.astype(np.float64) / q45.astype(np.float64))
    i = prepared_pairs(w, q45, fp64=False)
    assert np.array_equal(i[..., 0], w)
    want = [[(int(a) << 63) // int(b) for a, b in zip(row, q45)] for row in w]
    assert np.array_equal(i[..., 1].astype(object), np.array(want, dtype=object))
    big = np.array([2 ** 63 - 25], dtype=np.uint64)
    wb = np.array([0, 1, 2 ** 62 + 7, 2 ** 63 - 26], dtype=np.uint64)
    assert [int(x) for x in shoup_companion(wb, big)] == [(int(a) << 63) // int(big[0]) for a in wb]
