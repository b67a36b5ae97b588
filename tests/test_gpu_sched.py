"""The batch scheduler: mixed ring degrees and levels captured into one HIP
graph; every replay must equal the eager launches bit for bit, and the
eager launches must equal the C oracle and decrypt to the right values."""
import numpy as np
import pytest

import oracle_c as OC

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sched():
    import torch
    assert torch.cuda.is_available()
    from cipherpu.sched import BatchScheduler, make_workload
    ops = make_workload(96, seed=3)
    s = BatchScheduler(ops, seed=5)
    s.fill_random(seed=6)
    return s


def test_workload_mix(sched):
    from cipherpu.sched import make_workload
    ops = make_workload(2048, seed=0)
    kinds = [o.kind for o in ops]
    assert abs(kinds.count("hmult") / 2048 - 0.6) < 0.05
    assert abs(kinds.count("rotate") / 2048 - 0.3) < 0.05
    assert sum(o.N == 4096 for o in ops) == 1024
    assert {o.level for o in ops} == {1, 2, 3, 4}
    assert all(o.step & (o.step - 1) == 0 and 0 < o.step < o.N // 2 for o in ops if o.kind == "rotate")
    assert sum(bk.batch for bk in sched.buckets) == 96


def _oracle_args(ck, l):
    return ck.primes[:l], ck.primes[ck.L], ck.ring.roots[:l], ck.ring.roots[ck.L]


def test_eager_buckets_match_oracle_and_decrypt(sched):
    """The first, middle and last row of every bucket (every kind, ring
    degree and level) against the C oracle; every HMult output decrypts to
    the slot-wise product of its inputs."""
    import torch
    from cipherpu.ckks import Ciphertext
    from cipherpu.ring import to_numpy
    sched.run_eager()
    for bk in sched.buckets:
        ck = sched.ck[bk.N]
        l = bk.limbs
        out = to_numpy(bk.out)
        a = to_numpy(bk.a)
        pr, p, ps, pp = _oracle_args(ck, l)
        for r in sorted({0, bk.batch // 2, bk.batch - 1}):
            if bk.kind == "hmult":
                key = to_numpy(sched.rlk[bk.N].coeff)       # the oracle takes digits 0..l-1 itself
                want = OC.hmult(a[r], to_numpy(bk.b)[r], key, pr, p, ps, pp)
            elif bk.kind == "rotate":
                gk = sched.gk[(bk.N, bk.steps[r])]          # each ciphertext with its own step
                want = OC.rotate(a[r], to_numpy(gk.coeff), pr, p, ps, pp, gk.galois)
            else:
                b = to_numpy(bk.b)
                want = np.stack([np.stack([OC.polymul_ntt(a[r, c, i], b[2 * r + c, i], pr[i], ps[i])
                                           for i in range(l)]) for c in range(2)])
            assert np.array_equal(out[r], want), (bk.kind, bk.N, bk.level, r)
        if bk.kind == "hmult":
            sk = sched.sk[bk.N]
            za = to_numpy(ck.decode(ck.decrypt(Ciphertext(bk.a, ck.p.scale), sk), ck.p.scale)).real
            zb = to_numpy(ck.decode(ck.decrypt(Ciphertext(bk.b, ck.p.scale), sk), ck.p.scale)).real
            sc = ck.p.scale ** 2 / ck.primes[l - 1]
            zo = to_numpy(ck.decode(ck.decrypt(Ciphertext(bk.out, sc), sk), sc)).real
            assert np.abs(zo - za * zb).max() < 1e-5, (bk.N, bk.level)


def test_graph_replays_equal_eager(sched):
    import torch
    from cipherpu.ckks import Ciphertext
    sched.run_eager()
    torch.cuda.synchronize()
    ref = [o.clone() for o in sched.outputs()]
    sched.capture()
    ck = sched.ck[8192]
    side = ck.encrypt(ck.encode(np.zeros((2, 4096)), level=3), sched.sk[8192])
    for rep in range(3):
        for o in sched.outputs():
            o.fill_(0)           # a replay must rewrite every output
        torch.cuda.synchronize()
        sched.replay()
        torch.cuda.synchronize()
        for i, (o, r) in enumerate(zip(sched.outputs(), ref)):
            assert torch.equal(o, r), (rep, i)
        # the program's own eager work between replays
        ck.hmult(side, side, sched.rlk[8192])
        torch.cuda.synchronize()


def test_tune_keeps_results(sched):
    import torch
    sched.run_eager()
    torch.cuda.synchronize()
    ref = [o.clone() for o in sched.outputs()]
    sched.tune(reps=1)
    sched.run_eager()
    torch.cuda.synchronize()
    for o, r in zip(sched.outputs(), ref):
        assert torch.equal(o, r)
    assert all(bk.ms > 0 for bk in sched.buckets)
