"""The matrix-core probe on the GPU (ckks_probe_dft64): the radix-2 VALU
kernel and the int8 MFMA kernel give the exact 64-point transform's words,
y = G^reps x mod q, for every vector of the batch.  Of the two candidate
accumulator layouts only the hardware's (row 4 g + e of the tile) matches;
the software-pipelined form of that layout matches too."""
import numpy as np
import pytest

import ckks_oracle as O
from cipherpu import dft64_probe as P
from cipherpu.params import CHAIN4, SPECIAL_P

pytestmark = pytest.mark.gpu


def _run(T, x_h, method, reps):
    import torch
    x = torch.from_numpy(x_h.view(np.int64)).cuda()
    y = torch.zeros_like(x)
    tab = torch.from_numpy(T.blob.copy()).cuda()
    P.launch(x, y, tab, T, method, reps)
    torch.cuda.synchronize()
    return y.cpu().numpy().view(np.uint64)


@pytest.mark.parametrize("q", [CHAIN4[0], SPECIAL_P])
@pytest.mark.parametrize("reps", [1, 2, 5])
def test_probe_kernels_exact(q, reps):
    T = P.build_tables(q, O.root_of_unity(q, 128))
    rng = np.random.default_rng(100 + reps)
    V = 1024
    x = rng.integers(0, q, (64, V), dtype=np.uint64)
    x[:, 0], x[:, 1], x[:, 2] = q - 1, 0, 1
    print(f"q={q} reps={reps} V={V}", flush=True)
    r2 = _run(T, x, 0, reps)
    cols = list(range(8)) + list(rng.choice(V, 24, replace=False))
    want = P.matmul_mod(T.G, x[:, cols], q, reps)
    assert np.array_equal(r2[:, cols], want)
    mf = [_run(T, x, m, reps) for m in (1, 2, 3)]
    ok = [bool(np.array_equal(m, r2)) for m in mf]
    print("layout 0, layout 1, layout 0 interleaved equal to radix-2:", ok, flush=True)
    assert ok == [True, False, True]


def test_probe_rejects_bad_arguments():
    import torch
    from cipherpu._lib import CkksError
    q = CHAIN4[0]
    T = P.build_tables(q, O.root_of_unity(q, 128))
    x = torch.zeros((64, 96), dtype=torch.int64, device="cuda")
    tab = torch.from_numpy(T.blob.copy()).cuda()
    with pytest.raises(CkksError):
        P.launch(x, x.clone(), tab, T, 1, 1)          # 96 vectors: not a multiple of 64
    x = torch.zeros((64, 64), dtype=torch.int64, device="cuda")
    with pytest.raises(CkksError):
        P.launch(x, x.clone(), tab, T, 4, 1)
    with pytest.raises(CkksError):
        P.launch(x, x.clone(), tab, T, 1, 65)
