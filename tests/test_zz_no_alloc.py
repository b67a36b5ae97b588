"""ckks_hmult_noalloc (the HMult entry point that allocates nothing, its
workspace a library-owned device array) against ckks_hmult with a caller
workspace, word for word, through the C ABI: two calls in a row, the second
with a smaller batch and level, and the rejection of a batch beyond the
array."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def test_hmult_noalloc_matches_hmult():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    from cipherpu import _lib
    from cipherpu.ckks import CKKS, CKKSParams
    L = _lib.load()
    ck = CKKS(CKKSParams(), seed=23)
    sk = ck.keygen()
    rlk = ck.relin_key(sk)
    rng = np.random.default_rng(23)
    h, stream = ck.ring.handle, ck.ring.stream()
    for batch, l in ((20, 4), (3, 3)):
        z1 = rng.uniform(-1, 1, (batch, 4096))
        z2 = rng.uniform(-1, 1, (batch, 4096))
        a = ck.encrypt(ck.encode(z1, level=l), sk).data
        b = ck.encrypt(ck.encode(z2, level=l), sk).data
        want = torch.empty(batch, 2, l - 1, ck.N, dtype=torch.uint64, device=ck.device)
        ws = torch.empty(L.ckks_workspace_bytes(h, batch, l) // 8, dtype=torch.uint64, device=ck.device)
        assert L.ckks_hmult(h, _p(a), _p(b), _p(rlk.prepared), ck.L + 1, _p(want), batch, l, ck.sp, _p(ws),
                            stream) == 0
        got = torch.full_like(want, 7)
        assert L.ckks_hmult_noalloc(h, _p(a), _p(b), _p(rlk.prepared), ck.L + 1, _p(got), batch, l, ck.sp,
                                    stream) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, want), (batch, l)
    # 25 pairs at N = 8192, l = 4 need 66 MiB: beyond the 64 MiB array
    out = torch.empty(1, dtype=torch.uint64, device=ck.device)
    assert L.ckks_hmult_noalloc(h, _p(a), _p(b), _p(rlk.prepared), ck.L + 1, _p(out), 25, 4, ck.sp,
                                stream) == -4
    assert L.ckks_hmult_noalloc(h, _p(a), _p(b), _p(rlk.prepared), ck.L + 1, _p(out), 0, 4, ck.sp, stream) == 0
