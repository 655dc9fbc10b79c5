"""The allocation-free direct sum (glst_coulomb_direct_noalloc) through the C
ABI: it must give exactly what glst_coulomb_direct gives (the same kernels
and split plan; bit for bit), also on a second call with a smaller system
that reuses the library's fixed partial-sum array, and the oracle's direct
sum within 1e-13."""
import ctypes

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _charges(n, seed):
    rng = np.random.default_rng(seed)
    p = rng.uniform(0.0, 25.0, (n, 3))
    q = rng.choice([-1.0, 1.0], n)
    return [p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q]


def _noalloc(h, torch):
    import glst_amd as g
    lib = g.library()
    fn = lib.glst_coulomb_direct_noalloc
    fn.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint32] + \
        [ctypes.c_void_p] * 5
    fn.restype = ctypes.c_int
    d = [torch.as_tensor(a, dtype=torch.float64, device="cuda") for a in h]
    out = [torch.full((h[0].size,), np.nan, dtype=torch.float64,
                      device="cuda") for _ in range(4)]
    stream = torch.cuda.current_stream().cuda_stream
    st = fn(*[t.data_ptr() for t in d], h[0].size,
            *[t.data_ptr() for t in out], ctypes.c_void_p(stream))
    assert st == 0, g.library().glst_global_last_error()
    torch.cuda.synchronize()
    return d, [t.cpu().numpy() for t in out]


def test_noalloc_direct_sum_equals_allocating_entry_point(cuda):
    import torch
    import glst_amd as g
    from oracle import oracle as o
    # 9,000 atoms: 36 source splits; then 700 atoms: 3 splits, in the same
    # fixed array, which still holds the first call's partial sums
    for n, seed in ((9000, 1), (700, 2)):
        h = _charges(n, seed)
        d, got = _noalloc(h, torch)
        ref = [t.cpu().numpy() for t in g.coulomb_direct(*d)]
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
        orc = o.coulomb(*h)
        assert rel_err(got[:3], orc[:3]) < 1e-13
        assert rel_err(got[3:], orc[3:]) < 1e-13
