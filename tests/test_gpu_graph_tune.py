"""HIP-graph replay and the autotuner.

A replayed graph runs the same kernels on the same data as the eager step,
so every replay must equal the eager result bit for bit; replays are checked
one by one with other eager GPU work and host synchronisations in between.
The autotuner only changes schedules (near-field overlap, back-projection
splits, shell-sum shape), which leave results equal up to the split
boundaries of the back-projection (FP64 1e-12, mixed 1e-6).  The
single-system autotuner tunes CT = double handles and refuses CT = float
ones.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _dev(h, torch):
    return [torch.as_tensor(a, dtype=torch.float64, device="cuda") for a in h]


def _step(f, d):
    f.zero_ef()
    f.calc_ener_force(*d)
    return [t.cpu().numpy() for t in f.get_ef()]


@pytest.fixture(scope="module")
def water3k():
    import glst_amd as g
    return g.make_water_box(3000, 44.7, seed=2)


@pytest.mark.parametrize("ct", ["double", "float"])
def test_graph_replays_equal_eager(cuda, water3k, ct):
    import torch
    import glst_amd as g
    d = _dev(water3k, torch)
    eager = g.GLSTForce(d[0].numel(), 1e-4, 44.7, rcut=9.0, ct=ct)
    ref = _step(eager, d)
    f = g.GLSTForce(d[0].numel(), 1e-4, 44.7, rcut=9.0, ct=ct)
    f.set_option("graph", 1)
    assert f.get_option("graph") == 1
    for rep in range(4):
        got = _step(f, d)   # capture on the first, replays after
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
        # unrelated eager GPU work and a host synchronisation in between
        _step(eager, d)
        torch.cuda.synchronize()
    # new input arrays: the step is captured again
    d2 = [x.clone() for x in d]
    for a, b in zip(_step(f, d2), ref):
        np.testing.assert_array_equal(a, b)
    # each option change drops the graph; the overlap-free step is equal too
    f.set_option("overlap_near", 0)
    for a, b in zip(_step(f, d2), ref):
        np.testing.assert_array_equal(a, b)
    f.close()
    eager.close()


@pytest.mark.parametrize("mode", [1, 2])
def test_batch_graph_replays_equal_eager(cuda, mode):
    import torch
    import glst_amd as g
    nws = [300, 1500, 800]
    hs = [g.make_water_box(nw, (nw / 0.0334) ** (1 / 3), seed=7 + k)
          for k, nw in enumerate(nws)]
    L = [(nw / 0.0334) ** (1 / 3) for nw in nws]
    d = [torch.as_tensor(np.concatenate([h[i] for h in hs]), device="cuda")
         for i in range(4)]
    natoms = [h[0].size for h in hs]
    eager = g.GLSTBatch(natoms, L, 9.0, 1e-4, ct="float")
    ref = _step(eager, d)
    b = g.GLSTBatch(natoms, L, 9.0, 1e-4, ct="float")
    b.set_option("graph", mode)
    for rep in range(4):
        got = _step(b, d)
        for x, y in zip(got, ref):
            np.testing.assert_array_equal(x, y)
        _step(eager, d)
        torch.cuda.synchronize()
    b.close()
    eager.close()


_MIXED_TUNE_SKIP = (
    "the autotuner refuses CT = float handles: one run of this case ended in "
    "a host segmentation fault inside glst_autotune (after two runs that "
    "passed) whose cause is not found")


@pytest.mark.parametrize("ct,bar", [
    ("double", 1e-12),
    pytest.param("float", 1e-6, marks=pytest.mark.skip(reason=_MIXED_TUNE_SKIP)),
])
def test_autotune_keeps_results(cuda, water3k, ct, bar):
    import torch
    import glst_amd as g
    d = _dev(water3k, torch)
    base = g.GLSTForce(d[0].numel(), 1e-4, 44.7, rcut=9.0, ct=ct)
    ref = _step(base, d)
    f = g.GLSTForce(d[0].numel(), 1e-4, 44.7, rcut=9.0, ct=ct)
    ms = f.autotune(*d, reps=2)
    assert ms > 0
    assert f.get_option("overlap_near") in (0, 1, 2)
    assert f.get_option("lr_blocks") in (1024, 2048, 4096, 8192, 16384, 32768,
                                        65536)
    assert f.get_option("rmt_shape") in (0, 1, 3, 4)
    # accumulators were cleared; results unchanged up to split boundaries
    got = [t.cpu().numpy() for t in f.get_ef()]
    assert all(np.all(a == 0) for a in got)
    got = _step(f, d)
    assert rel_err(got[:3], ref[:3]) < bar
    assert rel_err(got[3:], ref[3:]) < bar
    base.close()
    f.close()


def test_autotune_refuses_mixed_handles(cuda, water3k):
    """A CT = float handle is refused before any step runs; its schedule is
    left as it was."""
    import torch
    import glst_amd as g
    d = _dev(water3k, torch)
    f = g.GLSTForce(d[0].numel(), 1e-4, 44.7, rcut=9.0, ct="float")
    f.set_option("overlap_near", 2)
    with pytest.raises(RuntimeError, match="not tuned"):
        f.autotune(*d, reps=1)
    assert f.get_option("overlap_near") == 2
    f.close()


def test_batch_autotune_keeps_results(cuda):
    import torch
    import glst_amd as g
    nws = [500, 2000, 1200, 150]
    hs = [g.make_water_box(nw, (nw / 0.0334) ** (1 / 3), seed=30 + k)
          for k, nw in enumerate(nws)]
    L = [(nw / 0.0334) ** (1 / 3) for nw in nws]
    d = [torch.as_tensor(np.concatenate([h[i] for h in hs]), device="cuda")
         for i in range(4)]
    natoms = [h[0].size for h in hs]
    base = g.GLSTBatch(natoms, L, 9.0, 1e-4, ct="double")
    ref = _step(base, d)
    b = g.GLSTBatch(natoms, L, 9.0, 1e-4, ct="double")
    assert b.autotune(*d, reps=2) > 0
    got = _step(b, d)
    assert rel_err(got[:3], ref[:3]) < 1e-12
    base.close()
    b.close()
