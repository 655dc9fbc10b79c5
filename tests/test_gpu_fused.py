"""The mixed batched step as one persistent launch, or as one launch per
stage for all radial counts (glst_batch option "fused" 1, 2 and 3), on the
GPU.

The fused kernel runs the per-stage launches' own unit bodies, with the same
units except for S(k) (four cells of one direction chunk per workgroup, one
wave per cell, as the per-stage launch's one-cell workgroups compute them)
and the binning sort (four keys per thread).  So for radial counts up to 28
its results equal the per-stage step's bit for bit; above that the fused
step runs the any-count bodies where the per-stage step still unrolls, and
the two agree to the mixed bar (1e-6).  Every result is also checked against
the direct FP64 sum (1e-5) and, for a sample, the oracle (2e-6).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

RCUT = 9.0
TOL = 2e-5  # the served tolerance: radial counts 13-26 on these boxes


def _box(nw):
    return (nw / 0.0334) ** (1.0 / 3.0)


def _systems(nwaters, seed0):
    import glst_amd as g
    return [(_box(nw), g.make_water_box(nw, _box(nw), seed=seed0 + k))
            for k, nw in enumerate(nwaters)]


def _pack(systems, torch):
    cols = [np.concatenate([h[i] for _, h in systems]) for i in range(4)]
    return [torch.as_tensor(c, dtype=torch.float64, device="cuda")
            for c in cols]


def _batch(systems, tol):
    """A mixed batch whose per-stage S(k) sums directly and whose shell sums
    use the LDS tables, as the fused kernels do (the per-stage defaults, the
    Chebyshev form and the window sums, agree with them to FP32 rounding)."""
    import glst_amd as g
    b = g.GLSTBatch([h[0].size for _, h in systems],
                    [L for L, _ in systems], RCUT, tol, ct="float")
    b.set_option("sf_cheb", 0)
    b.set_option("rmt_wave", 0)
    return b


def _step(b, d):
    b.zero_ef()
    b.calc_ener_force(*d)
    return [t.cpu().numpy() for t in b.get_ef()]


def _direct_err(systems, d, got):
    import glst_amd as g
    worst, o = 0.0, 0
    for L, h in systems:
        n = h[0].size
        sl = [x[o:o + n] for x in d]
        ref = [t.cpu().numpy() for t in g.coulomb_direct(*sl)]
        worst = max(worst, rel_err([a[o:o + n] for a in got[:3]], ref[:3]))
        o += n
    return worst


@pytest.fixture(scope="module")
def mix(cuda):
    """Seven boxes from 150 waters (2^3 grid, no far field) to 5,200 (6^3),
    packed into one mixed batch; the per-stage step's outputs."""
    import torch
    systems = _systems([150, 300, 700, 1100, 2100, 3000, 5200], seed0=71)
    d = _pack(systems, torch)
    b = _batch(systems, TOL)
    ref = _step(b, d)
    b.close()
    return systems, d, ref


@pytest.mark.parametrize("variant", [1, 2, 3])
def test_fused_step_equals_per_stage_step(mix, variant):
    systems, d, ref = mix
    b = _batch(systems, TOL)
    b.set_option("fused", variant)
    got = _step(b, d)
    for k in range(4):
        assert np.array_equal(got[k], ref[k]), k
    # deterministic, and outputs accumulate like every other step
    again = _step(b, d)
    for k in range(4):
        assert np.array_equal(again[k], ref[k]), k
    b.calc_ener_force(*d)
    twice = [t.cpu().numpy() for t in b.get_ef()]
    assert rel_err(twice[:3], [2 * a for a in ref[:3]]) < 1e-15
    b.close()
    assert _direct_err(systems, d, got) < 1e-5


def test_fused_step_matches_oracle(mix):
    """The smallest system with a far field and the largest against the
    oracle's FP64 path."""
    from oracle import oracle as o
    systems, d, _ = mix
    b = _batch(systems, TOL)
    b.set_option("fused", 1)
    got = _step(b, d)
    b.close()
    offs = np.cumsum([0] + [h[0].size for _, h in systems])
    for s in (1, len(systems) - 1):
        L, h = systems[s]
        part = [a[offs[s]:offs[s + 1]] for a in got]
        ref = o.glst(*h, TOL, L, RCUT)
        assert rel_err(part[:3], ref[:3]) < 2e-6, s
        assert rel_err(part[3:], ref[3:]) < 2e-6, s


def test_fused_graph_replays_equal_eager(mix):
    """graph 1: the fill of the barrier words and the persistent launch are
    captured once and replayed; each replay, with eager zeroing and host
    synchronisations between, equals the eager step."""
    systems, d, ref = mix
    b = _batch(systems, TOL)
    b.set_option("fused", 1)
    b.set_option("graph", 1)
    for _ in range(4):
        got = _step(b, d)
        for k in range(4):
            assert np.array_equal(got[k], ref[k]), k
    b.close()


def test_fused_step_above_the_unrolled_counts(cuda):
    """tol 1e-6: the first alpha group has 34-35 radial nodes, which the
    fused step runs with the any-count bodies and the per-stage step
    unrolled; the second group (23-28) is unrolled in both."""
    import torch
    systems = _systems([300, 1000, 1500, 3000], seed0=91)
    d = _pack(systems, torch)
    b = _batch(systems, 1e-6)
    ref = _step(b, d)
    b.set_option("fused", 2)
    got = _step(b, d)
    b.set_option("fused", 3)
    per_stage = _step(b, d)
    b.close()
    assert rel_err(got[:3], ref[:3]) < 1e-6
    assert rel_err(got[3:], ref[3:]) < 1e-6
    assert _direct_err(systems, d, got) < 1e-5
    # one launch per stage runs the same units as the persistent launch
    for k in range(4):
        assert np.array_equal(per_stage[k], got[k]), k


def test_fused_autotune_keeps_results(mix):
    """The autotuner times every "fused" mode (and the other knobs) and
    keeps the fastest; results stay those of the per-stage step."""
    systems, d, ref = mix
    b = _batch(systems, TOL)
    ms = b.autotune(*d, reps=2)
    assert ms > 0.0
    got = _step(b, d)
    b.close()
    assert rel_err(got[:3], ref[:3]) < 1e-6
    assert rel_err(got[3:], ref[3:]) < 1e-6


def test_fused_server_stream(cuda):
    """A server whose two slots run fused steps (each on half of the device,
    so both may be resident at once): 60 requests of 14 systems, a third of
    them repeats of resident systems, admitted while earlier batches run.
    Every answer equals the per-stage server's answer to the same request
    (the batches' compositions, and so their back-projection splits, may
    differ: mixed bar 1e-6) and is within 1e-5 of the direct sum."""
    import torch
    import glst_amd as g
    rng = np.random.default_rng(17)
    nws = np.round(np.exp(rng.uniform(np.log(100), np.log(6000), 14)))
    pool = _systems([int(nw) for nw in nws], seed0=300)
    dev = [[torch.as_tensor(a, dtype=torch.float64, device="cuda") for a in h]
           for _, h in pool]
    order = rng.integers(len(pool), size=60)

    def serve(fused):
        srv = g.GLSTServer(TOL, 1 << 36, max_batch=6, ct="float")
        srv.set_option("fused", fused)
        outs = []
        for p in order:
            o = [torch.full((dev[p][0].numel(),), np.nan, dtype=torch.float64,
                            device="cuda") for _ in range(4)]
            srv.submit(pool[p][0], RCUT, *dev[p], *o)
            outs.append(o)
            srv.poll()
        while srv.pending():
            srv.step()
        st = srv.stats()
        srv.close()
        assert st["completed"] == len(order)
        return [[t.cpu().numpy() for t in o] for o in outs]

    fused = serve(1)
    plain = serve(0)
    for k, p in enumerate(order):
        assert rel_err(fused[k][:3], plain[k][:3]) < 1e-6, k
        assert rel_err(fused[k][3:], plain[k][3:]) < 1e-6, k
        ref = [t.cpu().numpy() for t in g.coulomb_direct(*dev[p])]
        assert rel_err(fused[k][:3], ref[:3]) < 1e-5, k


def test_fused_option_checks(cuda):
    """The launch modes belong to the mixed path; out-of-range values and the
    FP64 path are refused with an error, leaving the batch usable."""
    import torch
    import glst_amd as g
    systems = _systems([300, 700], seed0=5)
    d = _pack(systems, torch)
    b = g.GLSTBatch([h[0].size for _, h in systems], [L for L, _ in systems],
                    RCUT, TOL, ct="double")
    with pytest.raises(g.GLSTError):
        b.set_option("fused", 1)
    b.calc_ener_force(*d)
    b.close()
    b = _batch(systems, TOL)
    with pytest.raises(g.GLSTError):
        b.set_option("fused", 4)
    with pytest.raises(g.GLSTError):
        b.set_option("fused", -1)
    b.set_option("fused", 3)
    b.calc_ener_force(*d)
    b.close()
    srv = g.GLSTServer(TOL, 1 << 30, ct="double")
    with pytest.raises(g.GLSTError):
        srv.set_option("fused", 2)
    with pytest.raises(g.GLSTError):
        srv.set_option("graph", 3)
    srv.close()
