"""Batched multi-system engine (include/glst_batch.h) on the GPU.

Every system of a batch must give what a single-system handle gives for it
alone: the same kernels on the same data, differing only in how the
back-projection's direction ranges are split (FP64: summation order, bar
1e-12; mixed: FP32 partial-sum boundaries, bar 1e-6), and the FP64 batch
must match the oracle like the single path does (1e-10).  Systems are water
boxes at 0.0334 molecules/A^3 with rcut 9 A, from a 1x1x1 grid (no far
field) to 5x5x5.
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

NWATERS = [100, 300, 1000, 2000, 500, 1500]
RCUT = 9.0


def _box(nw):
    return (nw / 0.0334) ** (1.0 / 3.0)


def _systems(nwaters, seed0=11):
    import glst_amd as g
    out = []
    for k, nw in enumerate(nwaters):
        L = _box(nw)
        out.append((L, g.make_water_box(nw, L, seed=seed0 + k)))
    return out


def _pack(systems, torch):
    cols = [np.concatenate([h[i] for _, h in systems]) for i in range(4)]
    return [torch.as_tensor(c, dtype=torch.float64, device="cuda")
            for c in cols]


def _single(systems, tol, ct, torch):
    import glst_amd as g
    res = []
    for L, h in systems:
        d = [torch.as_tensor(a, dtype=torch.float64, device="cuda") for a in h]
        f = g.GLSTForce(h[0].size, tol, L, rcut=RCUT, ct=ct)
        f.calc_ener_force(*d)
        res.append([t.cpu().numpy() for t in f.get_ef()])
        f.close()
    return res


@pytest.fixture(scope="module")
def systems():
    return _systems(NWATERS)


@pytest.mark.parametrize("ct,bar", [("double", 1e-12), ("float", 1e-6)])
def test_batch_matches_single_handles(cuda, systems, ct, bar):
    import torch
    import glst_amd as g
    tol = 1e-4
    natoms = [h[0].size for _, h in systems]
    b = g.GLSTBatch(natoms, [L for L, _ in systems], RCUT, tol, ct=ct)
    assert b.num_systems() == len(systems)
    b.calc_ener_force(*_pack(systems, torch))
    got = [t.cpu().numpy() for t in b.get_ef()]
    ref = _single(systems, tol, ct, torch)
    for s, (L, h) in enumerate(systems):
        o = b.offset(s)
        assert o == sum(natoms[:s])
        part = [a[o:o + natoms[s]] for a in got]
        assert rel_err(part[:3], ref[s][:3]) < bar, s
        assert rel_err(part[3:], ref[s][3:]) < bar, s
        inf = b.info(s)
        assert inf.natom == natoms[s]
        assert tuple(inf.ncell) == g.grid_dims(1 if ct == "float" else 0,
                                               L, L, L, RCUT)
    # the smallest box has no far field (its grid is 1x1x1)
    assert b.info(0).has_long_range == 0
    assert b.info(3).has_long_range == 1
    b.close()


def test_batch_fp64_matches_oracle(cuda):
    import torch
    import glst_amd as g
    from oracle import oracle as o
    systems = _systems([300, 700, 1000], seed0=5)
    tol = 1e-5
    natoms = [h[0].size for _, h in systems]
    b = g.GLSTBatch(natoms, [L for L, _ in systems], RCUT, tol, ct="double")
    b.calc_ener_force(*_pack(systems, torch))
    got = [t.cpu().numpy() for t in b.get_ef()]
    for s, (L, h) in enumerate(systems):
        o0 = b.offset(s)
        part = [a[o0:o0 + natoms[s]] for a in got]
        ref = o.glst(*h, tol, L, RCUT)
        assert rel_err(part[:3], ref[:3]) < 1e-10, s
        assert rel_err(part[3:], ref[3:]) < 1e-10, s
    b.close()


@pytest.mark.parametrize("ct", ["double", "float"])
def test_batch_accumulates_overlap_modes_agree(cuda, systems, ct):
    """Outputs accumulate like the reference and zero_ef clears them; the
    near field on the side stream changes no arithmetic."""
    import torch
    import glst_amd as g
    natoms = [h[0].size for _, h in systems]
    d = _pack(systems, torch)
    b = g.GLSTBatch(natoms, [L for L, _ in systems], RCUT, 1e-4, ct=ct)
    b.calc_ener_force(*d)
    one = [t.cpu().numpy() for t in b.get_ef()]
    b.calc_ener_force(*d)
    two = [t.cpu().numpy() for t in b.get_ef()]
    for x, y in zip(one, two):
        np.testing.assert_allclose(y, 2 * x, rtol=1e-12, atol=1e-12)
    b.zero_ef()
    b.set_option("overlap_near", 0)
    b.calc_ener_force(*d)
    seq = [t.cpu().numpy() for t in b.get_ef()]
    for x, y in zip(one, seq):
        np.testing.assert_array_equal(x, y)
    with pytest.raises(g.GLSTError):
        b.set_option("overlap_near", 2)
    b.close()


def test_batch_single_system_equals_handle(cuda):
    """A batch of one is the single-system engine."""
    import torch
    import glst_amd as g
    systems = _systems([1000], seed0=3)
    L, h = systems[0]
    b = g.GLSTBatch([h[0].size], [L], RCUT, 1e-5, ct="double")
    b.calc_ener_force(*_pack(systems, torch))
    got = [t.cpu().numpy() for t in b.get_ef()]
    ref = _single(systems, 1e-5, "double", torch)[0]
    assert rel_err(got[:3], ref[:3]) < 1e-12
    b.close()


@pytest.mark.parametrize("ct,bar", [("double", 1e-12), ("float", 1e-6)])
def test_batch_shared_shapes_match_single_handles(cuda, ct, bar):
    """Systems of one grid shape share its device-resident node and task
    tables (one task range, separate back-projection splits and partial
    rows): a batch of repeated shapes, interleaved with others, still gives
    every system what its own handle gives."""
    import torch
    import glst_amd as g
    # 41-44 A boxes all give a 5x5x5 grid at rcut 9; 31 A gives 4x4x4
    spec = [(2200, 41.0), (700, 31.0), (2300, 42.5), (2100, 44.0),
            (650, 31.0)]
    systems = [(L, g.make_water_box(nw, L, seed=200 + k))
               for k, (nw, L) in enumerate(spec)]
    b = g.GLSTBatch([h[0].size for _, h in systems], [L for L, _ in systems],
                    RCUT, 1e-4, ct=ct)
    assert tuple(b.info(0).ncell) == tuple(b.info(2).ncell) \
        == tuple(b.info(3).ncell)
    b.calc_ener_force(*_pack(systems, torch))
    out = [t.cpu().numpy() for t in b.get_ef()]
    refs = _single(systems, 1e-4, ct, torch)
    for s, ((_, h), ref) in enumerate(zip(systems, refs)):
        a0 = b.offset(s)
        got = [o[a0:a0 + h[0].size] for o in out]
        assert rel_err(got[:3], ref[:3]) < bar
        assert rel_err(got[3:], ref[3:]) < bar
    b.close()


def test_batch_tight_tolerance_mixes_kernels(cuda):
    """At tol 1e-7 the larger grids' first alpha group has more radial nodes
    than the unrolled kernels cover, so one batch runs S(k) and the
    back-projection of those groups one atom per lane and the others over
    atom groups; every system still matches its single handle (mixed bar
    1e-6) and the FP64 path (2e-6)."""
    import torch
    import glst_amd as g
    systems = _systems([300, 1000, 1500], seed0=31)
    natoms = [h[0].size for _, h in systems]
    tol = 1e-7
    b = g.GLSTBatch(natoms, [L for L, _ in systems], RCUT, tol, ct="float")
    b.calc_ener_force(*_pack(systems, torch))
    got = [t.cpu().numpy() for t in b.get_ef()]
    b.close()
    single = _single(systems, tol, "float", torch)
    f64 = _single(systems, tol, "double", torch)
    for s in range(len(systems)):
        o = sum(natoms[:s])
        part = [a[o:o + natoms[s]] for a in got]
        assert rel_err(part[:3], single[s][:3]) < 1e-6, s
        assert rel_err(part[3:], single[s][3:]) < 1e-6, s
        assert rel_err(part[:3], f64[s][:3]) < 2e-6, s
        assert rel_err(part[3:], f64[s][3:]) < 2e-6, s


# ---- the batched-ensemble configuration at the served tolerance ------------
# tol 2e-5 at rcut 9 A keeps water boxes of 100-10,000 waters at <= 4e-6
# relative force error against the direct sum (tol 1e-4 gives 1-2 % there).
SERVED_TOL = 2e-5


@pytest.mark.parametrize("ct,bar", [("double", 1e-10), ("float", 2e-6)])
def test_batch_matches_oracle_at_served_tol(cuda, ct, bar):
    """A batch of four grid shapes (2^3 without a far field, 3^3, 4^3, 5^3):
    every system against the oracle's FP64 path and the direct sum."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    systems = _systems([150, 450, 1100, 2100], seed0=61)
    natoms = [h[0].size for _, h in systems]
    b = g.GLSTBatch(natoms, [L for L, _ in systems], RCUT, SERVED_TOL, ct=ct)
    d = _pack(systems, torch)
    b.calc_ener_force(*d)
    got = [t.cpu().numpy() for t in b.get_ef()]
    b.close()
    for s, (L, h) in enumerate(systems):
        o0 = sum(natoms[:s])
        part = [a[o0:o0 + natoms[s]] for a in got]
        ref = o.glst(*h, SERVED_TOL, L, RCUT)
        assert rel_err(part[:3], ref[:3]) < bar, s
        assert rel_err(part[3:], ref[3:]) < bar, s
        sl = [x[o0:o0 + natoms[s]] for x in d]
        direct = [t.cpu().numpy() for t in g.coulomb_direct(*sl)]
        assert rel_err(part[:3], direct[:3]) < 1e-5, s


@pytest.fixture(scope="module")
def ensemble512(cuda):
    """512 water boxes of round(exp(U(ln 100, ln 10000))) waters (seed 4) at
    0.0334 waters/A^3, rcut 9 A, tol 2e-5, packed into one mixed batch:
    (systems, packed device inputs, host outputs, offsets)."""
    import torch
    import glst_amd as g
    rng = np.random.default_rng(4)
    nw = np.round(np.exp(rng.uniform(np.log(100), np.log(10000), 512)))
    systems = []
    for k, w in enumerate(nw.astype(int)):
        L = _box(int(w))
        systems.append((L, g.make_water_box(int(w), L, seed=1000 + k)))
    natoms = [h[0].size for _, h in systems]
    d = _pack(systems, torch)
    b = g.GLSTBatch(natoms, [L for L, _ in systems], RCUT, SERVED_TOL,
                    ct="float")
    assert b.num_systems() == 512
    b.calc_ener_force(*d)
    got = [t.cpu().numpy() for t in b.get_ef()]
    offsets = [b.offset(s) for s in range(512)]
    b.close()
    assert offsets == list(np.cumsum([0] + natoms[:-1]))
    return systems, d, got, offsets


def _part(ens, s):
    systems, _, got, offsets = ens
    n = systems[s][1][0].size
    return [a[offsets[s]:offsets[s] + n] for a in got]


def test_ensemble512_every_system_equals_its_handle(ensemble512):
    """Each of the 512 systems as its own single-system handle (mixed bar
    1e-6: only the back-projection's split boundaries differ)."""
    import torch
    import glst_amd as g
    systems = ensemble512[0]
    for s, (L, h) in enumerate(systems):
        d = [torch.as_tensor(a, dtype=torch.float64, device="cuda") for a in h]
        f = g.GLSTForce(h[0].size, SERVED_TOL, L, rcut=RCUT, ct="float")
        f.calc_ener_force(*d)
        ref = [t.cpu().numpy() for t in f.get_ef()]
        f.close()
        part = _part(ensemble512, s)
        assert rel_err(part[:3], ref[:3]) < 1e-6, s
        assert rel_err(part[3:], ref[3:]) < 1e-6, s


def test_ensemble512_every_system_meets_the_error_bar(ensemble512):
    """Each of the 512 systems within 1e-5 of the direct FP64 sum."""
    import glst_amd as g
    systems, d, _, offsets = ensemble512
    worst = 0.0
    for s, (L, h) in enumerate(systems):
        n = h[0].size
        sl = [x[offsets[s]:offsets[s] + n] for x in d]
        ref = [t.cpu().numpy() for t in g.coulomb_direct(*sl)]
        err = rel_err(_part(ensemble512, s)[:3], ref[:3])
        worst = max(worst, err)
        assert err < 1e-5, (s, n, err)
    assert worst > 0.0


def test_ensemble512_sample_matches_oracle(ensemble512):
    """The smallest and largest systems and 6 drawn at random (of at most
    3,000 waters) against the oracle's FP64 path (mixed bar 2e-6)."""
    from oracle import oracle as o
    systems = ensemble512[0]
    sizes = [h[0].size for _, h in systems]
    rng = np.random.default_rng(9)
    pick = {int(np.argmin(sizes)), int(np.argmax(sizes))}
    modest = [s for s, n in enumerate(sizes) if n <= 9000 and s not in pick]
    pick |= set(rng.choice(modest, 6, replace=False).tolist())
    assert len(pick) == 8
    for s in sorted(pick):
        L, h = systems[s]
        ref = o.glst(*h, SERVED_TOL, L, RCUT)
        part = _part(ensemble512, s)
        assert rel_err(part[:3], ref[:3]) < 2e-6, s
        assert rel_err(part[3:], ref[3:]) < 2e-6, s


def test_batch_chebyshev_sf_matches_direct(cuda):
    """The batch's S(k) in Chebyshev form (the default) against direct sums
    (option sf_cheb 0) and the direct FP64 Coulomb sum, on boxes of 2^3 to
    6^3 cells at the served tolerance."""
    import torch
    import glst_amd as g
    systems = _systems([150, 700, 1500, 3000, 5200], seed0=41)
    natoms = [h[0].size for _, h in systems]
    d = _pack(systems, torch)
    out = []
    for cheb in (1, 0):
        b = g.GLSTBatch(natoms, [L for L, _ in systems], RCUT, SERVED_TOL,
                        ct="float")
        b.set_option("sf_cheb", cheb)
        b.calc_ener_force(*d)
        out.append([t.cpu().numpy() for t in b.get_ef()])
        b.close()
    o = 0
    for s, (L, h) in enumerate(systems):
        part = [a[o:o + natoms[s]] for a in out[0]]
        ref = [a[o:o + natoms[s]] for a in out[1]]
        assert rel_err(part[:3], ref[:3]) < 1e-6, s
        assert rel_err(part[3:], ref[3:]) < 1e-6, s
        direct = [t.cpu().numpy() for t in g.coulomb_direct(
            *[torch.as_tensor(a, dtype=torch.float64, device="cuda")
              for a in h])]
        assert rel_err(part[:3], direct[:3]) < 1e-5, s
        o += natoms[s]
