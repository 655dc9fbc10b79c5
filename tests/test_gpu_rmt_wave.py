"""Shell sums of small grids as FP32 window sums (option "rmt_shape" 5, CT =
float, every grid dimension up to 8 cells; the default picks them there).

A workgroup takes 16 nodes, up to 16 at a time, and sums the two cubes of
every cell as three passes of windows along lines of at most 8 cells, each
window the difference of two FP32 running sums.  Its remote sums R(k) meet the bar of
the other shapes against the oracle (1e-5 of the field's largest entry) and
its forces the mixed bar (2e-6 against the oracle); against the FP64
summed-volume tables they agree to FP32 rounding.  The batched engine runs
the same tiles for its small systems (batch option "rmt_wave").
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu


def _dev(arrs, torch):
    return [torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64,
                            device="cuda") for a in arrs]


def _host(ts):
    return [t.cpu().numpy() for t in ts]


def _lattice(box, spacing, seed):
    """Neutral +/-1 charges on a jittered lattice filling a rectangular box."""
    rng = np.random.default_rng(seed)
    axes = [np.arange(spacing / 2, b, spacing) for b in box]
    p = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    p = p + rng.uniform(-0.3, 0.3, p.shape)
    p = np.clip(p, 0.0, None)
    q = np.where(np.arange(len(p)) % 2 == 0, 1.0, -1.0)
    if len(q) % 2:
        q[-1] = 0.0
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q


def _fields_and_forces(h, box, rcut, tol, rmt_shape, torch, g):
    f = g.GLSTForce(h[0].size, tol, box, rcut=rcut, ct="float")
    f.set_option("rmt_shape", rmt_shape)
    d = _dev(h, torch)
    f.assign_atoms(*d)
    f.calc_sf()
    f.sum_rmt_sf()
    rre, rim = f.rmt()
    ncell = tuple(f.info().ncell)
    f.zero_ef()
    f.calc_ener_force(*d)
    ef = _host(f.get_ef())
    f.close()
    return ncell, (rre, rim), ef


# (box, rcut, tol, lattice spacing): 3^3, 4^3, 6^3, 4 x 5 x 6, 6 x 4 x 8,
# 6 x 8 x 7 and 8^3 grids (16, 16, 7, 8, 4, 4 and 4 nodes at a time)
CASES = [((26.0, 26.0, 26.0), 9.0, 1e-4, 2.6),
         ((31.04, 31.04, 31.04), 9.0, 1e-5, 2.6),
         ((50.0, 50.0, 50.0), 9.0, 1e-4, 3.0),
         ((31.0, 40.0, 50.0), 9.0, 1e-4, 2.6),
         ((52.0, 28.0, 70.0), 9.0, 1e-4, 2.6),
         ((54.0, 72.0, 63.0), 9.0, 1e-4, 3.4),
         ((71.0, 71.0, 71.0), 9.0, 1e-4, 3.5)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(
    str(int(np.ceil(b / c[1]))) for b in c[0]))
def test_window_shell_sums_match_oracle(cuda, case):
    import torch
    import glst_amd as g
    from oracle import oracle as o
    box, rcut, tol, spacing = case
    h = _lattice(box, spacing, 13)
    ncell, (rre, rim), ef = _fields_and_forces(h, box, rcut, tol, 5, torch,
                                               g)
    assert max(ncell) <= 8
    ref, (_, _, orre, orim) = o.glst(*h, tol, box, rcut, ct=1, fields=True)
    rscale = np.abs(orre).max()
    err_r = max(np.abs(rre - orre).max(), np.abs(rim - orim).max()) / rscale
    err_f = rel_err(ef[:3], ref[:3])
    err_e = rel_err(ef[3:], ref[3:])
    print(f"grid {ncell}: R {err_r:.2e} forces {err_f:.2e} energy {err_e:.2e}")
    assert err_r <= 1e-5
    assert err_f < 2e-6
    assert err_e < 2e-6
    # against the FP64 summed-volume tables of the same handle type
    _, (tre, tim), tef = _fields_and_forces(h, box, rcut, tol, 0, torch, g)
    assert max(np.abs(rre - tre).max(), np.abs(rim - tim).max()) <= 1e-5 * \
        rscale
    assert rel_err(ef[:3], tef[:3]) < 1e-6


def test_window_step_schedules_agree(cuda):
    """The pipelined step, the stage-by-stage step and graph replays give
    the same bits with the window shell sums on a 7^3 grid."""
    import torch
    import glst_amd as g
    box = (62.0, 62.0, 62.0)
    h = _lattice(box, 3.5, 5)
    d = _dev(h, torch)
    out = {}
    for pl in (0, 1):
        for ov in (0, 2):
            f = g.GLSTForce(h[0].size, 1e-4, box, rcut=9.0, ct="float")
            assert tuple(f.info().ncell) == (7, 7, 7)
            f.set_option("rmt_shape", 5)
            f.set_option("pipeline", pl)
            f.set_option("overlap_near", ov)
            f.zero_ef()
            f.calc_ener_force(*d)
            out[pl, ov] = _host(f.get_ef())
            f.close()
    for got in out.values():
        for a, b in zip(got, out[0, 0]):
            np.testing.assert_array_equal(a, b)
    f = g.GLSTForce(h[0].size, 1e-4, box, rcut=9.0, ct="float")
    f.set_option("rmt_shape", 5)
    f.set_option("graph", 1)
    for _ in range(3):
        f.zero_ef()
        f.calc_ener_force(*d)
        for a, b in zip(_host(f.get_ef()), out[1, 2]):
            np.testing.assert_array_equal(a, b)
        torch.cuda.synchronize()
    f.close()


@pytest.mark.parametrize("ct,box", [("float", (100.0, 40.0, 40.0)),
                                    ("double", (31.04, 31.04, 31.04))])
def test_window_shape_falls_back(cuda, ct, box):
    """A grid 12 cells long, and the FP64 path, keep the summed-volume
    tables under rmt_shape 5: the same bits as rmt_shape 0."""
    import torch
    import glst_amd as g
    h = _lattice(box, 3.0, 2)
    d = _dev(h, torch)
    out = []
    for rs in (0, 5):
        f = g.GLSTForce(h[0].size, 1e-4, box, rcut=9.0, ct=ct)
        f.set_option("rmt_shape", rs)
        assert f.get_option("rmt_shape") == rs
        f.zero_ef()
        f.calc_ener_force(*d)
        out.append(_host(f.get_ef()))
        f.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def test_batch_window_shell_sums(cuda):
    """The batch's small systems with the window shell sums (the default)
    against the LDS tables (option rmt_wave 0), single handles with the
    window shape, and the direct FP64 Coulomb sum; one system's 12^3 grid
    keeps the tables in both."""
    import torch
    import glst_amd as g
    rcut, tol = 9.0, 2e-5
    nws = [400, 900, 2000, 4000, 7000, 11000, 35000]
    boxes = [(nw / 0.0334) ** (1 / 3) for nw in nws]
    hs = [g.make_water_box(nw, L, seed=60 + k)
          for k, (nw, L) in enumerate(zip(nws, boxes))]
    natoms = [h[0].size for h in hs]
    d = [torch.as_tensor(np.concatenate([h[i] for h in hs]), device="cuda")
         for i in range(4)]
    out = []
    for wave in (1, 0):
        b = g.GLSTBatch(natoms, boxes, rcut, tol, ct="float")
        b.set_option("rmt_wave", wave)
        b.zero_ef()
        b.calc_ener_force(*d)
        out.append(_host(b.get_ef()))
        b.close()
    o = 0
    for s, h in enumerate(hs):
        part = [a[o:o + natoms[s]] for a in out[0]]
        tab = [a[o:o + natoms[s]] for a in out[1]]
        assert rel_err(part[:3], tab[:3]) < 1e-6, s
        assert rel_err(part[3:], tab[3:]) < 1e-6, s
        f = g.GLSTForce(natoms[s], tol, boxes[s], rcut=rcut, ct="float")
        f.set_option("rmt_shape", 5)
        f.zero_ef()
        f.calc_ener_force(*_dev(h, torch))
        one = _host(f.get_ef())
        f.close()
        assert rel_err(part[:3], one[:3]) < 1e-6, s
        direct = _host(g.coulomb_direct(*_dev(h, torch)))
        assert rel_err(part[:3], direct[:3]) < 1e-5, s
        o += natoms[s]
    with pytest.raises(Exception):
        b = g.GLSTBatch(natoms[:1], boxes[:1], rcut, tol, ct="float")
        try:
            b.set_option("rmt_wave", 2)
        finally:
            b.close()
