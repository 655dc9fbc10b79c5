"""Shell sums on the matrix cores (option "rmt_shape" 3, CT = float) against
the FP64 oracle and the summed-volume tables.

The matrix-core kernel sums each cube as z windows of FP32 running sums and
x and y windows as products with 0/1 band matrices (FP32 sums), where the
summed-volume tables (shapes 0-2) difference FP64 running sums.  Its remote
sums R(k) meet the same bar against the oracle as the other shapes (1e-5 of
the field's largest entry) and its forces the mixed bar (2e-6 against the
oracle).  Grids wider than 16 cells in x or y, and the FP64 path, fall back
to the tables, bit for bit.
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


# (box, rcut, tol, lattice spacing): 4^3, 4 x 5 x 6, 6 x 4 x 8, 15^3 and
# 16 x 16 x 30 grids
CASES = [((31.04, 31.04, 31.04), 9.0, 1e-5, 2.6),
         ((31.0, 40.0, 50.0), 9.0, 1e-4, 2.6),
         ((52.0, 28.0, 70.0), 9.0, 1e-4, 2.6),
         ((142.0, 142.0, 142.0), 10.0, 1e-4, 7.0),
         ((152.0, 152.0, 285.0), 9.6, 1e-4, 9.0)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(
    str(int(np.ceil(b / c[1]))) for b in c[0]))
def test_matrix_core_shell_sums_match_oracle(cuda, case):
    import torch
    import glst_amd as g
    from oracle import oracle as o
    box, rcut, tol, spacing = case
    h = _lattice(box, spacing, 11)
    ncell, (rre, rim), ef = _fields_and_forces(h, box, rcut, tol, 3, torch,
                                               g)
    assert max(ncell[:2]) <= 16
    ref, (_, _, orre, orim) = o.glst(*h, tol, box, rcut, ct=1, fields=True)
    rscale = np.abs(orre).max()
    err_r = max(np.abs(rre - orre).max(), np.abs(rim - orim).max()) / rscale
    err_f = rel_err(ef[:3], ref[:3])
    err_e = rel_err(ef[3:], ref[3:])
    print(f"grid {ncell}: R {err_r:.2e} forces {err_f:.2e} energy {err_e:.2e}")
    assert err_r <= 1e-5
    assert err_f < 2e-6
    assert err_e < 2e-6
    # against the summed-volume tables of the same handle type
    _, (tre, tim), tef = _fields_and_forces(h, box, rcut, tol, 0, torch, g)
    assert max(np.abs(rre - tre).max(), np.abs(rim - tim).max()) <= 1e-5 * \
        rscale
    assert rel_err(ef[:3], tef[:3]) < 1e-6


def test_matrix_core_step_schedules_agree(cuda):
    """The pipelined step, the stage-by-stage step and graph replays give
    the same bits with the matrix-core shell sums (each node's sums do not
    depend on how the node range is cut into launches)."""
    import torch
    import glst_amd as g
    box = (142.0, 142.0, 142.0)
    h = _lattice(box, 7.0, 5)
    d = _dev(h, torch)
    out = {}
    for pl in (0, 1):
        for ov in (0, 2):
            f = g.GLSTForce(h[0].size, 1e-4, box, rcut=10.0, ct="float")
            f.set_option("rmt_shape", 3)
            f.set_option("pipeline", pl)
            f.set_option("overlap_near", ov)
            f.zero_ef()
            f.calc_ener_force(*d)
            out[pl, ov] = _host(f.get_ef())
            f.close()
    for key, got in out.items():
        for a, b in zip(got, out[0, 0]):
            np.testing.assert_array_equal(a, b)
    f = g.GLSTForce(h[0].size, 1e-4, box, rcut=10.0, ct="float")
    f.set_option("rmt_shape", 3)
    f.set_option("graph", 1)
    for _ in range(3):
        f.zero_ef()
        f.calc_ener_force(*d)
        for a, b in zip(_host(f.get_ef()), out[1, 2]):
            np.testing.assert_array_equal(a, b)
    f.close()


@pytest.mark.parametrize("ct,box", [("float", (180.0, 40.0, 40.0)),
                                    ("double", (31.04, 31.04, 31.04))])
def test_matrix_core_shape_falls_back(cuda, ct, box):
    """A grid 20 cells wide, and the FP64 path, keep the summed-volume
    tables under rmt_shape 3: the same bits as rmt_shape 0."""
    import torch
    import glst_amd as g
    h = _lattice(box, 3.0, 2)
    d = _dev(h, torch)
    out = []
    for rs in (0, 3):
        f = g.GLSTForce(h[0].size, 1e-4, box, rcut=9.0, ct=ct)
        f.set_option("rmt_shape", rs)
        assert f.get_option("rmt_shape") == rs
        f.zero_ef()
        f.calc_ener_force(*d)
        out.append(_host(f.get_ef()))
        f.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("box,rcut,same_as", [((142.0,) * 3, 10.0, 3),
                                               ((31.04,) * 3, 9.0, 5),
                                               ((90.0,) * 3, 9.0, 0)])
def test_default_shape(cuda, box, rcut, same_as):
    """The default (rmt_shape 4) takes the matrix cores on the 15^3 grid,
    whose z slices fill most of their 16 x 16 tile, the small-grid window
    sums on the 4^3 grid, whose slices fill a sixteenth of it, and the
    summed-volume tables on the 10^3 grid, which neither fits."""
    import torch
    import glst_amd as g
    h = _lattice(box, 7.0 if box[0] > 50 else 2.6, 4)
    d = _dev(h, torch)
    out = []
    for rs in (None, same_as):
        f = g.GLSTForce(h[0].size, 1e-4, box, rcut=rcut, ct="float")
        if rs is None:
            assert f.get_option("rmt_shape") == 4
        else:
            f.set_option("rmt_shape", rs)
        f.zero_ef()
        f.calc_ener_force(*d)
        out.append(_host(f.get_ef()))
        f.close()
    for a, b in zip(*out):
        np.testing.assert_array_equal(a, b)


def test_matrix_core_288k(cuda):
    """The 288,000-charge water box (15^3 cells): forces within 1e-6 of the
    summed-volume tables' and within the metric's 1e-5 of the direct FP64
    sum."""
    import torch
    import glst_amd as g
    L = 142.0
    h = g.make_water_box(96000, L, seed=1)
    d = _dev(h, torch)
    out = {}
    for rs in (0, 3):
        f = g.GLSTForce(h[0].size, 1e-4, L, rcut=10.0, ct="float")
        f.set_option("rmt_shape", rs)
        f.zero_ef()
        f.calc_ener_force(*d)
        out[rs] = _host(f.get_ef())
        f.close()
    ref = _host(g.coulomb_direct(*d))
    err_t = rel_err(out[3][:3], out[0][:3])
    err_d = rel_err(out[3][:3], ref[:3])
    print(f"288k: against the tables {err_t:.2e}, direct sum {err_d:.2e}")
    assert err_t < 1e-6
    assert err_d < 1e-5
