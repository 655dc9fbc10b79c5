"""GPU parity of the HIP path against the CPU oracle, called through the C ABI.

Bars: bit-exact for binning and ordering; FP64 path within 1e-10 relative of
the oracle's FP64 restatement (same algorithm, different summation order);
mixed path within 2e-6 relative of the FP64 oracle; both within the
tol-derived error of the direct all-pairs Coulomb sum.
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


@pytest.fixture(scope="module")
def water1k():
    import glst_amd as g
    return g.make_water_box(1000, 31.04, seed=1)


@pytest.mark.parametrize("ct", ["double", "float"])
def test_binning_bit_exact(cuda, water1k, ct):
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rx, ry, rz, qc = water1k
    n = rx.size
    f = g.GLSTForce(n, 1e-5, 31.04, rcut=9.0, ct=ct)
    f.assign_atoms(*_dev([rx, ry, rz, qc], torch))
    ac, si, cp, cc = f.cells()
    ict = 1 if ct == "float" else 0
    ncell = tuple(f.info().ncell)
    assert ncell == o.grid_dims(ict, 31.04, 31.04, 31.04, 9.0)
    oac, osi, ocp, occ = o.bin_atoms(ict, rx, ry, rz, ncell, 9.0)
    np.testing.assert_array_equal(ac, oac)
    np.testing.assert_array_equal(si, osi)
    np.testing.assert_array_equal(cp, ocp)
    np.testing.assert_array_equal(cc, occ)
    assert f.info().max_atoms_cell == occ.max()


def test_binning_edge_cases(cuda):
    """Atoms outside the box, on cell faces, negative and huge coordinates."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rng = np.random.default_rng(7)
    L, rc = 30.0, 10.0
    base = rng.uniform(0, L, (3000, 3))
    special = np.array([[0, 0, 0], [10, 10, 10], [20, 20, 20], [30, 30, 30],
                        [-0.5, 5, 5], [-1.0, -1.0, -1.0], [35, 2, 2],
                        [1e12, 1e12, -1e12], [9.999999999, 19.99999999, 0],
                        [np.nextafter(10, 0), np.nextafter(20, 0), 1e-300]])
    pos = np.vstack([base, special, base[:50]])  # duplicates too
    q = rng.uniform(-1, 1, len(pos))
    for ct, ict in (("double", 0), ("float", 1)):
        f = g.GLSTForce(len(pos), 1e-4, L, rcut=rc, ct=ct)
        f.assign_atoms(*_dev([pos[:, 0], pos[:, 1], pos[:, 2], q], torch))
        ac, si, cp, cc = f.cells()
        oac, osi, ocp, occ = o.bin_atoms(ict, pos[:, 0], pos[:, 1], pos[:, 2],
                                         tuple(f.info().ncell), rc)
        np.testing.assert_array_equal(ac, oac)
        np.testing.assert_array_equal(si, osi)
        np.testing.assert_array_equal(cp, ocp)
        np.testing.assert_array_equal(cc, occ)


@pytest.mark.parametrize("rmt_shape", [0, 1, 2])
def test_fields_fp64_match_oracle(cuda, water1k, rmt_shape):
    """S(k) and R(k) of the FP64 path against the oracle, for every shell-sum
    kernel shape (2 = the global-memory tables of large grids)."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rx, ry, rz, qc = water1k
    tol = 1e-4
    f = g.GLSTForce(rx.size, tol, 31.04, rcut=9.0, ct="double")
    f.set_option("rmt_shape", rmt_shape)
    f.assign_atoms(*_dev([rx, ry, rz, qc], torch))
    f.calc_sf()
    sre, sim = f.sf()
    f.sum_rmt_sf()
    rre, rim = f.rmt()
    _, (osre, osim, orre, orim) = o.glst(rx, ry, rz, qc, tol, 31.04, 9.0,
                                         fields=True)
    scale = np.abs(osre).max()
    assert np.abs(sre - osre).max() <= 1e-12 * scale
    assert np.abs(sim - osim).max() <= 1e-12 * scale
    rscale = np.abs(orre).max()
    assert np.abs(rre - orre).max() <= 1e-11 * rscale
    assert np.abs(rim - orim).max() <= 1e-11 * rscale


@pytest.mark.parametrize("ct,bar", [("double", 1e-10), ("float", 2e-6)])
def test_step_matches_oracle(cuda, water1k, ct, bar):
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rx, ry, rz, qc = water1k
    tol = 1e-5
    f = g.GLSTForce(rx.size, tol, 31.04, rcut=9.0, ct=ct)
    f.calc_ener_force(*_dev([rx, ry, rz, qc], torch))
    got = _host(f.get_ef())
    ref = o.glst(rx, ry, rz, qc, tol, 31.04, 9.0)
    assert rel_err(got[:3], ref[:3]) < bar
    assert rel_err(got[3:], ref[3:]) < bar


@pytest.mark.parametrize("ct", ["double", "float"])
def test_step_vs_direct_coulomb(cuda, water1k, ct):
    """End to end against the direct FP64 all-pairs sum (utils/coulomb.cu)."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rx, ry, rz, qc = water1k
    f = g.GLSTForce(rx.size, 1e-5, 31.04, rcut=9.0, ct=ct)
    f.calc_ener_force(*_dev([rx, ry, rz, qc], torch))
    got = _host(f.get_ef())
    ref = o.coulomb(rx, ry, rz, qc)
    assert rel_err(got[:3], ref[:3]) < 1e-5
    assert abs(got[3].sum() - ref[3].sum()) < 1e-5 * abs(ref[3]).sum()


def test_gpu_coulomb_matches_oracle(cuda, water1k):
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rx, ry, rz, qc = water1k
    got = _host(g.coulomb_direct(*_dev([rx, ry, rz, qc], torch)))
    ref = o.coulomb(rx, ry, rz, qc)
    assert rel_err(got, ref) < 1e-13


@pytest.mark.parametrize("ct", ["double", "float"])
def test_accumulate_and_stages(cuda, water1k, ct):
    """Outputs accumulate (+=) like the reference; the staged calls equal
    calc_ener_force; zero_ef clears."""
    import torch
    import glst_amd as g
    rx, ry, rz, qc = water1k
    d = _dev([rx, ry, rz, qc], torch)
    f = g.GLSTForce(rx.size, 1e-4, 31.04, rcut=9.0, ct=ct)
    f.calc_ener_force(*d)
    one = _host(f.get_ef())
    f.assign_atoms(*d)
    f.calc_sf()
    f.sum_rmt_sf()
    f.calc_lr_ef()
    f.calc_sr_ef()
    f.comm_ef()
    two = _host(f.get_ef())
    for a, b in zip(one, two):
        np.testing.assert_allclose(b, 2 * a, rtol=1e-12, atol=1e-12)
    f.zero_ef()
    f.calc_ener_force(*d)
    again = _host(f.get_ef())
    for a, b in zip(one, again):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("rmt_shape", [0, 1, 2, 3, 5])
def test_fields_mixed_match_oracle(cuda, water1k, rmt_shape):
    """Mixed-precision S(k) and shell sums R(k) against the FP64 oracle's
    fields, for every shell-sum shape (bar: 1e-5 of the field's largest
    entry: FP32 phases of cell-local coordinates, FP32 S(k) slots)."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rx, ry, rz, qc = water1k
    tol = 1e-4
    f = g.GLSTForce(rx.size, tol, 31.04, rcut=9.0, ct="float")
    f.set_option("rmt_shape", rmt_shape)
    f.assign_atoms(*_dev([rx, ry, rz, qc], torch))
    f.calc_sf()
    sre, sim = f.sf()
    f.sum_rmt_sf()
    rre, rim = f.rmt()
    _, (osre, osim, orre, orim) = o.glst(rx, ry, rz, qc, tol, 31.04, 9.0,
                                         fields=True)
    scale = np.abs(osre).max()
    assert np.abs(sre - osre).max() <= 1e-5 * scale
    assert np.abs(sim - osim).max() <= 1e-5 * scale
    rscale = np.abs(orre).max()
    assert np.abs(rre - orre).max() <= 1e-5 * rscale
    assert np.abs(rim - orim).max() <= 1e-5 * rscale


def test_mixed_options_agree(cuda, water1k):
    """Every combination of the tuning options gives the same forces: the
    near field's overlap changes no arithmetic at all, the summed-volume
    shapes agree to FP64 rounding before the FP32 store, and the matrix-core
    and small-grid window shapes to FP32 rounding of the remote sums."""
    import torch
    import glst_amd as g
    rx, ry, rz, qc = water1k
    d = _dev([rx, ry, rz, qc], torch)
    out = {}
    for rs in (0, 1, 2, 3, 5):
        for ov in (0, 1, 2):
            for pl in (0, 1):
                f = g.GLSTForce(rx.size, 1e-5, 31.04, rcut=9.0, ct="float")
                f.set_option("rmt_shape", rs)
                f.set_option("overlap_near", ov)
                f.set_option("pipeline", pl)
                assert f.get_option("pipeline") == pl
                f.calc_ener_force(*d)
                out[rs, ov, pl] = _host(f.get_ef())
                f.close()
    for rs in (0, 1, 2, 3, 5):
        for ov in (0, 1, 2):
            for pl in (0, 1):
                for a, b in zip(out[rs, 0, 0], out[rs, ov, pl]):
                    np.testing.assert_array_equal(a, b)
        # the matrix-core and window shell sums round in FP32 (the tables
        # in FP64)
        assert rel_err(out[rs, 0, 0][:3], out[0, 0, 0][:3]) < (
            1e-6 if rs in (3, 5) else 1e-7)
    f = g.GLSTForce(rx.size, 1e-5, 31.04, rcut=9.0, ct="float")
    with pytest.raises(Exception):
        f.set_option("rmt_shape", 6)
    with pytest.raises(Exception):
        f.set_option("overlap_near", 3)
    with pytest.raises(Exception):
        f.set_option("pipeline", 2)
    with pytest.raises(Exception):
        f.set_option("no_such_option", 1)


@pytest.mark.parametrize("nw,L,rcut,tol", [(1000, 31.04, 9.0, 1e-5),
                                            (6000, 57.0, 9.0, 1e-4),
                                            (1000, 31.04, 9.0, 1e-7)])
def test_atom_group_back_projection(cuda, nw, L, rcut, tol):
    """The mixed back-projection over atom groups (up to four same-cell atoms
    per group, eight lanes sharing each direction's remote sums) agrees with
    the one-atom-per-lane kernel and with the FP64 path, for split counts
    from one split per alpha group to one pass of eight directions per
    split, with cells whose atom counts are not multiples of four.  At
    tol 1e-7 the first alpha group has 42 radial nodes, past the unrolled
    kernels: it runs one atom per lane while the second (29) uses groups."""
    import torch
    import glst_amd as g
    h = g.make_water_box(nw, L, seed=11)
    d = _dev(h, torch)
    f = g.GLSTForce(h[0].size, tol, L, rcut=rcut, ct="double")
    f.calc_ener_force(*d)
    f64 = _host(f.get_ef())
    f.close()
    out = {}
    for blocks in (1, 4096, 1 << 22):
        for agm in (0, 1):
            f = g.GLSTForce(h[0].size, tol, L, rcut=rcut, ct="float")
            f.set_option("lr_blocks", blocks)
            f.set_option("lr_atom_groups", agm)
            assert f.get_option("lr_atom_groups") == agm
            f.calc_ener_force(*d)
            out[blocks, agm] = _host(f.get_ef())
            f.close()
    for key, r in out.items():
        assert rel_err(r[:3], f64[:3]) < 2e-6, key
        assert rel_err(r[3:], f64[3:]) < 2e-6, key
        lane = out[key[0], 0]
        assert rel_err(r[:3], lane[:3]) < 2e-7, key
        assert rel_err(r[3:], lane[3:]) < 2e-7, key
    f = g.GLSTForce(h[0].size, tol, L, rcut=rcut, ct="float")
    with pytest.raises(Exception):
        f.set_option("lr_atom_groups", 2)
    f.close()


@pytest.mark.parametrize("nw,L,rcut", [(4000, 49.0, 9.0), (12000, 71.0, 9.0),
                                       (20000, 84.0, 10.0)])
def test_pipelined_far_field_bitwise(cuda, nw, L, rcut):
    """The far field pipelined across alpha groups (shell sums of group g on a
    third stream, started per group; workgroup tiles that straddle two groups
    wait for both) equals the serial schedule bit for bit, eagerly over
    repeated steps and as a replayed graph, on grids with several groups."""
    import torch
    import glst_amd as g
    h = g.make_water_box(nw, L, seed=7)
    d = _dev(h, torch)
    res = {}
    for pl in (0, 1):
        for graph in (0, 1):
            f = g.GLSTForce(h[0].size, 1e-4, L, rcut=rcut, ct="float")
            f.set_option("pipeline", pl)
            f.set_option("graph", graph)
            if pl == 1 and graph == 0:
                info = f.info()
                assert info.ngroup >= 2
            steps = []
            for _ in range(3):
                f.zero_ef()
                f.calc_ener_force(*d)
                torch.zeros(1 << 20, device="cuda").sum().item()  # eager work
                steps.append(_host(f.get_ef()))
            f.close()
            for st in steps[1:]:
                for a, b in zip(steps[0], st):
                    np.testing.assert_array_equal(a, b)
            res[pl, graph] = steps[0]
    for key in res:
        for a, b in zip(res[0, 0], res[key]):
            np.testing.assert_array_equal(a, b)


def _lattice_charges(box, spacing, seed):
    """Neutral +/-1 charges on a jittered lattice filling a rectangular box
    (no near-coincident pairs)."""
    rng = np.random.default_rng(seed)
    axes = [np.arange(spacing / 2, b, spacing) for b in box]
    p = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    p = p + rng.uniform(-0.3, 0.3, p.shape)
    p = np.clip(p, 0.0, None)
    q = np.where(np.arange(len(p)) % 2 == 0, 1.0, -1.0)
    if len(q) % 2:
        q[-1] = 0.0
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q


@pytest.mark.parametrize("box", [(31.0, 40.0, 50.0), (52.0, 28.0, 70.0)])
@pytest.mark.parametrize("ct", ["double", "float"])
def test_rectangular_grids_match_oracle(cuda, box, ct):
    """Grids with three different dimensions (odd and even z, the shell-sum
    tables' padding and border differ per axis): S(k), R(k) for every
    shell-sum shape and the forces against the FP64 oracle."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rx, ry, rz, qc = _lattice_charges(box, 2.6, 3)
    tol, rcut = 1e-4, 9.0
    ict = 1 if ct == "float" else 0
    (ref, (osre, osim, orre, orim)) = o.glst(rx, ry, rz, qc, tol, box, rcut,
                                           ct=ict, fields=True)
    sbar, rbar, fbar = ((1e-12, 1e-11, 1e-10) if ct == "double"
                        else (1e-5, 1e-5, 2e-6))
    d = _dev([rx, ry, rz, qc], torch)
    for rs in (0, 1, 2, 3, 5):
        f = g.GLSTForce(rx.size, tol, box, rcut=rcut, ct=ct)
        f.set_option("rmt_shape", rs)
        assert tuple(f.info().ncell) == o.grid_dims(ict, *box, rcut)
        f.assign_atoms(*d)
        f.calc_sf()
        sre, sim = f.sf()
        f.sum_rmt_sf()
        rre, rim = f.rmt()
        scale = np.abs(osre).max()
        assert np.abs(sre - osre).max() <= sbar * scale
        assert np.abs(sim - osim).max() <= sbar * scale
        rscale = np.abs(orre).max()
        assert np.abs(rre - orre).max() <= rbar * rscale
        assert np.abs(rim - orim).max() <= rbar * rscale
        f.close()
    f = g.GLSTForce(rx.size, tol, box, rcut=rcut, ct=ct)
    f.calc_ener_force(*d)
    got = _host(f.get_ef())
    f.close()
    assert rel_err(got[:3], ref[:3]) < fbar
    assert rel_err(got[3:], ref[3:]) < fbar
    # batched with a second rectangular system: each equals its own handle
    # (FP64 1e-12, mixed 1e-6: only the back-projection splits differ)
    h2 = _lattice_charges(box[::-1], 2.6, 4)
    systems = [((rx, ry, rz, qc), box), (h2, box[::-1])]
    b = g.GLSTBatch([h[0].size for h, _ in systems],
                    [bx for _, bx in systems], rcut, tol, ct=ct)
    cols = _dev([np.concatenate([h[i] for h, _ in systems]) for i in range(4)],
                torch)
    b.calc_ener_force(*cols)
    out = _host(b.get_ef())
    bbar = 1e-12 if ct == "double" else 1e-6
    for s, (h, bx) in enumerate(systems):
        f = g.GLSTForce(h[0].size, tol, bx, rcut=rcut, ct=ct)
        f.calc_ener_force(*_dev(h, torch))
        one = _host(f.get_ef())
        f.close()
        a0 = b.offset(s)
        mine = [o[a0:a0 + h[0].size] for o in out]
        assert rel_err(mine[:3], one[:3]) < bbar
        assert rel_err(mine[3:], one[3:]) < bbar
    b.close()


@pytest.mark.parametrize("box", [(31.0, 40.0, 8.5), (40.0, 31.0, 15.0),
                                 (16.0, 28.0, 40.0)])
@pytest.mark.parametrize("ct", ["double", "float"])
def test_thin_grids_near_field_match_oracle(cuda, box, ct):
    """Grids one or two cells deep along some axis (4x5x1, 5x4x2, 2x4x5 at
    rcut 9): the near field sweeps the z-neighbour cells of each (x, y)
    column as one atom range, clamped at the grid edge; the near field alone
    and the whole step against the FP64 oracle."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rx, ry, rz, qc = _lattice_charges(box, 2.6, 5)
    tol, rcut = 1e-4, 9.0
    ict = 1 if ct == "float" else 0
    bar = 1e-10 if ct == "double" else 2e-6
    d = _dev([rx, ry, rz, qc], torch)
    f = g.GLSTForce(rx.size, tol, box, rcut=rcut, ct=ct)
    assert tuple(f.info().ncell) == o.grid_dims(ict, *box, rcut)
    assert min(f.info().ncell) <= 2
    f.assign_atoms(*d)
    f.calc_sr_ef()
    f.comm_ef()
    near = _host(f.get_ef())
    f.close()
    ref_near = o.glst(rx, ry, rz, qc, tol, box, rcut, ct=ict, far=False)
    assert rel_err(near[:3], ref_near[:3]) < bar
    assert rel_err(near[3:], ref_near[3:]) < bar
    f = g.GLSTForce(rx.size, tol, box, rcut=rcut, ct=ct)
    f.calc_ener_force(*d)
    got = _host(f.get_ef())
    f.close()
    ref = o.glst(rx, ry, rz, qc, tol, box, rcut, ct=ict)
    assert rel_err(got[:3], ref[:3]) < bar
    assert rel_err(got[3:], ref[3:]) < bar


@pytest.mark.parametrize("box,ncell", [((31.04, 31.04, 31.04), (4, 4, 4)),
                                       ((31.04, 35.0, 40.0), (4, 4, 5))])
@pytest.mark.parametrize("ct", ["double", "float"])
def test_ncell_constructor_matches_oracle(cuda, box, ncell, ct):
    """glst_force(natom, tol, Lx, Ly, Lz, nx, ny, nz) (src/glst_force.cu:
    80-95): rcut = min_d L_d / n_d, then the grid is derived from that rcut
    in CT arithmetic, so a non-cubic box need not get the counts asked for
    (31.04 x 35 x 40 with 4 x 4 x 5 gives rcut 7.76 and 4 x 5 x 6 cells).
    The step must equal the oracle run at that rcut."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rx, ry, rz, qc = _lattice_charges(box, 2.6, 5)
    tol = 1e-5
    ict = 1 if ct == "float" else 0
    rcut = min(L / n for L, n in zip(box, ncell))
    f = g.GLSTForce(rx.size, tol, box, ncell=ncell, ct=ct)
    grid = tuple(f.info().ncell)
    assert grid == o.grid_dims(ict, *box, rcut)
    if box[0] == box[1] == box[2]:
        assert grid == ncell
    else:
        assert grid == (4, 5, 6)
    f.calc_ener_force(*_dev([rx, ry, rz, qc], torch))
    got = _host(f.get_ef())
    f.close()
    ref = o.glst(rx, ry, rz, qc, tol, box, rcut)
    bar = 1e-10 if ct == "double" else 2e-6
    assert rel_err(got[:3], ref[:3]) < bar
    assert rel_err(got[3:], ref[3:]) < bar


@pytest.mark.parametrize("n", [2, 255, 257, 700, 9000])
def test_gpu_coulomb_split_sizes_match_oracle(cuda, n):
    """The direct sum with its source atoms split into ranges (up to 64 for
    small systems, partial sums added in range order) against the oracle:
    one tile, a tile and one atom, and sizes with ragged last ranges."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rng = np.random.default_rng(n)
    p = rng.uniform(0.0, 20.0, (n, 3))
    q = rng.choice([-1.0, 1.0], n)
    h = [p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q]
    got = _host(g.coulomb_direct(*_dev(h, torch)))
    ref = o.coulomb(*h)
    assert rel_err(got[:3], ref[:3]) < 1e-13
    assert rel_err(got[3:], ref[3:]) < 1e-13


@pytest.mark.parametrize("ct", ["double", "float"])
def test_atoms_far_outside_the_box(cuda, water1k, ct):
    """Atoms hundreds of Angstrom beyond the box are binned into the edge
    cells (src/glst_force.cu:372-377), so their offsets from the cell centre
    are that large.  At 1,000 A the mixed path's FP32 phases k.d reach about
    280 revolutions, past the input range of the hardware sine and cosine
    (|t| <= 256), and must be reduced first.  Each far atom's own force and
    energy, and all forces together, must match the oracle (its FP64
    restatement uses absolute coordinates, where nothing leaves range).  The
    FP32 offset of an atom 1,000 A out carries 6e-5 A of rounding, about 1e-4
    rad of phase, so the mixed path's bars are 1e-5 over all forces (the
    metric's) and 1e-3 per far atom, far below what unreduced phases give."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    rx, ry, rz, qc = [a.copy() for a in water1k]
    far = [5, 1001, 2002]
    rx[far[0]] += 1000.0
    ry[far[1]] += 700.0
    rz[far[2]] += 450.0
    rx[far[2]] += 450.0
    f = g.GLSTForce(rx.size, 1e-5, 31.04, rcut=9.0, ct=ct)
    f.calc_ener_force(*_dev([rx, ry, rz, qc], torch))
    got = _host(f.get_ef())
    f.close()
    ref = o.glst(rx, ry, rz, qc, 1e-5, 31.04, 9.0)
    bar, atom_bar = (1e-10, 1e-9) if ct == "double" else (1e-5, 1e-3)
    assert rel_err(got[:3], ref[:3]) < bar
    assert rel_err(got[3:], ref[3:]) < bar
    frms = np.sqrt(np.mean(sum(r ** 2 for r in ref[:3])))
    erms = np.sqrt(np.mean(ref[3] ** 2))
    for i in far:
        fi = np.sqrt(sum(r[i] ** 2 for r in ref[:3]))
        df = np.sqrt(sum((a[i] - b[i]) ** 2 for a, b in zip(got[:3], ref[:3])))
        assert df < atom_bar * max(fi, frms), (i, df / max(fi, frms))
        assert abs(got[3][i] - ref[3][i]) < atom_bar * max(abs(ref[3][i]),
                                                           erms), i
