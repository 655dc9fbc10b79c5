"""The Chebyshev form of the mixed structure factors (option "sf_cheb", on by
default) against the FP64 oracle and the direct sum.

Per cell and direction, the atoms' Chebyshev moments times a fixed matrix of
Bessel values replace the sine and cosine of every (atom, radial node)
pair, in the alpha groups where that saves work.  The host plan is pinned on
the CPU (tests/test_host_cheb.py); here S(k) meets the mixed path's field bar
against the oracle (1e-5 of its largest entry), the forces the mixed bar
(2e-6), and a cell holding an atom from outside the grid falls back to the
direct sum.
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
    rng = np.random.default_rng(seed)
    axes = [np.arange(spacing / 2, b, spacing) for b in box]
    p = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    p = p + rng.uniform(-0.3, 0.3, p.shape)
    p = np.clip(p, 0.0, None)
    q = np.where(np.arange(len(p)) % 2 == 0, 1.0, -1.0)
    if len(q) % 2:
        q[-1] = 0.0
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q


def _run(h, box, rcut, tol, cheb, torch, g):
    f = g.GLSTForce(h[0].size, tol, box, rcut=rcut, ct="float")
    f.set_option("sf_cheb", cheb)
    d = _dev(h, torch)
    f.assign_atoms(*d)
    f.calc_sf()
    sre, sim = f.sf()
    f.zero_ef()
    f.calc_ener_force(*d)
    ef = _host(f.get_ef())
    f.close()
    return (sre, sim), ef


CASES = [((31.04,) * 3, 9.0, 1e-5, 2.6),
         ((31.0, 40.0, 50.0), 9.0, 1e-4, 2.6),
         ((52.0, 28.0, 70.0), 9.0, 1e-6, 2.6),
         ((142.0,) * 3, 10.0, 1e-4, 7.0)]


@pytest.mark.parametrize("case", CASES, ids=["4^3 1e-5", "4x5x6 1e-4",
                                             "6x4x8 1e-6", "15^3 1e-4"])
def test_chebyshev_sf_matches_oracle(cuda, case):
    import torch
    import glst_amd as g
    from oracle import oracle as o
    box, rcut, tol, spacing = case
    h = _lattice(box, spacing, 21)
    (sre, sim), ef = _run(h, box, rcut, tol, 1, torch, g)
    ref, (osre, osim, _, _) = o.glst(*h, tol, box, rcut, ct=1, fields=True)
    scale = np.abs(osre).max()
    err_s = max(np.abs(sre - osre).max(), np.abs(sim - osim).max()) / scale
    err_f = rel_err(ef[:3], ref[:3])
    (dre, dim), dref = _run(h, box, rcut, tol, 0, torch, g)
    err_d = max(np.abs(dre - osre).max(), np.abs(dim - osim).max()) / scale
    print(f"{case[0]} tol {tol}: S {err_s:.2e} (direct {err_d:.2e}), "
          f"forces {err_f:.2e}, against direct {rel_err(ef[:3], dref[:3]):.2e}")
    assert err_s <= 1e-5
    assert err_f < 2e-6
    assert rel_err(ef[3:], ref[3:]) < 2e-6
    assert rel_err(ef[:3], dref[:3]) < 1e-6


def test_cell_with_outside_atom_falls_back(cuda):
    """Atoms 3 A outside the box are binned into edge cells, more than half a
    cell from those cells' centres: their cells take the direct sum, and the
    forces stay within the mixed bar of the oracle."""
    import torch
    import glst_amd as g
    from oracle import oracle as o
    box = (52.0, 28.0, 70.0)
    rx, ry, rz, qc = _lattice(box, 2.6, 8)
    rx = np.concatenate([rx, [-3.0, 55.0]])
    ry = np.concatenate([ry, [10.0, 14.0]])
    rz = np.concatenate([rz, [20.0, 72.5]])
    qc = np.concatenate([qc, [1.0, -1.0]])
    h = (rx, ry, rz, qc)
    _, ef = _run(h, box, 9.0, 1e-4, 1, torch, g)
    _, dref = _run(h, box, 9.0, 1e-4, 0, torch, g)
    ref = o.glst(*h, 1e-4, box, 9.0, ct=1)
    assert rel_err(ef[:3], ref[:3]) < 2e-6
    assert rel_err(ef[:3], dref[:3]) < 1e-6


def test_chebyshev_288k(cuda):
    """The 288,000-charge water box: forces within 1e-6 of the direct sum
    form's and within the metric's 1e-5 of the direct FP64 sum."""
    import torch
    import glst_amd as g
    L = 142.0
    h = g.make_water_box(96000, L, seed=1)
    d = _dev(h, torch)
    out = {}
    for cheb in (0, 1):
        f = g.GLSTForce(h[0].size, 1e-4, L, rcut=10.0, ct="float")
        assert f.get_option("sf_cheb") == 1
        f.set_option("sf_cheb", cheb)
        f.zero_ef()
        f.calc_ener_force(*d)
        out[cheb] = _host(f.get_ef())
        f.close()
    ref = _host(g.coulomb_direct(*d))
    err_t = rel_err(out[1][:3], out[0][:3])
    err_d = rel_err(out[1][:3], ref[:3])
    print(f"288k: against the direct form {err_t:.2e}, direct sum {err_d:.2e}")
    assert err_t < 1e-6
    assert err_d < 1e-5


def test_sf_cheb_option_checks(cuda):
    import glst_amd as g
    f = g.GLSTForce(300, 1e-4, 31.04, rcut=9.0, ct="float")
    with pytest.raises(g.GLSTError):
        f.set_option("sf_cheb", 2)
    f.set_option("sf_cheb", 0)
    assert f.get_option("sf_cheb") == 0
    f.close()
