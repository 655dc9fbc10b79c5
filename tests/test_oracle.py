"""Pins the CPU oracle (oracle/glst_oracle.c) before it is trusted as the
checker: golden vectors from SciPy/NumPy, closed forms, invariants and
brute-force restatements on tiny inputs.  CPU only.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import special

from oracle import oracle as o

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def water(nw, L, seed=1):
    import glst_amd as g
    return g.make_water_box(nw, L, seed=seed)


# ---- erfc_inv (src/glst_force.cu:1616-1640) --------------------------------

def test_erfc_inv_golden():
    """The Newton iteration stops at |1 - erf(x) - y| < eps (absolute in y),
    so x is accurate to eps / |erfc'(x)|; compare within that."""
    eps = np.finfo(float).eps
    d = json.load(open(os.path.join(GOLD, "erfc_inv.json")))
    for y, x in zip(d["y"], d["x"]):
        got = o.erfc_inv(y)
        slope = 2.0 / math.sqrt(math.pi) * math.exp(-x * x)
        assert abs(got - x) <= 4 * eps / slope + 1e-13, (y, got, x)
        assert abs(special.erfc(got) - y) <= 4 * eps


def test_erfc_inv_rejects_outside_domain():
    for y in (0.0, -1.0, 2.0, 3.0, float("nan")):
        with pytest.raises(o.OracleError):
            o.erfc_inv(y)


# ---- alpha groups (src/glst_force.cu:1642-1724) -----------------------------

# Geometry of the benchmark configurations derived independently (widths,
# r_min and r_max of every group) from the group rule's text.
GROUP_TABLE = {
    (4, 9.0): [(1, 9.0, 46.77), (1, 18.0, 62.35)],
    (15, 10.0): [(1, 10.0, 51.96), (2, 20.0, 86.60), (5, 40.0, 173.21),
                 (5, 90.0, 259.81)],
    (18, 12.0): [(1, 12.0, 62.35), (2, 24.0, 103.92), (4, 48.0, 187.06),
                 (4, 96.0, 270.20), (5, 144.0, 374.12)],
}


@pytest.mark.parametrize("key", list(GROUP_TABLE))
def test_alpha_groups_table(key):
    n, cell = key
    for tol in (1e-4, 1e-5, 1e-6):
        g = o.alpha_groups(n, n, n, cell, tol)
        exp = GROUP_TABLE[key]
        assert len(g["width"]) == len(exp)
        for i, (w, rmin, rmax) in enumerate(exp):
            assert g["width"][i] == w
            assert abs(g["rmin"][i] - rmin) < 1e-9
            assert abs(g["rmax"][i] - rmax) < 5e-3
            a, z = g["alpha"][i], g["zcut"][i]
            # erfc(alpha rmin)/rmin = tol and erfc(zcut) = sqrt(pi)/2 tol/alpha
            assert abs(special.erfc(a * rmin) / rmin - tol) < 1e-9 * tol
            assert abs(special.erfc(z) - 0.5 * math.sqrt(math.pi) * tol / a) \
                < 1e-9 * tol / a
        # shells (r_in, r_out] tile Chebyshev distances 2 .. n-1
        assert g["r_in"][0] == 1
        assert all(g["r_in"][i + 1] == g["r_out"][i] for i in range(len(exp) - 1))
        assert g["r_out"][-1] == n - 1


def test_alpha_groups_small_grids():
    assert len(o.alpha_groups(1, 1, 1, 10.0, 1e-4)["width"]) == 0
    g2 = o.alpha_groups(2, 2, 2, 10.0, 1e-4)
    assert list(g2["width"]) == [1] and list(g2["r_out"]) == [2]
    g3 = o.alpha_groups(2, 5, 3, 10.0, 1e-4)  # the largest dimension decides
    assert g3["r_out"][-1] == 4


# ---- Gauss-Legendre (src/cubature.cu:161-402) --------------------------------

def test_gauss_legendre_golden():
    gold = json.load(open(os.path.join(GOLD, "gauss_legendre.json")))
    for n, v in gold.items():
        x, w = o.gauss_legendre(int(n))
        np.testing.assert_allclose(x, v["x"], rtol=0, atol=2e-15)
        np.testing.assert_allclose(w, v["w"], rtol=1e-11, atol=2e-14)


def test_gauss_legendre_all_even_orders():
    """The reference's scan + Newton finds every positive root for each order
    the selection loop can visit (2..324)."""
    for n in range(2, 325, 2):
        x, w = o.gauss_legendre(n)
        xl, wl = np.polynomial.legendre.leggauss(n)
        assert x.size == n // 2
        np.testing.assert_allclose(x, xl[xl > 0], atol=3e-15, rtol=0)
        np.testing.assert_allclose(w, wl[xl > 0], rtol=1e-11, atol=2e-14)
        assert abs(2 * w.sum() - 2.0) < 1e-13
        # exact for even monomials up to degree 2n-2
        for j in range(0, n, 7):
            assert abs(2 * (w * x ** (2 * j)).sum() - 2.0 / (2 * j + 1)) < 1e-12


@pytest.mark.parametrize("tol,dist,key", [(1e-4, 51.96, 0), (1e-5, 259.8, 3),
                                          (1e-6, 86.6, 1)])
def test_select_gl_is_minimal(tol, dist, key):
    g = o.alpha_groups(15, 15, 15, 10.0, tol)
    a, z = g["alpha"][key], g["zcut"][key]
    n = o.select_gl(tol, dist, a, z)

    def err(m):
        x, w = np.polynomial.legendre.leggauss(m)
        x, w = x[x > 0], w[x > 0]
        s = (w * np.exp(-z * z * x * x) * np.sin(2 * a * dist * z * x) / x).sum()
        return abs(s - 0.5 * math.pi * special.erf(a * dist))

    assert n % 2 == 0
    assert err(n) < 0.5 * math.pi * dist * tol
    for m in range(2, n, 2):
        assert err(m) >= 0.5 * math.pi * dist * tol


# ---- spherical rule ----------------------------------------------------------

def sphere_mean(a, b, c):
    """Mean of x^a y^b z^c over the unit sphere (closed form)."""
    if a % 2 or b % 2 or c % 2:
        return 0.0
    def dfact(k):
        return 1.0 if k <= 0 else k * dfact(k - 2)
    return dfact(a - 1) * dfact(b - 1) * dfact(c - 1) / dfact(a + b + c + 1)


@pytest.mark.parametrize("m", [0, 1, 2, 3, 5, 8, 13, 20])
def test_spherical_rule_exactness(m):
    xyz, w = o.sph_rule(m)
    assert abs(w.sum() - 1.0) < 1e-14
    np.testing.assert_allclose((xyz ** 2).sum(1), 1.0, atol=1e-15)
    # one point of every antipodal pair
    key = {tuple(np.round(p, 12)) for p in xyz}
    assert not any(tuple(np.round(-p, 12)) in key for p in xyz)
    t = 2 * m + 1
    for a in range(t + 1):
        for b in range(t + 1 - a):
            for c in range(t + 1 - a - b):
                if (a + b + c) % 2:
                    continue  # odd monomials: the half rule is for even f
                got = (w * xyz[:, 0] ** a * xyz[:, 1] ** b * xyz[:, 2] ** c).sum()
                assert abs(got - sphere_mean(a, b, c)) < 1e-13, (m, a, b, c)


@pytest.mark.parametrize("tol,key", [(1e-4, 0), (1e-4, 3), (1e-5, 1),
                                     (1e-6, 2)])
def test_select_sph_is_minimal(tol, key):
    """The degree the oracle selects is the three-axis walk of
    src/cubature.cu:486-516 re-evaluated with NumPy Gauss-Legendre nodes
    over the oracle's product rules (pinned above by their exactness): on
    each axis the first m >= m0 that passes, every m skipped on that axis
    fails, and m0 starts at half the number of positive roots."""
    g = o.alpha_groups(15, 15, 15, 10.0, tol)
    a, z, dist = g["alpha"][key], g["zcut"][key], g["rmax"][key]
    order = o.select_gl(tol, dist, a, z)
    m_sel = o.select_sph(order, tol, dist, a, z)
    x, w = np.polynomial.legendre.leggauss(order)
    x, w = x[x > 0], w[x > 0]
    target = special.erf(a * dist) / dist

    def err(m, axis):
        xyz, om = o.sph_rule(m)
        v = (om[None, :] * np.cos(2 * a * z * dist * x[:, None]
                                  * xyz[None, :, axis])).sum(1)
        s0 = (w * np.exp(-z * z * x * x) * v).sum() * 4 * a * z / math.pi
        return abs(s0 - target)

    m = x.size // 2
    for axis in range(3):
        start = m
        while err(m, axis) >= tol:
            m += 1
        for mm in range(start, m):
            assert err(mm, axis) >= tol
    assert m_sel == m
    assert err(m_sel, 2) < tol


def test_spherical_rule_not_exact_beyond_degree():
    """Degree t+1 (even) is no longer integrated exactly."""
    m = 4
    xyz, w = o.sph_rule(m)
    t = 2 * m + 1
    errs = [abs((w * xyz[:, 0] ** a * xyz[:, 2] ** (t + 1 - a)).sum()
                - sphere_mean(a, 0, t + 1 - a)) for a in range(0, t + 2, 2)]
    assert max(errs) > 1e-6


# ---- node table: the cubature reproduces erf(a r)/r -------------------------

@pytest.mark.parametrize("tol", [1e-3, 1e-4])
def test_node_table_reproduces_kernel(tol):
    n, cell = 6, 10.0
    kx, ky, kz, w, grp = o.build_nodes(tol, n, n, n, cell)
    g = o.alpha_groups(n, n, n, cell, tol)
    rng = np.random.default_rng(3)
    for gi in range(len(g["width"])):
        sel = grp == gi
        k = np.stack([kx[sel], ky[sel], kz[sel]], 1)
        a = g["alpha"][gi]
        for _ in range(20):
            r = rng.uniform(g["rmin"][gi], g["rmax"][gi])
            d = rng.normal(size=3)
            d *= r / np.linalg.norm(d)
            approx = (w[sel] * np.cos(k @ d)).sum()
            assert abs(approx - special.erf(a * r) / r) < 3 * tol, (gi, r)


# ---- binning (src/glst_force.cu:356-534) -------------------------------------

def bin_numpy(ct, pos, ncell, rcut):
    dt = np.float32 if ct == 1 else np.float64
    inv = dt(1) / dt(rcut)
    v = pos.astype(dt) * inv
    c = np.zeros(pos.shape, dtype=np.int64)
    for d in range(3):
        vd = v[:, d].astype(np.float64)
        cd = np.where(np.isnan(vd), 0,
                      np.trunc(np.clip(vd, -2.0 ** 31, 2.0 ** 31 - 1)))
        c[:, d] = np.clip(cd, 0, ncell[d] - 1)
    cell = (c[:, 0] * ncell[1] + c[:, 1]) * ncell[2] + c[:, 2]
    order = np.argsort(cell, kind="stable")
    counts = np.bincount(cell, minlength=int(np.prod(ncell)))
    points = np.concatenate([[0], np.cumsum(counts)[:-1]])
    return cell, order, points, counts


@pytest.mark.parametrize("ct", [0, 1])
def test_binning_matches_numpy(ct):
    rng = np.random.default_rng(11)
    L, rc = 31.0, 9.0
    pos = rng.uniform(-2, L + 2, (5000, 3))
    pos[:10] = [[0, 0, 0], [9, 18, 27], [np.nextafter(9, 0)] * 3, [-0.5] * 3,
                [-1, -1, -1], [1e15, -1e15, 3], [27.0, 27.0, 27.0],
                [8.999999, 9.000001, 0], [31, 31, 31], [36, 36, 36]]
    ncell = o.grid_dims(ct, L, L, L, rc)
    got = o.bin_atoms(ct, pos[:, 0], pos[:, 1], pos[:, 2], ncell, rc)
    exp = bin_numpy(ct, pos, ncell, rc)
    for a, b in zip(got, exp):
        np.testing.assert_array_equal(a, b)


def test_grid_dims_ct_precision():
    # n = (u32)(CT(L)/CT(rcut)), +1 when it falls short, in CT
    assert o.grid_dims(0, 142.0, 142.0, 142.0, 10.0) == (15, 15, 15)
    assert o.grid_dims(0, 31.04, 31.04, 31.04, 9.0) == (4, 4, 4)
    assert o.grid_dims(0, 30.0, 20.0, 10.0, 10.0) == (3, 2, 1)
    for L, rc in [(0.3, 0.1), (0.7, 0.1), (2.9, 0.29), (100.1, 10.01)]:
        nd = o.grid_dims(0, L, L, L, rc)[0]
        nf = o.grid_dims(1, L, L, L, rc)[0]
        e64 = int(L / rc) + (int(L / rc) * rc < L)
        f, r = np.float32(L), np.float32(rc)
        e32 = int(f / r) + (np.float32(int(f / r)) * r < f)
        assert (nd, nf) == (e64, e32)


# ---- the long-range pieces against brute force --------------------------------

TINY = dict(L=40.0, rc=10.0, tol=1e-3)


@pytest.fixture(scope="module")
def tiny():
    rng = np.random.default_rng(5)
    n = 400
    pos = rng.uniform(0, TINY["L"], (n, 3))
    q = rng.uniform(-1, 1, n)
    q -= q.mean()
    (f, fields) = o.glst(pos[:, 0], pos[:, 1], pos[:, 2], q, TINY["tol"],
                         TINY["L"], TINY["rc"], fields=True)
    ncell = o.grid_dims(0, TINY["L"], TINY["L"], TINY["L"], TINY["rc"])
    cell, order, points, counts = bin_numpy(0, pos, ncell, TINY["rc"])
    kx, ky, kz, w, grp = o.build_nodes(TINY["tol"], *ncell, TINY["rc"])
    g = o.alpha_groups(*ncell, TINY["rc"], TINY["tol"])
    return dict(pos=pos, q=q, f=f, fields=fields, ncell=ncell, cell=cell,
                k=np.stack([kx, ky, kz], 1), w=w, grp=grp, g=g)


def test_sf_brute_force(tiny):
    S = np.zeros((np.prod(tiny["ncell"]), tiny["w"].size), complex)
    np.add.at(S, tiny["cell"],
              tiny["q"][:, None] * np.exp(-1j * tiny["pos"] @ tiny["k"].T))
    sre, sim = tiny["fields"][0], tiny["fields"][1]
    np.testing.assert_allclose(sre, S.real, atol=1e-12)
    np.testing.assert_allclose(sim, S.imag, atol=1e-12)
    # a planted sign error in the restatement is caught by this comparison
    assert not np.allclose(sim, -S.imag, atol=1e-6)


def _cheb(ncell):
    idx = np.stack(np.unravel_index(np.arange(np.prod(ncell)), ncell), 1)
    return np.abs(idx[:, None, :] - idx[None, :, :]).max(-1)


def test_shell_sums_brute_force(tiny):
    S = tiny["fields"][0] + 1j * tiny["fields"][1]
    ch = _cheb(tiny["ncell"])
    R = np.zeros_like(S)
    for gi in range(len(tiny["g"]["width"])):
        sel = tiny["grp"] == gi
        m = (ch > tiny["g"]["r_in"][gi]) & (ch <= tiny["g"]["r_out"][gi])
        R[:, sel] = (m.astype(float) @ S[:, sel]) * tiny["w"][sel]
    np.testing.assert_allclose(tiny["fields"][2], R.real, atol=1e-11)
    np.testing.assert_allclose(tiny["fields"][3], R.imag, atol=1e-11)


def test_far_field_equals_pair_cubature(tiny):
    """Long-range energies/forces equal the pair sum of the cubature kernel
    sum_k W_k cos(k.(r_a - r_b)) over cell pairs in each node's shell."""
    pos, q, k, w, grp = tiny["pos"], tiny["q"], tiny["k"], tiny["w"], tiny["grp"]
    ncell = tiny["ncell"]
    idx = np.stack(np.unravel_index(tiny["cell"], ncell), 1)
    ch = np.abs(idx[:, None, :] - idx[None, :, :]).max(-1)
    far = o.glst(pos[:, 0], pos[:, 1], pos[:, 2], q, TINY["tol"], TINY["L"],
                 TINY["rc"], near=False)
    en = np.zeros(q.size)
    F = np.zeros((q.size, 3))
    for gi in range(len(tiny["g"]["width"])):
        sel = grp == gi
        m = (ch > tiny["g"]["r_in"][gi]) & (ch <= tiny["g"]["r_out"][gi])
        kk, ww = k[sel], w[sel]
        for a in range(q.size):
            b = np.nonzero(m[a])[0]
            if b.size == 0:
                continue
            d = pos[a] - pos[b]
            ph = d @ kk.T
            en[a] += q[a] * (q[b][:, None] * ww * np.cos(ph)).sum()
            F[a] += q[a] * ((q[b][:, None] * ww * np.sin(ph)) @ kk).sum(0)
    np.testing.assert_allclose(far[3], en, atol=1e-10)
    np.testing.assert_allclose(np.stack(far[:3], 1), F, atol=1e-10)


def test_near_field_brute_force(tiny):
    pos, q = tiny["pos"], tiny["q"]
    idx = np.stack(np.unravel_index(tiny["cell"], tiny["ncell"]), 1)
    ch = np.abs(idx[:, None, :] - idx[None, :, :]).max(-1)
    near = o.glst(pos[:, 0], pos[:, 1], pos[:, 2], q, TINY["tol"], TINY["L"],
                  TINY["rc"], far=False)
    d = pos[:, None, :] - pos[None, :, :]
    r = np.sqrt((d ** 2).sum(-1))
    m = (ch <= 1) & ~np.eye(q.size, dtype=bool)
    qq = q[:, None] * q[None, :] * m
    with np.errstate(divide="ignore", invalid="ignore"):
        en = np.where(m, qq / r, 0).sum(1)
        F = (np.where(m, qq / r ** 3, 0)[..., None] * d).sum(1)
    np.testing.assert_allclose(near[3], en, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(np.stack(near[:3], 1), F, rtol=1e-12, atol=1e-12)


# ---- direct Coulomb (utils/coulomb.cu:14-81) ---------------------------------

def test_coulomb_two_charges():
    f = o.coulomb([0.0, 3.0], [0.0, 4.0], [0.0, 0.0], [1.0, -2.0])
    # en_i = q_i q_j / r, F_i = q_i q_j (r_i - r_j)/r^3
    assert f[3][0] == pytest.approx(-2.0 / 5.0, rel=1e-15)
    assert f[3][1] == pytest.approx(-2.0 / 5.0, rel=1e-15)
    assert f[0][0] == pytest.approx(-2.0 * -3.0 / 125.0, rel=1e-15)
    assert f[1][1] == pytest.approx(-2.0 * 4.0 / 125.0, rel=1e-15)


def test_coulomb_numpy():
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, 20, (300, 3))
    q = rng.uniform(-1, 1, 300)
    f = o.coulomb(pos[:, 0], pos[:, 1], pos[:, 2], q)
    d = pos[:, None] - pos[None]
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    qq = q[:, None] * q[None]
    np.testing.assert_allclose(f[3], (qq / r).sum(1), rtol=1e-12)
    np.testing.assert_allclose(np.stack(f[:3], 1),
                               ((qq / r ** 3)[..., None] * d).sum(1), rtol=1e-11)


# ---- whole path against the direct sum -----------------------------------------

def test_glst_vs_coulomb_water1k():
    """Plain GLST in FP64 on the 1,000-water box reaches 2e-6 relative force
    error at tol 1e-5 (the tol-derived error; measured 1.9e-6)."""
    rx, ry, rz, qc = water(1000, 31.04)
    g = o.glst(rx, ry, rz, qc, 1e-5, 31.04, 9.0)
    c = o.coulomb(rx, ry, rz, qc)
    num = sum(((a - b) ** 2).sum() for a, b in zip(g[:3], c[:3]))
    den = sum((b ** 2).sum() for b in c[:3])
    assert math.sqrt(num / den) < 5e-6
    assert abs(g[3].sum() - c[3].sum()) < 1e-5 * abs(c[3]).sum()


def test_long_range_skipped_when_a_dimension_is_small():
    """src/glst_force.cu:667-670: with any n_d <= 2 only the near field is
    computed (reproduced as the reference has it)."""
    rng = np.random.default_rng(9)
    pos = rng.uniform(0, [20, 50, 50], (300, 3))
    q = rng.uniform(-1, 1, 300)
    full = o.glst(pos[:, 0], pos[:, 1], pos[:, 2], q, 1e-4, (20.0, 50.0, 50.0),
                  10.0)
    near = o.glst(pos[:, 0], pos[:, 1], pos[:, 2], q, 1e-4, (20.0, 50.0, 50.0),
                  10.0, far=False)
    for a, b in zip(full, near):
        np.testing.assert_array_equal(a, b)
