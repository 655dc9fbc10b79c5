"""The product's host numerics (C ABI of lib/libglst_hip.so, no GPU needed)
against the pinned oracle: erfc_inv, alpha groups, grid dimensions,
Gauss-Legendre rules and order selection, the spherical rule, and the
synthetic water generator."""
import math

import numpy as np
import pytest

import glst_amd as g
from oracle import oracle as o


def test_erfc_inv_bit_identical():
    for y in (1e-9, 9e-5, 1e-4, 4.43e-4, 0.01, 0.5, 1.3):
        assert g.erfc_inv(y, np.finfo(float).eps) == o.erfc_inv(y)


def test_erfc_inv_errors():
    for y in (0.0, 2.0, -1.0):
        with pytest.raises(g.GLSTError) as e:
            g.erfc_inv(y, 1e-16)
        assert e.value.status == 3  # GLST_ERR_NUMERICS


@pytest.mark.parametrize("n,cell", [(4, 9.0), (15, 10.0), (18, 12.0), (2, 5.0),
                                    (7, 3.3), (40, 10.0)])
@pytest.mark.parametrize("tol", [1e-3, 1e-4, 1e-5, 1e-6])
def test_alpha_groups_identical(n, cell, tol):
    a = g.alpha_groups(n, n, n, cell, cell, cell, tol)
    b = o.alpha_groups(n, n, n, cell, tol)
    for key in ("width", "rmax", "alpha", "zcut", "r_in", "r_out"):
        np.testing.assert_array_equal(a[key], b[key])


def test_grid_dims_identical():
    rng = np.random.default_rng(4)
    for _ in range(300):
        L = rng.uniform(1, 300, 3)
        rc = rng.uniform(0.5, 20)
        for ct in (0, 1):
            assert g.grid_dims(ct, *L, rc) == o.grid_dims(ct, *L, rc)
    # exact multiples and float-rounding cases
    for L, rc in [(30.0, 10.0), (0.3, 0.1), (0.7, 0.1), (142.0, 10.0),
                  (31.04, 9.0), (215.0, 12.0)]:
        for ct in (0, 1):
            assert g.grid_dims(ct, L, L, L, rc) == o.grid_dims(ct, L, L, L, rc)


@pytest.mark.parametrize("n", list(range(1, 60)) + [100, 161, 162, 250, 324])
def test_gauss_legendre_matches(n):
    x, w = g.gauss_legendre(n)
    xo, wo = o.gauss_legendre(n)
    np.testing.assert_allclose(x, xo, atol=3e-16 * n, rtol=0)
    np.testing.assert_allclose(w, wo, rtol=1e-11, atol=2e-14)


def test_gl_order_selection_matches():
    for (n, cell) in [(4, 9.0), (15, 10.0), (18, 12.0)]:
        for tol in (1e-4, 1e-5, 1e-6):
            gr = o.alpha_groups(n, n, n, cell, tol)
            for i in range(len(gr["width"])):
                args = (tol, gr["rmax"][i], gr["alpha"][i], gr["zcut"][i])
                assert g.select_gl_order(*args) == o.select_gl(*args)


@pytest.mark.parametrize("m", [0, 1, 2, 3, 4, 7, 16, 27, 40, 52])
def test_spherical_rule_matches(m):
    xyz, w = g.spherical_rule(m)
    xo, wo = o.sph_rule(m)
    assert xyz.shape == xo.shape
    np.testing.assert_allclose(xyz, xo, atol=1e-15)
    np.testing.assert_allclose(w, wo, rtol=1e-12, atol=1e-17)
    assert xyz.shape[0] == (m + 1) ** 2


def test_water_box():
    rx, ry, rz, q = g.make_water_box(500, 25.0, seed=3, nion=7)
    n = 3 * 500 + 14
    assert rx.size == n
    assert abs(q.sum()) < 1e-9
    assert rx.min() == 0.0 and ry.min() == 0.0 and rz.min() == 0.0
    pos = np.stack([rx, ry, rz], 1)
    O, H1, H2 = pos[0:1500:3], pos[1:1500:3], pos[2:1500:3]
    np.testing.assert_allclose(np.linalg.norm(H1 - O, axis=1), 1.0, atol=1e-12)
    np.testing.assert_allclose(np.linalg.norm(H2 - O, axis=1), 1.0, atol=1e-12)
    cosang = ((H1 - O) * (H2 - O)).sum(1)
    np.testing.assert_allclose(np.degrees(np.arccos(cosang)), 109.47, atol=1e-9)
    np.testing.assert_array_equal(q[:3], [-0.8476, 0.4238, 0.4238])
    np.testing.assert_array_equal(q[1500:1507], 1.0)
    np.testing.assert_array_equal(q[1507:], -1.0)
    # deterministic for a seed, different across seeds
    again = g.make_water_box(500, 25.0, seed=3, nion=7)
    for a, b in zip((rx, ry, rz, q), again):
        np.testing.assert_array_equal(a, b)
    other = g.make_water_box(500, 25.0, seed=4, nion=7)
    assert not np.array_equal(rx, other[0])
    # oxygens uniform: mean near L/2 (after the small recentring shift)
    assert abs(O[:, 0].mean() - 12.5) < 1.5


def test_orientation_uniform():
    """Bisector directions of random waters cover the sphere uniformly."""
    rx, ry, rz, _ = g.make_water_box(20000, 100.0, seed=5)
    pos = np.stack([rx, ry, rz], 1)
    O, H1, H2 = pos[0::3], pos[1::3], pos[2::3]
    b = (H1 + H2) / 2 - O
    b /= np.linalg.norm(b, axis=1)[:, None]
    # mean of a uniform unit vector is 0, mean of z^2 is 1/3
    assert np.abs(b.mean(0)).max() < 0.02
    assert abs((b[:, 2] ** 2).mean() - 1 / 3) < 0.01
