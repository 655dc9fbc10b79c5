"""Womersley design files (src/cubature.cu:422-523) through the C ABI, on the
host.

Exact symmetric designs with closed-form points are written into a temporary
directory in the reference's file format (ssTTT.NNNNN, one point per line in
three 25-character columns, the second half of the file the antipodes of the
first): an antipodal pair (degree 1), the octahedron (degree 3) and the
icosahedron (degree 5).  The loader must return the first N/2 rows, the
half-design with equal weights must integrate every even monomial up to its
degree exactly (and not the next even degree), and the degree selection with
these files must follow the three-axis test of src/cubature.cu:486-516,
re-evaluated here with NumPy's Gauss-Legendre nodes.
"""
import math
import os

import numpy as np
import pytest
from scipy import special

import glst_amd as g

PHI = (1.0 + 5.0 ** 0.5) / 2.0
HALF = {
    1: [[0.0, 0.0, 1.0]],
    3: [[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]],
    5: [[0.0, 1.0, PHI], [0.0, -1.0, PHI], [1.0, PHI, 0.0], [-1.0, PHI, 0.0],
        [PHI, 0.0, 1.0], [PHI, 0.0, -1.0]],
}


def _unit(p):
    p = np.asarray(p, dtype=float)
    return p / np.linalg.norm(p, axis=1)[:, None]


def _write(d, t, half, name=None):
    full = np.concatenate([half, -half])
    path = os.path.join(d, name or "ss%03d.%05d" % (t, len(full)))
    with open(path, "w") as fh:
        for p in full:
            fh.write("".join("%25.17e" % v for v in p) + "\n")
    return path


@pytest.fixture
def designs(tmp_path):
    d = str(tmp_path)
    for t, half in HALF.items():
        _write(d, t, _unit(half))
    # a file of another kind in the same directory is skipped
    with open(os.path.join(d, "README"), "w") as fh:
        fh.write("not a design\n")
    yield d
    g.set_tdesign_dir(None)


def sphere_mean(a, b, c):
    """Mean of x^a y^b z^c over the unit sphere (closed form)."""
    if a % 2 or b % 2 or c % 2:
        return 0.0

    def dfact(k):
        return 1.0 if k <= 0 else k * dfact(k - 2)
    return dfact(a - 1) * dfact(b - 1) * dfact(c - 1) / dfact(a + b + c + 1)


def _even_monomial_errors(pts, deg):
    w = 1.0 / len(pts)
    out = []
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            c = deg - a - b
            got = w * (pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c).sum()
            out.append(abs(got - sphere_mean(a, b, c)))
    return out


@pytest.mark.parametrize("t", [1, 3, 5])
def test_loader_reads_first_half_and_is_exact_to_its_degree(designs, t):
    pts = g.read_tdesign(designs, t)
    ref = _unit(HALF[t])
    assert pts.shape == ref.shape
    np.testing.assert_allclose(pts, ref, rtol=0, atol=1e-16)
    for deg in range(0, t + 1, 2):
        assert max(_even_monomial_errors(pts, deg)) < 1e-15, (t, deg)
    assert max(_even_monomial_errors(pts, t + 1)) > 1e-3  # degree is t


def test_loader_errors(tmp_path):
    d = str(tmp_path)
    with pytest.raises(g.GLSTError) as e:
        g.read_tdesign(d, 3)  # no file of that degree
    assert e.value.status == 5
    path = _write(d, 3, _unit(HALF[3]))
    with open(path) as fh:
        lines = fh.readlines()
    with open(path, "w") as fh:  # second row cut short
        fh.write(lines[0] + lines[1][:40] + "\n")
    with pytest.raises(g.GLSTError) as e:
        g.read_tdesign(d, 3)
    assert e.value.status == 5
    with pytest.raises(g.GLSTError) as e:
        g.read_tdesign(str(tmp_path / "missing"), 3)
    assert e.value.status == 5


def _axis_err(order, pts, axis, dist, alpha, zcut):
    """|(4 a z / (pi M)) sum_i w_i e^{-z^2 x_i^2} sum_m cos(2 a z d x_i s_m,j)
    - erf(a d)/d| (src/cubature.cu:494-515), NumPy nodes."""
    x, w = np.polynomial.legendre.leggauss(order)
    x, w = x[x > 0], w[x > 0]
    v = np.cos(2 * alpha * zcut * dist * x[:, None] * pts[None, :, axis]).sum(1)
    s0 = (w * np.exp(-zcut * zcut * x * x) * v).sum()
    s0 *= 4 * alpha * zcut / (math.pi * len(pts))
    return abs(s0 - special.erf(alpha * dist) / dist)


@pytest.mark.parametrize("order,alpha,dist,zcut,tol,path", [
    (2, 0.3, 3.0, 1.0, 0.05, [0, 0, 1]),  # the z axis moves to degree 3
    (2, 0.5, 3.0, 1.0, 0.1, [1, 1, 1]),   # the x axis needs degree 3
    (4, 0.5, 3.0, 1.0, 0.02, [2, 2, 2]),  # from m0 = 1 to degree 5
])
def test_selection_with_design_files(designs, order, alpha, dist, zcut, tol,
                                     path):
    halves = {(t - 1) // 2: g.read_tdesign(designs, t) for t in (1, 3, 5)}
    # the reference's walk: per axis, the first m >= m0 that passes, with m0
    # starting at len(x)/2 and carried to the next axis
    m = (order // 2) // 2
    walk = []
    for axis in range(3):
        while _axis_err(order, halves[m], axis, dist, alpha, zcut) >= tol:
            m += 1
        walk.append(m)
    assert walk == path
    # every degree skipped on an axis failed there
    start = (order // 2) // 2
    for axis, stop in enumerate(walk):
        for mm in range(start, stop):
            assert _axis_err(order, halves[mm], axis, dist, alpha, zcut) >= tol
        start = stop
    g.set_tdesign_dir(designs)
    got_m, ndir = g.select_spherical_degree(order, tol, dist, alpha, zcut)
    assert got_m == walk[-1]
    assert ndir == len(HALF[2 * got_m + 1])


def test_selection_needs_the_next_file(designs):
    """A tolerance no file up to degree 5 meets asks for degree 7, which the
    directory lacks (src/cubature.cu:441-459 then fails to open a file)."""
    g.set_tdesign_dir(designs)
    with pytest.raises(g.GLSTError) as e:
        g.select_spherical_degree(2, 1e-3, 3.0, 0.3, 1.0)
    assert e.value.status == 5
    assert "degree 7" in str(e.value)


def test_selection_without_files_uses_the_product_rule():
    g.set_tdesign_dir(None)
    m, ndir = g.select_spherical_degree(2, 0.05, 3.0, 0.3, 1.0)
    xyz, w = g.spherical_rule(m)
    assert ndir == len(xyz)
