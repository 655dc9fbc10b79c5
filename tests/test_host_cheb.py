"""The host plan of the Chebyshev form of the mixed structure factors (no
GPU): tests/host/cheb_check.cpp, built with g++, prints it.

- bessel_j (Miller's backward recurrence) against scipy.special.jv;
- cheb_terms: the smallest M whose tail bound 2 sum_{m >= M} |J_m(z)| is
  below 3e-8, checked with SciPy's J_m;
- every planned coefficient matrix: sum_m C[i][m] T_m(x) reproduces
  exp(-i kappa_i rho x) on [-1, 1] (Jacobi-Anger), so a wrong sign, index
  or order fails; and the direct sum is kept where the expansion would cost
  more issue slots.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "gauss-legendre-spherical-t_amd", "csrc")


@pytest.fixture(scope="module")
def plan(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the ROCm headers")
    exe = str(tmp_path_factory.mktemp("cheb") / "cheb_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-I" + os.path.join(REPO, "include"),
                    "-I" + CSRC,
                    os.path.join(REPO, "tests", "host", "cheb_check.cpp"),
                    os.path.join(CSRC, "numerics.cpp"), "-o", exe],
                   check=True, timeout=300)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120,
                         check=True).stdout.splitlines()
    bessel, terms, cases, cur = {}, {}, {}, None
    for line in out:
        w = line.split()
        if w[0] == "bessel":
            bessel[float(w[1])] = np.array([float(v) for v in w[2:]])
        elif w[0] == "terms":
            terms[float(w[1])] = int(w[2])
        elif w[0] == "case":
            cur = {"h": float(w[3]), "rho": float(w[5]), "groups": {}}
            cases[w[1]] = cur
        elif w[0] == "group":
            g = int(w[1])
            cur["groups"][g] = {"nr": int(w[3]), "m": int(w[5]),
                                "kappa": np.array([float(v) for v in w[9:]]),
                                "rows": []}
        elif w[0] == "row":
            cur["groups"][int(w[1])]["rows"].append(
                np.array([float(v) for v in w[3:]]))
    return bessel, terms, cases


def test_bessel_matches_scipy(plan):
    from scipy.special import jv
    bessel, _, _ = plan
    for z, J in bessel.items():
        ref = jv(np.arange(J.size), z)
        assert np.abs(J - ref).max() < 1e-14, z


def test_terms_meet_the_tail_bound(plan):
    from scipy.special import jv
    _, terms, _ = plan
    for z, M in terms.items():
        tail = lambda m: 2 * np.abs(jv(np.arange(m, m + 80), z)).sum()
        if M == 41:  # more than the kernels hold (40): direct sums
            assert tail(40) >= 3e-8, z
            continue
        assert tail(M) < 3e-8, z
        assert M == 1 or tail(M - 1) >= 3e-8, z


def test_planned_matrices_reproduce_the_phase_factors(plan):
    _, _, cases = plan
    x = np.linspace(-1.0, 1.0, 2001)
    used = 0
    for name, c in cases.items():
        rho = c["rho"]
        assert rho == pytest.approx(1.001 * np.sqrt(3) / 2 * c["h"], rel=1e-15)
        for g, grp in c["groups"].items():
            nr, M = grp["nr"], grp["m"]
            cost_cheb = 2 * M + 4 + nr * M / 32
            if M == 0:
                continue
            used += 1
            assert cost_cheb < 6 * nr, (name, g)
            assert len(grp["rows"]) == nr
            T = np.cos(np.outer(np.arange(M), np.arccos(x)))  # T_m(x)
            for i, row in enumerate(grp["rows"]):
                want = np.exp(-1j * grp["kappa"][i] * rho * x)
                re = row[0::2] @ T[0::2]
                im = row[1::2] @ T[1::2]
                err = np.abs(re + 1j * im - want).max()
                # truncation (3e-8) plus the FP32 coefficients
                assert err < 4e-7, (name, g, i, err)
    assert used >= 4
