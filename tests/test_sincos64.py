"""The FP64 path's sine and cosine (csrc/sincos64.hpp: Cody-Waite reduction
by pi/2 and Taylor series on [-pi/4, pi/4]) against x87 extended-precision
sinl/cosl (NumPy longdouble), on the host.

Bars: at most 2 ulp of the larger of |sin x|, |cos x| over 0 <= |x| <= 1e4
(the phases of the force step stay below that), exact quadrant signs at the
multiples of pi/2, and sin^2 + cos^2 = 1 to 4e-16.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "gauss-legendre-spherical-t_amd", "csrc")


@pytest.fixture(scope="module")
def sincos_bin(tmp_path_factory):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the ROCm headers")
    exe = str(tmp_path_factory.mktemp("sc") / "sincos_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off",
                    "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I" + CSRC,
                    os.path.join(REPO, "tests", "host", "sincos_check.cpp"),
                    "-o", exe], check=True, timeout=300)
    return exe


def _run(exe, x):
    out = subprocess.run([exe], input="\n".join(repr(float(v)) for v in x),
                         capture_output=True, text=True, check=True,
                         timeout=300).stdout.split()
    v = np.array(out, dtype=np.float64).reshape(-1, 2)
    return v[:, 0], v[:, 1]


def test_sincos_accuracy(sincos_bin):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-1e4, 1e4, 20000),
                        rng.uniform(-20.0, 20.0, 20000),
                        rng.uniform(-1e-3, 1e-3, 2000),
                        [0.0, -0.0, 1e-300, 0.785398, 0.7853981633974483]])
    s, c = _run(sincos_bin, x)
    xl = x.astype(np.longdouble)
    sl, cl = np.sin(xl), np.cos(xl)
    scale = np.maximum(np.abs(sl), np.abs(cl)).astype(np.float64)
    ulp = np.spacing(scale)
    es = np.abs((s - sl).astype(np.float64)) / ulp
    ec = np.abs((c - cl).astype(np.float64)) / ulp
    assert es.max() <= 2.0, (es.max(), x[es.argmax()])
    assert ec.max() <= 2.0, (ec.max(), x[ec.argmax()])
    assert np.abs(s * s + c * c - 1.0).max() < 4e-16


def test_sincos_quadrants(sincos_bin):
    """Exact signs and values around every quadrant boundary, positive and
    negative n."""
    n = np.arange(-40, 41)
    x = n * (np.pi / 2)
    s, c = _run(sincos_bin, x)
    q = np.mod(n, 4)
    np.testing.assert_allclose(s, np.select([q == 0, q == 1, q == 2, q == 3],
                                            [0, 1, 0, -1]), atol=1e-14)
    np.testing.assert_allclose(c, np.select([q == 0, q == 1, q == 2, q == 3],
                                            [1, 0, -1, 0]), atol=1e-14)
