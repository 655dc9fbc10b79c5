"""The C++ facade (include/glst_force.hpp) and the full_example driver on the
GPU.  The driver's report is read with label regexes equal to the ones a
reference log parser uses (stage labels of test/full_example.cu:215-233)."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "gauss-legendre-spherical-t_amd", "bin")
FLOAT = r"[-+]?(?:\d+(?:\.\d*)?|\.\d+)(?:[eE][-+]?\d+)?"


def test_cpp_facade(cuda):
    r = subprocess.run([os.path.join(BIN, "glst_cpp_test")], capture_output=True,
                       text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.parametrize("ct", ["double", "float"])
def test_full_example_report(cuda, ct):
    env = dict(os.environ, GLST_BENCH_ITERS="3", GLST_CT=ct,
               GLST_BENCH_TERMINAL="1")
    r = subprocess.run([os.path.join(BIN, "full_example"), "water:1000", "1e-5",
                        "31.04", "9.0"], capture_output=True, text=True,
                       timeout=120, env=env)
    out = r.stdout
    assert r.returncode == 0, out + r.stderr
    assert re.search(r"^BENCH_INPUT\s+.*tol=1e-05", out, re.M)
    assert re.search(r"Number of atoms:\s*3000", out)
    assert re.search(r"Number of cells:\s*4,\s*4,\s*4", out)
    assert re.search(r"Number of alpha groups:\s*2", out)
    assert re.search(r"Finished\s+3\s+calculations", out)
    for label in ("Assign atoms to cells", "Calculate structure factors",
                  "Sum remote structure factors",
                  "Calculate long-range energy and forces",
                  "Calculate short-range energy and forces",
                  "Communicate energies and forces", "Long-Range GLST Runtime",
                  "Total GLST Runtime"):
        assert re.search(rf"{re.escape(label)}:\s*({FLOAT})\s*ms\s*\(\+/-\s*"
                         rf"({FLOAT})\s*ms\)", out), label
    assert re.search(rf"Coulomb Runtime:\s*({FLOAT})\s*ms", out)
    # normalised RMS force error of each component is below tol
    for sec in ("X Force", "Y Force", "Z Force"):
        block = re.search(rf"{sec}\s+Error:(.*?)(?:\n\s*\n)", out, re.S).group(1)
        m = re.search(rf"Root Mean Squared:\s*({FLOAT})\s*\(Norm:\s*({FLOAT})\)",
                      block)
        assert float(m.group(2)) < 1e-5, (sec, m.group(0))
