"""The engine's host planning under AddressSanitizer and UBSan (no GPU).

tests/host/plan_check.cpp drives the host code a handle and its autotuner
run (setup_system, layout_nodes, make_splits for every back-projection split
target the autotuner tries, plan_pipeline for both shell-sum tile limits and
the matrix-core tiles of 1, 2 and 4 nodes, plan_chebyshev for the mixed
structure factors) at the 3,000-charge workloads (L 31.04 and 44.7), the
288,000- and the 1,000,796-charge workloads in both calculation types, and
checks every table index against the launch sizes.  It is built
here with g++ (host code only) and the sanitizers, so an out-of-bounds
access in that code fails this test.
"""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "gauss-legendre-spherical-t_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None or
                    not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the ROCm headers")
def test_host_planning_is_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "plan_check")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", "-I" + os.path.join(REPO, "include"),
           "-I" + CSRC, "-fsanitize=address,undefined",
           "-fno-sanitize-recover=undefined",
           os.path.join(REPO, "tests", "host", "plan_check.cpp"),
           os.path.join(CSRC, "numerics.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
    assert r.stdout.count("checked") == 8
