"""bench.py on the host: the parts that need no GPU (argument handling, the
error metrics) so that a broken bench line shows up before a GPU run."""
import os
import subprocess
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"),
                        "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_error_metrics():
    import bench
    rng = np.random.default_rng(0)
    ref = [rng.normal(size=100) for _ in range(4)]
    assert bench.rel_force_err(ref, ref) == 0.0
    got = [r * (1 + 1e-6) for r in ref]
    assert abs(bench.rel_force_err(got, ref) - 1e-6) < 1e-12
    e = bench.ref_style_errors(got, ref)
    assert set(e) == {"fx", "fy", "fz", "en"}
    assert all(abs(v - 1e-6) < 1e-9 for v in e.values())


def test_error_report_finds_the_worst_atom():
    """One atom's force off by 1e-3 of its magnitude, the rest exact: the
    per-atom maximum names it, the median stays zero, and the reference
    report's absolute maximum of that component is the planted error."""
    import bench
    rng = np.random.default_rng(1)
    ref = [rng.normal(size=200) for _ in range(4)]
    got = [r.copy() for r in ref]
    mag = np.sqrt(sum(r[17] ** 2 for r in ref[:3]))
    got[1][17] += 1e-3 * mag
    rep = bench.error_report(got, ref)
    assert abs(rep["per_atom_rel_force_err"]["max"] - 1e-3) < 1e-12
    assert rep["per_atom_rel_force_err"]["p50"] == 0.0
    assert abs(rep["fy"]["abs_max"] - 1e-3 * mag) < 1e-15
    assert rep["fx"]["abs_max"] == 0.0 and rep["en"]["rms"] == 0.0
    rms_y = np.sqrt((got[1] ** 2).mean())
    assert abs(rep["fy"]["abs_max_norm"] - 1e-3 * mag / rms_y) < 1e-15
