"""The C ABI library loads without a GPU, exports every entry point that
include/glst_hip.h declares, and rejects bad arguments before touching the
device."""
import os
import re
import subprocess

import pytest

import glst_amd as g

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(os.path.join(REPO, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)  # drop comments
    return sorted(set(re.findall(
        r"(?:glst_status_t|const char\s*\*)\s*(glst_[a-z0-9_]+)\s*\(", text)))


HEADERS = ["glst_hip.h", "glst_batch.h", "glst_serve.h", "glst_io.h"]


def test_every_declared_symbol_is_exported():
    lib = g.library()
    names = [n for h in HEADERS for n in declared(h)]
    assert len(declared("glst_hip.h")) > 30
    assert len(declared("glst_batch.h")) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_exported_symbols_via_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", g.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (glst_\w+)", out))
    for h in HEADERS:
        assert set(declared(h)) <= exported, h


def test_version():
    assert "gfx950" in g.version()


@pytest.mark.parametrize("kw,status", [
    (dict(natom=10, tol=0.0, box=10.0, rcut=5.0), 1),
    (dict(natom=10, tol=1e-4, box=-1.0, rcut=5.0), 1),
    (dict(natom=10, tol=1e-4, box=10.0, rcut=0.0), 1),
])
def test_create_rejects_bad_arguments(kw, status):
    with pytest.raises(g.GLSTError) as e:
        g.GLSTForce(kw["natom"], kw["tol"], kw["box"], rcut=kw["rcut"],
                    stream=0)
    assert e.value.status == status
    assert g._lib.glst_global_last_error()


def test_create_ncell_rejects_zero_cells():
    with pytest.raises(g.GLSTError) as e:
        g.GLSTForce(10, 1e-4, 10.0, ncell=(0, 1, 1), stream=0)
    assert e.value.status == 1


def test_grid_dims_rejects_nonpositive():
    with pytest.raises(g.GLSTError):
        g.grid_dims(0, 0.0, 1.0, 1.0, 1.0)


def test_null_handle_is_invalid():
    lib = g.library()
    assert lib.glst_calc_sf(None) == 1
    assert lib.glst_destroy(None) == 0


def test_tdesign_dir_missing_is_io_error(tmp_path):
    """A design directory without the needed file fails like
    src/cubature.cu:427-428 / :445 (as GLST_ERR_IO), before any device work."""
    g.set_tdesign_dir(str(tmp_path / "does-not-exist"))
    try:
        with pytest.raises(g.GLSTError) as e:
            g.GLSTForce(30, 1e-4, 40.0, rcut=10.0, stream=0)
        # without a GPU the handle may fail at the device query first
        assert e.value.status in (2, 5)
    finally:
        g.set_tdesign_dir(None)


def test_single_hip_runtime_in_process():
    """Importing the bindings before torch must not load a second HIP/HSA
    runtime (two runtimes in one process see no device)."""
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import glst_amd, torch\n"
        "m = open('/proc/self/maps').read().splitlines()\n"
        "libs = {l.split()[-1] for l in m if 'libamdhip64' in l}\n"
        "print(len(libs))\n" % os.path.join(REPO, "gauss-legendre-spherical-t_amd"))
    out = subprocess.run(["python", "-c", code], capture_output=True, text=True,
                         check=True, timeout=300).stdout.strip()
    assert out == "1"


# ---- batched engine: host-side planning (no GPU needed) ----------------------

def test_batch_device_bytes_grows_with_systems():
    one = g.batch_device_bytes([3000], [31.04], 9.0, 1e-4)
    two = g.batch_device_bytes([3000, 3000], [31.04, 31.04], 9.0, 1e-4)
    assert 0 < one < two
    # the mixed path stores FP32 fields
    assert g.batch_device_bytes([3000], [31.04], 9.0, 1e-4, ct="float") < one


def test_batch_device_bytes_no_far_field_system():
    # a 14.4 A box with rcut 9 has a 1x1x1 grid: near field only, no fields
    small = g.batch_device_bytes([300], [14.4], 9.0, 1e-4)
    assert 0 < small < 1_000_000


@pytest.mark.parametrize("natoms,boxes,rcuts,tol", [
    ([], [], [], 1e-4),            # no systems
    ([0], [31.0], [9.0], 1e-4),    # a system without atoms
    ([30], [-1.0], [9.0], 1e-4),   # negative box
    ([30], [31.0], [0.0], 1e-4),   # zero rcut
    ([30], [31.0], [9.0], 0.0),    # zero tol
])
def test_batch_device_bytes_rejects_bad_arguments(natoms, boxes, rcuts, tol):
    import numpy as np
    with pytest.raises((g.GLSTError, ValueError)) as e:
        g.batch_device_bytes(np.array(natoms, dtype=np.uint32),
                             np.array(boxes, dtype=np.float64),
                             np.array(rcuts, dtype=np.float64), tol)
    if isinstance(e.value, g.GLSTError):
        assert e.value.status == 1


def test_batch_create_rejects_bad_arguments_before_the_device():
    with pytest.raises(g.GLSTError) as e:
        g.GLSTBatch([30, 30], [31.0, -2.0], 9.0, 1e-4, stream=0)
    assert e.value.status == 1


def test_server_create_rejects_bad_arguments_before_the_device():
    for kw in (dict(tol=0.0, budget_bytes=1 << 30),
               dict(tol=1e-4, budget_bytes=0)):
        with pytest.raises(g.GLSTError) as e:
            g.GLSTServer(stream=0, **kw)
        assert e.value.status == 1
