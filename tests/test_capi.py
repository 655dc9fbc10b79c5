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


def test_every_declared_symbol_is_exported():
    lib = g.library()
    names = declared("glst_hip.h")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_exported_symbols_via_nm():
    out = subprocess.run(["nm", "-D", "--defined-only", g.LIB_PATH],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (glst_\w+)", out))
    assert set(declared("glst_hip.h")) <= exported


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
