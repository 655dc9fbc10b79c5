"""CHARMM PSF/COR fixed-column readers (utils/io.cpp:35-117) and recentring
(utils/utils.cpp:35-58) through the C ABI of include/glst_io.h.  CPU only."""
import ctypes

import numpy as np
import pytest

import glst_amd as g

lib = g.library()
lib.glst_psf_natom.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint64)]
lib.glst_read_psf_charges.argtypes = [ctypes.c_char_p, ctypes.c_uint64,
                                      ctypes.c_void_p]
lib.glst_read_cor.argtypes = [ctypes.c_char_p, ctypes.c_uint64] + [ctypes.c_void_p] * 3
lib.glst_recenter.argtypes = [ctypes.c_uint64] + [ctypes.c_void_p] * 3


def write_psf(path, q):
    lines = ["PSF EXT", "", "       1 !NTITLE", "* synthetic", ""]
    lines.append(f"{len(q):10d} !NATOM")
    for i, qi in enumerate(q):
        # columns 50..63 hold the charge (14 wide)
        head = f"{i + 1:10d} WAT      1        OH2      OT      "
        head = head.ljust(50)
        lines.append(head + f"{qi:14.6f}" + "       15.9994           0")
    lines.append("")
    lines.append("       0 !NBOND: bonds")
    open(path, "w").write("\n".join(lines) + "\n")


def write_cor(path, pos):
    lines = ["* synthetic coordinates", "*"]
    lines.append(f"{len(pos):10d}  EXT")
    for i, (x, y, z) in enumerate(pos):
        head = f"{i + 1:10d}{1:10d}  WAT       OH2".ljust(40)
        lines.append(head + f"{x:20.10f}{y:20.10f}{z:20.10f}  W         1"
                     "               0.0000000000")
    open(path, "w").write("\n".join(lines) + "\n")


def test_psf_cor_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    n = 57
    q = np.round(rng.uniform(-1, 1, n), 6)
    pos = np.round(rng.uniform(-30, 30, (n, 3)), 10)
    write_psf(tmp_path / "s.psf", q)
    write_cor(tmp_path / "s.cor", pos)
    nat = ctypes.c_uint64()
    assert lib.glst_psf_natom(str(tmp_path / "s.psf").encode(), ctypes.byref(nat)) == 0
    assert nat.value == n
    qq = np.zeros(n)
    assert lib.glst_read_psf_charges(str(tmp_path / "s.psf").encode(), n,
                                     qq.ctypes.data) == 0
    np.testing.assert_array_equal(qq, q)
    r = [np.zeros(n) for _ in range(3)]
    assert lib.glst_read_cor(str(tmp_path / "s.cor").encode(), n,
                             *[a.ctypes.data for a in r]) == 0
    np.testing.assert_allclose(np.stack(r, 1), pos, atol=1e-10)
    assert lib.glst_recenter(n, *[a.ctypes.data for a in r]) == 0
    np.testing.assert_allclose(np.stack(r, 1), pos - pos.min(0), atol=1e-10)
    assert min(a.min() for a in r) == 0.0


def test_count_mismatch_and_missing(tmp_path):
    write_psf(tmp_path / "s.psf", [0.1, -0.1, 0.2])
    write_cor(tmp_path / "s.cor", [[0, 0, 0], [1, 1, 1]])
    q = np.zeros(3)
    r = [np.zeros(3) for _ in range(3)]
    # COR holds 2 atoms, the PSF 3: "Number of atoms differs" (utils/io.cpp:72)
    assert lib.glst_read_cor(str(tmp_path / "s.cor").encode(), 3,
                             *[a.ctypes.data for a in r]) == 5
    assert lib.glst_read_psf_charges(str(tmp_path / "s.psf").encode(), 2,
                                     q.ctypes.data) == 5
    nat = ctypes.c_uint64()
    assert lib.glst_psf_natom(str(tmp_path / "none.psf").encode(),
                              ctypes.byref(nat)) == 5
