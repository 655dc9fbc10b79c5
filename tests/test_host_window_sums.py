"""The window-sum rule of the small-grid shell sums, on the CPU.

The kernel (rmt_wave_tile in csrc/kernels.hip) sums each cube |i' - i| <= r
of a grid of at most 8 cells a side as three passes along lines held in
8-slot registers: the slots past the grid hold zero, P is the running sum
of the line, and the window of slot i is P[min(i + r, 7)] - P[i - r - 1]
(only P[min(i + r, 7)] when i - r - 1 < 0), r clamped to 7.  The shell of a
node is the outer cube (r_out) minus the inner one (r_in).  This restates
that rule in NumPy and checks it against brute-force cube sums: shells for
six (r_in, r_out) pairs on the grids up to 8 cells a side that have an
8-cell side or a cell count divisible by 7, and cubes of every half-width
0 .. 8 on every grid up to 4 cells a side.  A window one slot short fails
the same check.  The GPU kernel itself is checked against the oracle in
tests/test_gpu_rmt_wave.py.
"""
import itertools

import numpy as np


def _window8(line, r):
    """The kernel's window of an 8-slot line (zeros past the grid)."""
    P = np.cumsum(line, axis=-1)
    r = min(r, 7)
    out = np.empty_like(P)
    for i in range(8):
        hi = min(i + r, 7)
        lo = i - r - 1
        out[..., i] = P[..., hi] - (P[..., lo] if lo >= 0 else 0.0)
    return out


def _cube_by_passes(Y, r):
    """z, then x, then y windows of a grid padded to 8 x 8 x 8."""
    nx, ny, nz = Y.shape
    pad = np.zeros((8, 8, 8), dtype=Y.dtype)
    pad[:nx, :ny, :nz] = Y
    w = _window8(pad, r)                                    # along z
    w = np.moveaxis(_window8(np.moveaxis(w, 0, -1), r), -1, 0)   # along x
    w = np.moveaxis(_window8(np.moveaxis(w, 1, -1), r), -1, 1)   # along y
    return w[:nx, :ny, :nz]


def _cube_brute(Y, r):
    nx, ny, nz = Y.shape
    out = np.zeros_like(Y)
    for x, y, z in itertools.product(range(nx), range(ny), range(nz)):
        out[x, y, z] = Y[max(0, x - r):x + r + 1, max(0, y - r):y + r + 1,
                         max(0, z - r):z + r + 1].sum()
    return out


def test_window_rule_equals_brute_force_cubes():
    rng = np.random.default_rng(5)
    for nx, ny, nz in itertools.product(range(1, 9), repeat=3):
        if (nx * ny * nz) % 7 and max(nx, ny, nz) < 8:
            continue  # a spread of shapes, every one with an 8-cell side
        Y = rng.standard_normal((nx, ny, nz)) + 1j * rng.standard_normal(
            (nx, ny, nz))
        for r_in, r_out in [(0, 1), (1, 2), (2, 4), (3, 7), (5, 9), (0, 12)]:
            shell = _cube_by_passes(Y, r_out) - _cube_by_passes(Y, r_in)
            ref = _cube_brute(Y, r_out) - _cube_brute(Y, r_in)
            np.testing.assert_allclose(shell, ref, rtol=0, atol=1e-12)


def test_window_rule_every_small_cube():
    """All grids up to 4 cells a side, every half-width 0 .. 8."""
    rng = np.random.default_rng(6)
    for nx, ny, nz in itertools.product(range(1, 5), repeat=3):
        Y = rng.standard_normal((nx, ny, nz))
        for r in range(9):
            np.testing.assert_allclose(_cube_by_passes(Y, r),
                                       _cube_brute(Y, r), rtol=0, atol=1e-12)


def test_window_rule_catches_an_off_by_one():
    """A window one slot too short fails the brute-force check."""
    Y = np.arange(1.0, 9.0)[None, None, :].repeat(3, 0).repeat(3, 1)

    def short(line, r):
        P = np.cumsum(line, axis=-1)
        out = np.empty_like(P)
        for i in range(8):
            hi = min(i + r - 1, 7)  # the planted mistake
            lo = i - r - 1
            out[..., i] = P[..., max(hi, 0)] - (P[..., lo] if lo >= 0
                                                 else 0.0)
        return out

    pad = np.zeros((8, 8, 8))
    pad[:3, :3, :8] = Y
    got = short(pad, 2)[:3, :3, :8]
    ref = np.stack([[[Y[x, y, max(0, z - 2):z + 3].sum() for z in range(8)]
                     for y in range(3)] for x in range(3)])
    assert np.abs(got - ref).max() > 1.0
