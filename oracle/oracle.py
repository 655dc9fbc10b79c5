"""ctypes bindings of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_double, c_int, c_size_t, c_uint32, c_void_p

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"oracle not built: {LIB_PATH} (run `make -C oracle`)")
    lib = ctypes.CDLL(LIB_PATH)
    P, d, u = c_void_p, c_double, c_uint32
    sig = {
        "orc_erfc_inv": [d, d, POINTER(d)],
        "orc_alpha_groups": [u, u, u, d, d, d, d, u, POINTER(u), P, P, P, P, P,
                             P, P],
        "orc_gauss_legendre": [u, P, P, POINTER(u)],
        "orc_select_gl": [d, d, d, d, POINTER(u)],
        "orc_sph_rule": [u, P, P, POINTER(u)],
        "orc_select_sph": [u, d, d, d, d, POINTER(u)],
        "orc_build_nodes": [d, u, u, u, d, c_size_t, POINTER(u), P, P, P, P, P],
        "orc_grid_dims": [c_int, d, d, d, d, POINTER(u), POINTER(u), POINTER(u)],
        "orc_bin": [c_int, P, P, P, u, u, u, u, d, P, P, P, P],
        "orc_glst": [c_int, P, P, P, P, u, d, d, d, d, d, u, c_int, c_int, P, P,
                     P, P, P, P, P, P],
        "orc_coulomb": [P, P, P, P, u, P, P, P, P],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = c_int
    lib.orc_legendre.argtypes = [d, u, POINTER(d), POINTER(d)]
    lib.orc_legendre.restype = None
    lib.orc_set_threads.argtypes = [c_int]
    lib.orc_set_threads.restype = None
    lib.orc_get_threads.argtypes = []
    lib.orc_get_threads.restype = c_int
    lib.orc_last_stage_times.argtypes = [P]
    lib.orc_last_stage_times.restype = None
    return lib


_lib = _load()


class OracleError(RuntimeError):
    pass


def _ok(st):
    if st != 0:
        raise OracleError(f"oracle status {st}")


def _p(a):
    return None if a is None else a.ctypes.data


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def set_threads(n: int):
    _lib.orc_set_threads(n)


def get_threads() -> int:
    return _lib.orc_get_threads()


def erfc_inv(y, tol=np.finfo(float).eps):
    x = c_double()
    _ok(_lib.orc_erfc_inv(y, tol, byref(x)))
    return x.value


def alpha_groups(nx, ny, nz, cell, tol):
    n = c_uint32()
    cap = 64
    w = np.zeros(cap, np.uint32)
    rmin, rmax, a, z = (np.zeros(cap) for _ in range(4))
    rin = np.zeros(cap, np.uint32)
    rout = np.zeros(cap, np.uint32)
    _ok(_lib.orc_alpha_groups(nx, ny, nz, cell, cell, cell, tol, cap, byref(n),
                              _p(w), _p(rmin), _p(rmax), _p(a), _p(z), _p(rin),
                              _p(rout)))
    g = n.value
    return dict(width=w[:g], rmin=rmin[:g], rmax=rmax[:g], alpha=a[:g],
                zcut=z[:g], r_in=rin[:g], r_out=rout[:g])


def legendre(x, n):
    p, dp = c_double(), c_double()
    _lib.orc_legendre(x, n, byref(p), byref(dp))
    return p.value, dp.value


def gauss_legendre(n):
    x = np.zeros(n // 2 + 1)
    w = np.zeros(n // 2 + 1)
    c = c_uint32()
    _ok(_lib.orc_gauss_legendre(n, _p(x), _p(w), byref(c)))
    return x[: c.value].copy(), w[: c.value].copy()


def select_gl(tol, dist, alpha, zcut):
    o = c_uint32()
    _ok(_lib.orc_select_gl(tol, dist, alpha, zcut, byref(o)))
    return o.value


def sph_rule(m):
    c = c_uint32()
    _ok(_lib.orc_sph_rule(m, None, None, byref(c)))
    xyz = np.zeros((c.value, 3))
    w = np.zeros(c.value)
    _ok(_lib.orc_sph_rule(m, _p(xyz), _p(w), byref(c)))
    return xyz, w


def select_sph(order, tol, dist, alpha, zcut):
    m = c_uint32()
    _ok(_lib.orc_select_sph(order, tol, dist, alpha, zcut, byref(m)))
    return m.value


def build_nodes(tol, nx, ny, nz, cell):
    K = c_uint32()
    _ok(_lib.orc_build_nodes(tol, nx, ny, nz, cell, 0, byref(K), None, None,
                             None, None, None))
    k = K.value
    kx, ky, kz, w = (np.zeros(k) for _ in range(4))
    g = np.zeros(k, np.uint32)
    _ok(_lib.orc_build_nodes(tol, nx, ny, nz, cell, k, byref(K), _p(kx), _p(ky),
                             _p(kz), _p(w), _p(g)))
    return kx, ky, kz, w, g


def grid_dims(ct, lx, ly, lz, rcut):
    n = [c_uint32() for _ in range(3)]
    _ok(_lib.orc_grid_dims(ct, lx, ly, lz, rcut, *[byref(v) for v in n]))
    return tuple(v.value for v in n)


def bin_atoms(ct, rx, ry, rz, ncell, rcut):
    rx, ry, rz = _f64(rx), _f64(ry), _f64(rz)
    n = rx.size
    nc = int(np.prod(ncell))
    ac = np.zeros(n, np.uint32)
    si = np.zeros(n, np.uint32)
    cp = np.zeros(nc, np.uint32)
    cc = np.zeros(nc, np.uint32)
    _ok(_lib.orc_bin(ct, _p(rx), _p(ry), _p(rz), n, ncell[0], ncell[1],
                     ncell[2], rcut, _p(ac), _p(si), _p(cp), _p(cc)))
    return ac, si, cp, cc


def glst(rx, ry, rz, qc, tol, box, rcut, ct=0, node_stride=1, far=True,
         near=True, fields=False):
    """FP64 GLST energies/forces in original order: (fx, fy, fz, en) and, with
    fields=True, also (S_re, S_im, R_re, R_im) of shape (ncell, K)."""
    rx, ry, rz, qc = _f64(rx), _f64(ry), _f64(rz), _f64(qc)
    if isinstance(box, (int, float)):
        box = (float(box),) * 3
    n = rx.size
    out = [np.zeros(n) for _ in range(4)]
    fld = [None] * 4
    if fields:
        nx, ny, nz = grid_dims(ct, *box, rcut)
        cell = float(np.float32(rcut)) if ct == 1 else rcut
        K = build_nodes(tol, nx, ny, nz, cell)[0].size
        fld = [np.zeros((nx * ny * nz, K)) for _ in range(4)]
    _ok(_lib.orc_glst(ct, _p(rx), _p(ry), _p(rz), _p(qc), n, tol, box[0],
                      box[1], box[2], rcut, node_stride, int(far), int(near),
                      *[_p(o) for o in out], *[_p(f) for f in fld]))
    return (tuple(out), tuple(fld)) if fields else tuple(out)


STAGES = ("assign_atoms", "calc_sf", "sum_rmt_sf", "calc_lr_ef", "calc_sr_ef",
          "get_ef")


def last_stage_times() -> dict:
    """Wall seconds of each stage of the last glst() call."""
    t = np.zeros(6)
    _lib.orc_last_stage_times(_p(t))
    return dict(zip(STAGES, t.tolist()))


def coulomb(rx, ry, rz, qc):
    rx, ry, rz, qc = _f64(rx), _f64(ry), _f64(rz), _f64(qc)
    n = rx.size
    out = [np.zeros(n) for _ in range(4)]
    _ok(_lib.orc_coulomb(_p(rx), _p(ry), _p(rz), _p(qc), n,
                         *[_p(o) for o in out]))
    return tuple(out)
