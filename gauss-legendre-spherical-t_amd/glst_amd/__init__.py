"""Python mirror of the reference's ``glst_force<CT>`` object over the C ABI of
``include/glst_hip.h`` (``lib/libglst_hip.so``).

The class :class:`GLSTForce` keeps the reference's method names and meaning
(include/glst_force.hcu:24-60): construct from (natom, tol, box, rcut) or
(natom, tol, box, ncell), then ``assign_atoms`` -> ``calc_sf`` ->
``sum_rmt_sf`` -> ``calc_lr_ef`` -> ``calc_sr_ef`` -> ``comm_ef`` (or
``calc_ener_force``), and ``get_ef`` for energies and forces in original atom
order.  Outputs accumulate (+=) between steps exactly as in the reference;
``zero_ef`` clears them (the driver's ``fx()[dev].set(0.0)``).

Device arrays are torch tensors on ``cuda:0`` (float64, contiguous); torch is
only the allocator and stream provider.  Every compute call goes to the HIP
library; there is no fallback path: if the library is missing, importing this
module raises.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, byref, c_char_p, c_double, c_int,
                    c_size_t, c_uint32, c_uint64, c_void_p)

_HERE = os.path.dirname(os.path.abspath(__file__))
# GLST_HIP_LIB may point at another build of the same library.
LIB_PATH = os.environ.get(
    "GLST_HIP_LIB", os.path.join(os.path.dirname(_HERE), "lib", "libglst_hip.so"))

CT_DOUBLE = 0
CT_FLOAT = 1

_STATUS = {
    0: "GLST_OK",
    1: "GLST_ERR_INVALID",
    2: "GLST_ERR_HIP",
    3: "GLST_ERR_NUMERICS",
    4: "GLST_ERR_ALLOC",
    5: "GLST_ERR_IO",
    6: "GLST_ERR_STATE",
}


class GLSTError(RuntimeError):
    """A non-zero glst_status_t; ``status`` holds the code."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{_STATUS.get(status, status)}: {message}")
        self.status = status


class Info(Structure):
    _fields_ = [
        ("natom", c_uint32),
        ("ncell", c_uint32 * 3),
        ("ncell_total", c_uint32),
        ("cell_dim", c_double * 3),
        ("ngroup", c_uint32),
        ("num_nodes", c_uint32),
        ("has_long_range", c_uint32),
        ("max_atoms_cell", c_uint32),
        ("ct", c_int),
        ("spherical_rule", c_int),
        ("device_bytes", c_size_t),
    ]


class GroupInfo(Structure):
    _fields_ = [
        ("width", c_uint32),
        ("r_in", c_uint32),
        ("r_out", c_uint32),
        ("rmin", c_double),
        ("rmax", c_double),
        ("alpha", c_double),
        ("zcut", c_double),
        ("gl_order", c_uint32),
        ("sph_degree", c_uint32),
        ("num_dirs", c_uint32),
        ("first_node", c_uint32),
        ("num_nodes", c_uint32),
    ]


def _load():
    # PyTorch ships its own HIP runtime and requests it by an unversioned
    # name; loaded after ours it would be a second runtime in the process
    # (and the second one finds no device).  Import torch first so the
    # library binds to the runtime torch loaded (same soname).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"GLST HIP library not built: {LIB_PATH} is missing "
            "(run `make -C gauss-legendre-spherical-t_amd`)")
    lib = ctypes.CDLL(LIB_PATH)
    P = c_void_p
    d = c_double
    u = c_uint32
    sig = {
        "glst_create": [c_int, u, d, d, d, d, d, P, POINTER(P)],
        "glst_create_ncell": [c_int, u, d, d, d, d, u, u, u, P, POINTER(P)],
        "glst_destroy": [P],
        "glst_set_stream": [P, P],
        "glst_assign_atoms": [P, P, P, P, P],
        "glst_calc_sf": [P],
        "glst_sum_rmt_sf": [P],
        "glst_calc_lr_ef": [P],
        "glst_calc_sr_ef": [P],
        "glst_comm_ef": [P],
        "glst_calc_ener_force": [P, P, P, P, P],
        "glst_calc_ener_force_async": [P, P, P, P, P],
        "glst_get_ef": [P, P, P, P, P],
        "glst_sorted_ef_ptrs": [P, POINTER(P), POINTER(P), POINTER(P), POINTER(P)],
        "glst_zero_ef": [P],
        "glst_sorted_index_ptr": [P, POINTER(P)],
        "glst_get_info": [P, POINTER(Info)],
        "glst_get_group_info": [P, u, POINTER(GroupInfo)],
        "glst_get_nodes": [P, P, P, P, P, P],
        "glst_get_cells": [P, P, P, P, P],
        "glst_get_sf": [P, P, P],
        "glst_get_rmt": [P, P, P],
        "glst_set_timing": [P, c_int],
        "glst_set_option": [P, c_char_p, ctypes.c_int64],
        "glst_get_option": [P, c_char_p, POINTER(ctypes.c_int64)],
        "glst_autotune": [P, P, P, P, P, u, POINTER(d)],
        "glst_kernel_time": [P, c_char_p, POINTER(d), POINTER(u)],
        "glst_reset_timing": [P],
        "glst_coulomb_direct": [P, P, P, P, u, P, P, P, P, P],
        "glst_erfc_inv": [d, d, POINTER(d)],
        "glst_alpha_groups": [u, u, u, d, d, d, d, u, POINTER(u), P, P, P, P,
                              P, P],
        "glst_gauss_legendre": [u, P, P, POINTER(u)],
        "glst_select_gl_order": [d, d, d, d, POINTER(u)],
        "glst_spherical_rule": [u, P, P, POINTER(u)],
        "glst_set_tdesign_dir": [c_char_p],
        "glst_read_tdesign": [c_char_p, u, P, POINTER(u)],
        "glst_select_spherical_degree": [u, d, d, d, d, POINTER(u),
                                         POINTER(u)],
        "glst_grid_dims": [c_int, d, d, d, d, POINTER(u), POINTER(u),
                           POINTER(u)],
        "glst_make_water_box": [u, u, d, c_uint64, P, P, P, P],
        # batched engine (include/glst_batch.h)
        "glst_batch_create": [c_int, u, P, P, P, d, P, POINTER(P)],
        "glst_batch_destroy": [P],
        "glst_batch_device_bytes": [c_int, u, P, P, P, d, POINTER(c_size_t)],
        "glst_batch_num_systems": [P, POINTER(u)],
        "glst_batch_atom_offset": [P, u, POINTER(u)],
        "glst_batch_get_info": [P, u, POINTER(Info)],
        "glst_batch_calc_ener_force": [P, P, P, P, P],
        "glst_batch_calc_ener_force_async": [P, P, P, P, P],
        "glst_batch_get_ef": [P, P, P, P, P],
        "glst_batch_zero_ef": [P],
        "glst_batch_set_option": [P, c_char_p, ctypes.c_int64],
        "glst_batch_autotune": [P, P, P, P, P, u, POINTER(d)],
        "glst_batch_set_stream": [P, P],
        # serving loop (include/glst_serve.h)
        "glst_server_create": [c_int, d, c_size_t, u, P, POINTER(P)],
        "glst_server_destroy": [P],
        "glst_server_submit": [P, u, P, d, P, P, P, P, P, P, P, P,
                               POINTER(c_uint64)],
        "glst_server_step": [P, POINTER(u)],
        "glst_server_poll": [P, POINTER(u)],
        "glst_server_completed": [P, P, u, POINTER(u)],
        "glst_server_pending": [P, POINTER(u)],
        "glst_server_get_stats": [P, P],
        "glst_server_set_option": [P, c_char_p, ctypes.c_int64],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = c_int
    lib.glst_last_error.argtypes = [P]
    lib.glst_last_error.restype = c_char_p
    lib.glst_global_last_error.argtypes = []
    lib.glst_global_last_error.restype = c_char_p
    lib.glst_version.argtypes = []
    lib.glst_version.restype = c_char_p
    lib.glst_batch_last_error.argtypes = [P]
    lib.glst_batch_last_error.restype = c_char_p
    lib.glst_server_last_error.argtypes = [P]
    lib.glst_server_last_error.restype = c_char_p
    return lib


_lib = _load()


def library():
    """The loaded ctypes CDLL (for tests that check exported symbols)."""
    return _lib


def version() -> str:
    return _lib.glst_version().decode()


def _check(status: int, handle=None):
    if status != 0:
        msg = (_lib.glst_last_error(handle) if handle else
               _lib.glst_global_last_error())
        raise GLSTError(status, (msg or b"").decode())


def _ptr(a) -> int:
    """Address of a contiguous numpy array or torch tensor (None -> NULL)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return a.data_ptr()
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


# ---- host numerics ---------------------------------------------------------

def erfc_inv(y: float, tol: float) -> float:
    out = c_double()
    _check(_lib.glst_erfc_inv(y, tol, byref(out)))
    return out.value


def alpha_groups(nx, ny, nz, cdx, cdy, cdz, tol):
    import numpy as np
    n = c_uint32()
    _check(_lib.glst_alpha_groups(nx, ny, nz, cdx, cdy, cdz, tol, 0, byref(n),
                                  None, None, None, None, None, None))
    g = n.value
    width = np.zeros(g, np.uint32)
    rmax = np.zeros(g)
    alpha = np.zeros(g)
    zcut = np.zeros(g)
    rin = np.zeros(g, np.uint32)
    rout = np.zeros(g, np.uint32)
    _check(_lib.glst_alpha_groups(nx, ny, nz, cdx, cdy, cdz, tol, g, byref(n),
                                  _ptr(width), _ptr(rmax), _ptr(alpha),
                                  _ptr(zcut), _ptr(rin), _ptr(rout)))
    return dict(width=width, rmax=rmax, alpha=alpha, zcut=zcut, r_in=rin,
                r_out=rout)


def gauss_legendre(n: int):
    import numpy as np
    x = np.zeros(n // 2 + 1)
    w = np.zeros(n // 2 + 1)
    c = c_uint32()
    _check(_lib.glst_gauss_legendre(n, _ptr(x), _ptr(w), byref(c)))
    return x[: c.value].copy(), w[: c.value].copy()


def select_gl_order(tol, dist, alpha, zcut) -> int:
    o = c_uint32()
    _check(_lib.glst_select_gl_order(tol, dist, alpha, zcut, byref(o)))
    return o.value


def spherical_rule(m: int):
    import numpy as np
    c = c_uint32()
    _check(_lib.glst_spherical_rule(m, None, None, byref(c)))
    xyz = np.zeros((c.value, 3))
    w = np.zeros(c.value)
    _check(_lib.glst_spherical_rule(m, _ptr(xyz), _ptr(w), byref(c)))
    return xyz, w


def read_tdesign(directory: str, t: int):
    """The first N/2 points of the design of degree t in `directory`
    (cubature<CT>::get_tdesign_props + read_tdesign), shape (N/2, 3)."""
    import numpy as np
    c = c_uint32()
    _check(_lib.glst_read_tdesign(directory.encode(), t, None, byref(c)))
    xyz = np.zeros((c.value, 3))
    _check(_lib.glst_read_tdesign(directory.encode(), t, _ptr(xyz), byref(c)))
    return xyz


def select_spherical_degree(gl_order, tol, dist, alpha, zcut):
    """(m, directions) chosen by get_spherical_tdesign's three-axis test for
    the Gauss-Legendre rule of order gl_order, from the current design
    directory or the built-in product rule."""
    m, c = c_uint32(), c_uint32()
    _check(_lib.glst_select_spherical_degree(gl_order, tol, dist, alpha, zcut,
                                             byref(m), byref(c)))
    return m.value, c.value


def set_tdesign_dir(path: str | None):
    _check(_lib.glst_set_tdesign_dir(path.encode() if path else None))


def grid_dims(ct, lx, ly, lz, rcut):
    n = [c_uint32(), c_uint32(), c_uint32()]
    _check(_lib.glst_grid_dims(ct, lx, ly, lz, rcut, byref(n[0]), byref(n[1]),
                               byref(n[2])))
    return tuple(v.value for v in n)


def make_water_box(nwater: int, L: float, seed: int = 1, nion: int = 0):
    """Host float64 arrays (rx, ry, rz, qc) of 3*nwater + 2*nion charges."""
    import numpy as np
    n = 3 * nwater + 2 * nion
    arrs = [np.zeros(n) for _ in range(4)]
    _check(_lib.glst_make_water_box(nwater, nion, L, seed,
                                    *[_ptr(a) for a in arrs]))
    return tuple(arrs)


def coulomb_direct(rx, ry, rz, qc):
    """Direct FP64 all-pairs sum on the GPU; returns (fx, fy, fz, en)."""
    import torch
    n = rx.numel()
    out = [torch.empty(n, dtype=torch.float64, device=rx.device)
           for _ in range(4)]
    stream = torch.cuda.current_stream(rx.device).cuda_stream
    _check(_lib.glst_coulomb_direct(_ptr(rx), _ptr(ry), _ptr(rz), _ptr(qc), n,
                                    *[_ptr(o) for o in out], c_void_p(stream)))
    return tuple(out)


# ---- the force object -------------------------------------------------------

class GLSTForce:
    """MI355X GLST force object (reference: ``glst_force<CT>``).

    ``ct`` is ``"double"`` (plain FP64 path) or ``"float"`` (mixed path).
    Pass either ``rcut`` or ``ncell=(nx, ny, nz)``.
    """

    def __init__(self, natom: int, tol: float, box, rcut: float | None = None,
                 ncell=None, ct: str = "double", stream=None):
        import torch
        if isinstance(box, (int, float)):
            box = (float(box),) * 3
        self.natom = int(natom)
        self.ct = CT_FLOAT if ct in ("float", CT_FLOAT) else CT_DOUBLE
        if stream is None and torch.cuda.is_available():
            stream = torch.cuda.current_stream().cuda_stream
        self._stream = stream
        h = c_void_p()
        if ncell is not None:
            _check(_lib.glst_create_ncell(self.ct, self.natom, tol, box[0],
                                          box[1], box[2], ncell[0], ncell[1],
                                          ncell[2], c_void_p(stream),
                                          byref(h)))
        else:
            if rcut is None:
                raise ValueError("pass rcut or ncell")
            _check(_lib.glst_create(self.ct, self.natom, tol, box[0], box[1],
                                    box[2], rcut, c_void_p(stream), byref(h)))
        self._h = h

    # lifetime
    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.glst_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, status):
        _check(status, self._h)

    # the step (src/glst_force.cu:1604-1614)
    def assign_atoms(self, rx, ry, rz, qc):
        self._c(_lib.glst_assign_atoms(self._h, _ptr(rx), _ptr(ry), _ptr(rz),
                                       _ptr(qc)))

    def calc_sf(self):
        self._c(_lib.glst_calc_sf(self._h))

    def sum_rmt_sf(self):
        self._c(_lib.glst_sum_rmt_sf(self._h))

    def calc_lr_ef(self):
        self._c(_lib.glst_calc_lr_ef(self._h))

    def calc_sr_ef(self):
        self._c(_lib.glst_calc_sr_ef(self._h))

    def comm_ef(self):
        self._c(_lib.glst_comm_ef(self._h))

    def calc_ener_force(self, rx, ry, rz, qc):
        self._c(_lib.glst_calc_ener_force(self._h, _ptr(rx), _ptr(ry),
                                          _ptr(rz), _ptr(qc)))

    def calc_ener_force_async(self, rx, ry, rz, qc):
        self._c(_lib.glst_calc_ener_force_async(self._h, _ptr(rx), _ptr(ry),
                                                _ptr(rz), _ptr(qc)))

    def zero_ef(self):
        self._c(_lib.glst_zero_ef(self._h))

    # results
    def get_ef(self, out=None):
        """(fx, fy, fz, en) in original atom order as float64 cuda tensors."""
        import torch
        if out is None:
            out = tuple(torch.empty(self.natom, dtype=torch.float64,
                                    device="cuda") for _ in range(4))
        self._c(_lib.glst_get_ef(self._h, *[_ptr(o) for o in out]))
        return out

    def sorted_ef_ptrs(self):
        p = [c_void_p() for _ in range(4)]
        self._c(_lib.glst_sorted_ef_ptrs(self._h, *[byref(v) for v in p]))
        return tuple(v.value for v in p)

    # setup inspection
    def info(self) -> Info:
        inf = Info()
        self._c(_lib.glst_get_info(self._h, byref(inf)))
        return inf

    def group_info(self, g: int) -> GroupInfo:
        gi = GroupInfo()
        self._c(_lib.glst_get_group_info(self._h, g, byref(gi)))
        return gi

    def nodes(self):
        import numpy as np
        K = self.info().num_nodes
        kx, ky, kz, w = (np.zeros(K) for _ in range(4))
        grp = np.zeros(K, np.uint32)
        self._c(_lib.glst_get_nodes(self._h, _ptr(kx), _ptr(ky), _ptr(kz),
                                    _ptr(w), _ptr(grp)))
        return kx, ky, kz, w, grp

    def cells(self):
        import numpy as np
        inf = self.info()
        ac = np.zeros(self.natom, np.uint32)
        si = np.zeros(self.natom, np.uint32)
        cp = np.zeros(inf.ncell_total, np.uint32)
        cc = np.zeros(inf.ncell_total, np.uint32)
        self._c(_lib.glst_get_cells(self._h, _ptr(ac), _ptr(si), _ptr(cp),
                                    _ptr(cc)))
        return ac, si, cp, cc

    def _field(self, fn):
        import numpy as np
        inf = self.info()
        re = np.zeros((inf.ncell_total, inf.num_nodes))
        im = np.zeros((inf.ncell_total, inf.num_nodes))
        self._c(fn(self._h, _ptr(re), _ptr(im)))
        return re, im

    def sf(self):
        return self._field(_lib.glst_get_sf)

    def rmt(self):
        return self._field(_lib.glst_get_rmt)

    def set_option(self, name: str, value: int):
        self._c(_lib.glst_set_option(self._h, name.encode(), int(value)))

    def get_option(self, name: str) -> int:
        v = ctypes.c_int64()
        self._c(_lib.glst_get_option(self._h, name.encode(), byref(v)))
        return v.value

    def autotune(self, rx, ry, rz, qc, reps: int = 3) -> float:
        """Keep the fastest schedule for these inputs; returns ms/step.
        CT = double handles of up to 300,000 atoms only (others raise
        GLSTError; set their schedule with set_option)."""
        t = c_double()
        self._c(_lib.glst_autotune(self._h, _ptr(rx), _ptr(ry), _ptr(rz),
                                   _ptr(qc), reps, byref(t)))
        return t.value

    # timing
    def set_timing(self, on: bool):
        self._c(_lib.glst_set_timing(self._h, 1 if on else 0))

    def kernel_time(self, name: str):
        ms = c_double()
        n = c_uint32()
        self._c(_lib.glst_kernel_time(self._h, name.encode(), byref(ms),
                                      byref(n)))
        return ms.value, n.value

    def reset_timing(self):
        self._c(_lib.glst_reset_timing(self._h))


# ---- the batched engine -----------------------------------------------------

def _batch_arrays(natoms, boxes, rcuts):
    import numpy as np
    n = np.ascontiguousarray(natoms, dtype=np.uint32)
    b = np.asarray(boxes, dtype=np.float64)
    if b.ndim == 1:
        b = np.repeat(b[:, None], 3, axis=1)
    b = np.ascontiguousarray(b.reshape(len(n), 3))
    r = np.ascontiguousarray(np.broadcast_to(
        np.asarray(rcuts, dtype=np.float64), (len(n),)))
    return n, b, r


def batch_device_bytes(natoms, boxes, rcuts, tol: float,
                       ct: str = "double") -> int:
    """Device bytes a batch of these systems would hold (host-only)."""
    n, b, r = _batch_arrays(natoms, boxes, rcuts)
    out = c_size_t()
    _check(_lib.glst_batch_device_bytes(
        CT_FLOAT if ct in ("float", CT_FLOAT) else CT_DOUBLE, len(n), _ptr(n),
        _ptr(b), _ptr(r), tol, byref(out)))
    return out.value


class GLSTBatch:
    """Many independent systems, each a ``glst_force<CT>``, computed together
    (``include/glst_batch.h``).  ``natoms[s]`` atoms in a box ``boxes[s]``
    (a scalar edge or three edges) with cell edge ``rcuts[s]``; inputs and
    outputs are packed: system s owns atoms ``offset(s) .. offset(s) +
    natoms[s] - 1``."""

    def __init__(self, natoms, boxes, rcuts, tol: float, ct: str = "double",
                 stream=None):
        import torch
        n, b, r = _batch_arrays(natoms, boxes, rcuts)
        self.natoms = [int(v) for v in n]
        self.ct = CT_FLOAT if ct in ("float", CT_FLOAT) else CT_DOUBLE
        if stream is None and torch.cuda.is_available():
            stream = torch.cuda.current_stream().cuda_stream
        h = c_void_p()
        _check(_lib.glst_batch_create(self.ct, len(n), _ptr(n), _ptr(b),
                                      _ptr(r), tol, c_void_p(stream),
                                      byref(h)))
        self._h = h
        self.natom = int(sum(self.natoms))

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.glst_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _c(self, status):
        if status != 0:
            msg = _lib.glst_batch_last_error(self._h) if self._h else b""
            if not msg:
                msg = _lib.glst_global_last_error()
            raise GLSTError(status, (msg or b"").decode())

    def num_systems(self) -> int:
        v = c_uint32()
        self._c(_lib.glst_batch_num_systems(self._h, byref(v)))
        return v.value

    def offset(self, s: int) -> int:
        v = c_uint32()
        self._c(_lib.glst_batch_atom_offset(self._h, s, byref(v)))
        return v.value

    def info(self, s: int) -> Info:
        i = Info()
        self._c(_lib.glst_batch_get_info(self._h, s, byref(i)))
        return i

    def calc_ener_force(self, rx, ry, rz, qc):
        self._c(_lib.glst_batch_calc_ener_force(self._h, _ptr(rx), _ptr(ry),
                                                _ptr(rz), _ptr(qc)))

    def calc_ener_force_async(self, rx, ry, rz, qc):
        self._c(_lib.glst_batch_calc_ener_force_async(
            self._h, _ptr(rx), _ptr(ry), _ptr(rz), _ptr(qc)))

    def get_ef(self):
        """(fx, fy, fz, en) of all systems, packed, original order."""
        import torch
        out = [torch.empty(self.natom, dtype=torch.float64, device="cuda")
               for _ in range(4)]
        self._c(_lib.glst_batch_get_ef(self._h, *[_ptr(o) for o in out]))
        return tuple(out)

    def zero_ef(self):
        self._c(_lib.glst_batch_zero_ef(self._h))

    def set_option(self, name: str, value: int):
        self._c(_lib.glst_batch_set_option(self._h, name.encode(), int(value)))

    def autotune(self, rx, ry, rz, qc, reps: int = 3) -> float:
        """Keep the fastest schedule for these inputs; returns ms/step."""
        t = c_double()
        self._c(_lib.glst_batch_autotune(self._h, _ptr(rx), _ptr(ry), _ptr(rz),
                                         _ptr(qc), reps, byref(t)))
        return t.value


# ---- the serving loop -------------------------------------------------------

class ServerStats(Structure):
    _fields_ = [(n, c_uint64) for n in (
        "submitted", "rejected", "completed", "batches", "reconfigures",
        "setups", "setup_hits", "atoms", "max_in_flight")]


class GLSTServer:
    """Continuous-batching server over the batched engine
    (``include/glst_serve.h``).  ``submit`` queues one system's force request
    (device arrays, borrowed until it completes; outputs overwritten) and
    returns a ticket; ``step`` computes the oldest queued requests that fit
    ``budget_bytes`` (at most ``max_batch``) as one batch."""

    def __init__(self, tol: float, budget_bytes: int, max_batch: int = 0,
                 ct: str = "double", stream=None):
        import torch
        self.ct = CT_FLOAT if ct in ("float", CT_FLOAT) else CT_DOUBLE
        if stream is None and torch.cuda.is_available():
            stream = torch.cuda.current_stream().cuda_stream
        h = c_void_p()
        _check(_lib.glst_server_create(self.ct, tol, int(budget_bytes),
                                       int(max_batch), c_void_p(stream),
                                       byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.glst_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _c(self, status):
        if status != 0:
            msg = _lib.glst_server_last_error(self._h) if self._h else b""
            if not msg:
                msg = _lib.glst_global_last_error()
            raise GLSTError(status, (msg or b"").decode())

    def submit(self, box, rcut: float, rx, ry, rz, qc, fx, fy, fz, en) -> int:
        import numpy as np
        if isinstance(box, (int, float)):
            box = (float(box),) * 3
        b = np.ascontiguousarray(box, dtype=np.float64)
        t = c_uint64()
        self._c(_lib.glst_server_submit(
            self._h, rx.numel(), _ptr(b), rcut, _ptr(rx), _ptr(ry), _ptr(rz),
            _ptr(qc), _ptr(fx), _ptr(fy), _ptr(fz), _ptr(en), byref(t)))
        return t.value

    def step(self) -> int:
        """Launch what fits, retire finished batches; wait for the oldest
        batch in flight if none had finished.  Returns requests answered."""
        n = c_uint32()
        self._c(_lib.glst_server_step(self._h, byref(n)))
        return n.value

    def poll(self) -> int:
        """As step, but never waits for the device."""
        n = c_uint32()
        self._c(_lib.glst_server_poll(self._h, byref(n)))
        return n.value

    def completed(self, cap: int = 4096):
        import numpy as np
        out = np.zeros(cap, dtype=np.uint64)
        n = c_uint32()
        self._c(_lib.glst_server_completed(self._h, _ptr(out), cap, byref(n)))
        return [int(v) for v in out[: n.value]]

    def set_option(self, name: str, value: int):
        """Server options ("fused", "graph", "max_batch"; glst_server_set_option)."""
        self._c(_lib.glst_server_set_option(self._h, name.encode(), int(value)))

    def pending(self) -> int:
        """Requests not yet answered (queued or in flight)."""
        n = c_uint32()
        self._c(_lib.glst_server_pending(self._h, byref(n)))
        return n.value

    def stats(self) -> dict:
        st = ServerStats()
        self._c(_lib.glst_server_get_stats(self._h, ctypes.addressof(st)))
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}
