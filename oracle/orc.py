"""ctypes binding of the CPU oracle (oracle/liborc.so, built from fdtd_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, where it is the checker.  The product never
imports it.  It takes the same spec dicts and ABI structs as php_fdtd_amd.
"""
import ctypes
import os

import numpy as np

from php_fdtd_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("oracle not built: make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        d = ctypes.c_double
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [P(_abi.Grid), P(_abi.Material), ctypes.c_int, P(ctypes.c_int32),
                                 ctypes.c_char_p, ctypes.c_int]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_set_sources.argtypes = [ctypes.c_void_p, ctypes.c_int, P(_abi.Source)]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_get_step.argtypes = [ctypes.c_void_p]
        L.orc_get_step.restype = ctypes.c_long
        L.orc_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, P(d)]
        L.orc_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, P(d)]
        L.orc_state.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, P(d), ctypes.c_int]
        L.orc_state.restype = ctypes.c_long
        L.orc_monitor_row.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5 + [P(d)]
        L.orc_cell_ions.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(d), P(d)]
        L.orc_diag.argtypes = [ctypes.c_void_p, P(d), P(ctypes.c_int)]
        L.orc_slice.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char, ctypes.c_int, P(ctypes.c_float)]
        L.orc_source_value.argtypes = [P(_abi.Source), d]
        L.orc_source_value.restype = d
        L.orc_pml_profile.argtypes = [ctypes.c_int, d, d, d, d, d, d, d, P(d), P(d), P(d)]
        L.orc_eps_inverse.argtypes = [P(d), P(d), P(d), P(d)]
        L.orc_inverse3.argtypes = [P(d), P(d)]
        L.orc_ade_coefs.argtypes = [P(d), P(d), d, d] + [P(d)] * 6
        _LIB = L
    return _LIB


class Oracle:
    """One FP64 CPU simulation (reference step ordering)."""

    def __init__(self, spec, medium):
        L = lib()
        self.spec = spec
        g = spec["grid"]
        self.shape = (int(g["nx"]), int(g["ny"]), int(g["nz"]))
        self.grid = _abi.make_grid(g)
        self.mats = _abi.MaterialTable(spec["materials"])
        self.medium = _abi.medium_array(medium)
        assert self.medium.shape == self.shape
        err = ctypes.create_string_buffer(256)
        self.h = L.orc_create(ctypes.byref(self.grid), self.mats.arr, self.mats.n, _abi.i32p(self.medium), err, 256)
        if not self.h:
            raise ValueError(err.value.decode())
        srcs = spec.get("sources", [])
        self._src = _abi.make_sources(srcs)
        L.orc_set_sources(self.h, len(srcs), self._src)

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    __del__ = close

    def step(self, n=1):
        lib().orc_step(self.h, n)

    @property
    def t(self):
        return lib().orc_get_step(self.h)

    def field(self, comp):
        c = _abi.COMPONENTS[comp] if isinstance(comp, str) else comp
        out = np.empty(self.shape, np.float64)
        lib().orc_get_field(self.h, c, _abi.f64p(out))
        return out

    def set_field(self, comp, arr):
        c = _abi.COMPONENTS[comp] if isinstance(comp, str) else comp
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape
        lib().orc_set_field(self.h, c, _abi.f64p(a))

    def state(self, mat, kind):
        n = lib().orc_state(self.h, mat, kind, None, 0)
        out = np.empty(max(n, 1), np.float64)
        lib().orc_state(self.h, mat, kind, _abi.f64p(out), 0)
        return out[:n]

    def set_state(self, mat, kind, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        lib().orc_state(self.h, mat, kind, _abi.f64p(a), 1)

    def monitor_row(self, x, y, z, atom=0, dir=0):
        row = np.empty(12, np.float64)
        lib().orc_monitor_row(self.h, x, y, z, atom, dir, _abi.f64p(row))
        return row

    def cell_ions(self, x, y, z):
        r = np.zeros(3 * 64, np.float64)
        v = np.zeros(3 * 64, np.float64)
        n = lib().orc_cell_ions(self.h, x, y, z, _abi.f64p(r), _abi.f64p(v))
        return r[:3 * n], v[:3 * n]

    def diag(self):
        v = np.empty(3, np.float64)
        idx = (ctypes.c_int * 5)()
        lib().orc_diag(self.h, _abi.f64p(v), idx)
        return v, list(idx)

    def slice(self, comp, axis, pos):
        nx, ny, nz = self.shape
        shp = {"z": (nx, ny), "x": (ny, nz), "y": (nx, nz)}[axis]
        out = np.empty(shp, np.float32)
        c = _abi.COMPONENTS[comp] if isinstance(comp, str) else comp
        rc = lib().orc_slice(self.h, c, axis.encode(), pos, _abi.f32p(out))
        if rc != 0:
            raise ValueError("bad slice")
        return out
