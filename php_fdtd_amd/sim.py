"""Python mirror of the libphpfdtd_hip C ABI.

``Simulation`` owns one phpfdtd_ctx.  It takes the same plain-dict spec the
tests and workload builders use:

    spec = {"grid": {nx, ny, nz, tsteps, dx, dy, dz, dt, eps0, mu0,
                     pml_thickness, pml_reflection, pml_order, kappa_max,
                     alpha_max, f_source},
            "materials": [{eps_re[9], eps_im[9], sigma, anglez, anglex,
                           use_ade, natom, qezv, qezm, dmm}, ...],
            "sources": [{type, x, y, z, component, amplitude, frequency,
                         t0, tau, start_time, end_time}, ...]}

and a medium array of material ids in file order (shape (nx, ny, nz)).  Every
call goes through the HIP library; there is no host fallback.
"""
import ctypes

import numpy as np

from . import _abi


class Simulation:
    def __init__(self, spec, medium, precision=_abi.FP64, lib=None):
        self.lib = lib or _abi.load_library()
        g = spec["grid"]
        self.spec = spec
        self.shape = (int(g["nx"]), int(g["ny"]), int(g["nz"]))
        self.grid = _abi.make_grid(g, precision)
        self.mats = _abi.MaterialTable(spec["materials"])
        self.medium = _abi.medium_array(medium)
        if self.medium.shape != self.shape:
            raise ValueError("medium shape %s != grid %s" % (self.medium.shape, self.shape))
        h = ctypes.c_void_p()
        _abi.check(self.lib, self.lib.phpfdtd_create(ctypes.byref(self.grid), self.mats.arr, self.mats.n,
                                                     _abi.i32p(self.medium), ctypes.byref(h)))
        self.h = h
        srcs = spec.get("sources", [])
        if srcs:
            self.set_sources(srcs)

    # -- lifecycle -----------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.phpfdtd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- configuration -------------------------------------------------------
    def set_sources(self, sources):
        self._src = _abi.make_sources(sources)
        _abi.check(self.lib, self.lib.phpfdtd_set_sources(self.h, len(sources), self._src))

    def set_monitor(self, x, y, z, atom=0, dir=0, ref_compat=True):
        _abi.check(self.lib, self.lib.phpfdtd_set_monitor(self.h, x, y, z, atom, dir, 1 if ref_compat else 0))

    # -- stepping --------------------------------------------------------------
    def run(self, n):
        _abi.check(self.lib, self.lib.phpfdtd_run(self.h, int(n)))

    def sync(self, diag=True):
        if not diag:
            _abi.check(self.lib, self.lib.phpfdtd_sync(self.h, None))
            return None
        d = _abi.Diag()
        rc = self.lib.phpfdtd_sync(self.h, ctypes.byref(d))
        if rc < 0 and rc != _abi.EDIVERGED:
            _abi.check(self.lib, rc)
        return d

    @property
    def t(self):
        return self.lib.phpfdtd_step(self.h)

    # -- sampling --------------------------------------------------------------
    def slice(self, comp, axis, pos, ref_compat=True):
        nx, ny, nz = self.shape
        shp = {"z": (nx, ny), "x": (ny, nz), "y": (nx, nz)}.get(axis, (1,))  # bad axes: library says EINVAL
        out = np.empty(shp, np.float32)
        c = _abi.COMPONENTS[comp] if isinstance(comp, str) else comp
        _abi.check(self.lib, self.lib.phpfdtd_read_slice(self.h, c, axis.encode(), pos, 1 if ref_compat else 0,
                                                         _abi.f32p(out)))
        return out

    def _plane_shape(self, axis):
        nx, ny, nz = self.shape
        return {"z": (nx, ny), "x": (ny, nz), "y": (nx, nz)}.get(axis, (1,))

    def set_dft(self, comp, axis, pos, interval, nsamples, f_hz, ref_compat=True):
        """Streamed single-bin DFT of one plane (phpfdtd_set_dft); returns the bin index."""
        c = _abi.COMPONENTS[comp] if isinstance(comp, str) else comp
        k = ctypes.c_int()
        _abi.check(self.lib, self.lib.phpfdtd_set_dft(self.h, c, axis.encode(), pos, 1 if ref_compat else 0,
                                                      int(interval), int(nsamples), float(f_hz), ctypes.byref(k)))
        self._dft_shape = self._plane_shape(axis)
        return k.value

    def dft(self):
        """(X, samples accumulated): X complex in the slice layout."""
        re = np.empty(self._dft_shape, np.float64)
        im = np.empty(self._dft_shape, np.float64)
        n = ctypes.c_int()
        _abi.check(self.lib, self.lib.phpfdtd_read_dft(self.h, _abi.f64p(re), _abi.f64p(im), ctypes.byref(n)))
        return re + 1j * im, n.value

    def monitor(self, max_rows=1 << 20):
        out = np.empty((max_rows, 12), np.float64)
        n = ctypes.c_int()
        _abi.check(self.lib, self.lib.phpfdtd_read_monitor(self.h, _abi.f64p(out), max_rows, ctypes.byref(n)))
        return out[:n.value].copy()

    def point(self, x, y, z):
        out = np.empty(9, np.float64)
        _abi.check(self.lib, self.lib.phpfdtd_read_point(self.h, x, y, z, _abi.f64p(out)))
        return out

    def ions(self, x, y, z):
        r = np.zeros(96, np.float64)
        v = np.zeros(96, np.float64)
        n = _abi.check(self.lib, self.lib.phpfdtd_read_ions(self.h, x, y, z, _abi.f64p(r), _abi.f64p(v)))
        return r[:3 * n], v[:3 * n]

    def field(self, comp):
        c = _abi.COMPONENTS[comp] if isinstance(comp, str) else comp
        out = np.empty(self.shape, np.float64)
        _abi.check(self.lib, self.lib.phpfdtd_read_field(self.h, c, _abi.f64p(out)))
        return out

    def set_field(self, comp, arr):
        c = _abi.COMPONENTS[comp] if isinstance(comp, str) else comp
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError("shape")
        _abi.check(self.lib, self.lib.phpfdtd_write_field(self.h, c, _abi.f64p(a)))

    def count_cells(self, mat):
        n = ctypes.c_int64()
        _abi.check(self.lib, self.lib.phpfdtd_count_cells(self.h, mat, ctypes.byref(n)))
        return n.value

    def state(self, mat, kind):
        n = self.count_cells(mat)
        m = self.spec["materials"][mat]
        width = 3 * int(m.get("natom", 0)) if kind < 2 else 3
        out = np.empty(max(1, n * width), np.float64)
        _abi.check(self.lib, self.lib.phpfdtd_read_state(self.h, mat, kind, _abi.f64p(out)))
        return out[:n * width]

    def set_state(self, mat, kind, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _abi.check(self.lib, self.lib.phpfdtd_write_state(self.h, mat, kind, _abi.f64p(a)))

    # -- measurement -----------------------------------------------------------
    def profile(self, nsteps):
        out = (_abi.KTime * 64)()
        n = ctypes.c_int()
        _abi.check(self.lib, self.lib.phpfdtd_profile(self.h, int(nsteps), out, 64, ctypes.byref(n)))
        return [dict(name=out[q].name.decode(), ms=out[q].avg_ms, bytes=out[q].bytes, flops=out[q].flops)
                for q in range(n.value)]

    def time_steps(self, nsteps):
        ms = ctypes.c_double()
        _abi.check(self.lib, self.lib.phpfdtd_time_steps(self.h, int(nsteps), ctypes.byref(ms)))
        return ms.value


def footprint_bytes(spec, medium, precision=_abi.FP64, lib=None):
    lib = lib or _abi.load_library()
    grid = _abi.make_grid(spec["grid"], precision)
    mats = _abi.MaterialTable(spec["materials"])
    med = _abi.medium_array(medium)
    out = ctypes.c_size_t()
    _abi.check(lib, lib.phpfdtd_footprint_bytes(ctypes.byref(grid), mats.arr, mats.n, _abi.i32p(med),
                                                ctypes.byref(out)))
    return out.value


class Batch:
    """A phpfdtd_batch: member Simulations stepped together, one launch per
    kernel kind over all members' tiles (see include/phpfdtd.h)."""

    def __init__(self, precision=_abi.FP64, lib=None):
        self.lib = lib or _abi.load_library()
        h = ctypes.c_void_p()
        _abi.check(self.lib, self.lib.phpfdtd_batch_create(int(precision), ctypes.byref(h)))
        self.h = h
        self.members = []

    def add(self, sim):
        _abi.check(self.lib, self.lib.phpfdtd_batch_add(self.h, sim.h))
        self.members.append(sim)

    def remove(self, sim):
        _abi.check(self.lib, self.lib.phpfdtd_batch_remove(self.h, sim.h))
        self.members.remove(sim)

    def run(self, n):
        _abi.check(self.lib, self.lib.phpfdtd_batch_run(self.h, int(n)))

    def wait(self, keep=0):
        _abi.check(self.lib, self.lib.phpfdtd_batch_wait(self.h, int(keep)))

    def sync(self):
        _abi.check(self.lib, self.lib.phpfdtd_batch_sync(self.h))

    def __len__(self):
        return self.lib.phpfdtd_batch_size(self.h)

    def close(self):
        if getattr(self, "h", None):
            self.lib.phpfdtd_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
