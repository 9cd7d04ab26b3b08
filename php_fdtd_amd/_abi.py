"""ctypes mirror of include/phpfdtd.h and the loader of libphpfdtd_hip.so.

The structures here are byte-for-byte the C structs of the header; tests check
their sizes against the library.  The library is loaded from the package
directory (built in-tree by ``php_fdtd_amd/csrc/Makefile``); there is no other
implementation to fall back to, so a missing library raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphpfdtd_hip.so")

OK = 0
EINVAL = -1
ENOMEM_HBM = -2
EHIP = -3
EDIVERGED = -4
EFORMAT = -5
ESTATE = -6
ENOMEM_HOST = -7

FP64 = 0
FP32 = 1       # precision form, lattice r, v stored as compensated FP32
FP32_RV64 = 2  # precision form, lattice r, v stored in FP64

COMPONENTS = {"Ex": 0, "Ey": 1, "Ez": 2, "Dx": 3, "Dy": 4, "Dz": 5, "Hx": 6, "Hy": 7, "Hz": 8, "E": 9}


class Grid(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("tsteps", ctypes.c_int32),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("dt", ctypes.c_double), ("eps0", ctypes.c_double), ("mu0", ctypes.c_double),
        ("pml_thickness", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("pml_reflection", ctypes.c_double), ("pml_order", ctypes.c_double),
        ("kappa_max", ctypes.c_double), ("alpha_max", ctypes.c_double),
        ("f_source", ctypes.c_double),
        ("flags", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class Material(ctypes.Structure):
    _fields_ = [
        ("eps_re", ctypes.c_double * 9), ("eps_im", ctypes.c_double * 9),
        ("sigma", ctypes.c_double), ("anglez", ctypes.c_double), ("anglex", ctypes.c_double),
        ("use_ade", ctypes.c_int32), ("natom", ctypes.c_int32),
        ("qezv", ctypes.POINTER(ctypes.c_double)), ("qezm", ctypes.POINTER(ctypes.c_double)),
        ("dmm", ctypes.POINTER(ctypes.c_double)),
    ]


class Source(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_int32), ("x", ctypes.c_int32), ("y", ctypes.c_int32), ("z", ctypes.c_int32),
        ("component", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("amplitude", ctypes.c_double), ("frequency", ctypes.c_double), ("t0", ctypes.c_double),
        ("tau", ctypes.c_double), ("start_time", ctypes.c_double), ("end_time", ctypes.c_double),
    ]


class Diag(ctypes.Structure):
    _fields_ = [
        ("step", ctypes.c_int64), ("max_e", ctypes.c_double), ("max_ex", ctypes.c_double),
        ("max_r", ctypes.c_double), ("r_i", ctypes.c_int32), ("r_j", ctypes.c_int32),
        ("r_k", ctypes.c_int32), ("r_atom", ctypes.c_int32), ("r_dir", ctypes.c_int32),
        ("diverged", ctypes.c_int32),
    ]


class KTime(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 48), ("avg_ms", ctypes.c_double), ("bytes", ctypes.c_double),
        ("flops", ctypes.c_double), ("launches", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


# --------------------------------------------------------------------------
# spec (plain dict) -> structs
# --------------------------------------------------------------------------
class MaterialTable:
    """Keeps the ctypes array and the numpy buffers it points into alive."""

    def __init__(self, materials):
        self.n = len(materials)
        self.arr = (Material * max(1, self.n))()
        self._keep = []
        for i, m in enumerate(materials):
            s = self.arr[i]
            for e in range(9):
                s.eps_re[e] = float(m["eps_re"][e])
                s.eps_im[e] = float(m["eps_im"][e])
            s.sigma = float(m.get("sigma", 0.0))
            s.anglez = float(m.get("anglez", 0.0))
            s.anglex = float(m.get("anglex", 0.0))
            s.use_ade = int(m.get("use_ade", 0))
            s.natom = int(m.get("natom", 0))
            if s.natom > 0:
                for name in ("qezv", "qezm", "dmm"):
                    buf = np.ascontiguousarray(np.asarray(m[name], dtype=np.float64))
                    self._keep.append(buf)
                    setattr(s, name, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))


def make_grid(g, precision=FP64):
    s = Grid()
    s.nx, s.ny, s.nz = int(g["nx"]), int(g["ny"]), int(g["nz"])
    s.tsteps = int(g["tsteps"])
    s.dx, s.dy, s.dz = float(g["dx"]), float(g["dy"]), float(g["dz"])
    s.dt = float(g["dt"])
    s.eps0 = float(g.get("eps0", 8.854e-12))
    s.mu0 = float(g.get("mu0", 3.141592653589793 * 4e-7))
    s.pml_thickness = int(g["pml_thickness"])
    s.precision = int(precision)
    s.pml_reflection = float(g.get("pml_reflection", 1e-3))
    s.pml_order = float(g.get("pml_order", 3.0))
    s.kappa_max = float(g.get("kappa_max", 7.0))
    s.alpha_max = float(g.get("alpha_max", 0.05))
    s.f_source = float(g["f_source"])
    return s


def make_sources(sources):
    arr = (Source * max(1, len(sources)))()
    for i, src in enumerate(sources):
        s = arr[i]
        s.type = int(src["type"])
        s.x, s.y, s.z = int(src["x"]), int(src["y"]), int(src["z"])
        comp = src["component"]
        s.component = ord(comp) if isinstance(comp, str) else int(comp)
        for k in ("amplitude", "frequency", "t0", "tau", "start_time", "end_time"):
            setattr(s, k, float(src[k]))
    return arr


def medium_array(medium):
    return np.ascontiguousarray(medium, dtype=np.int32)


def i32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def f64p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def f32p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


_LIB = None


def load_library(path=None):
    """Load libphpfdtd_hip.so and declare its signatures.  Raises OSError when
    the library has not been built: the HIP path is the only implementation."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise OSError("libphpfdtd_hip.so is not built (%s); run `python -c 'import __graft_entry__ as g; "
                      "g.build()'` or `make -C php_fdtd_amd/csrc`" % p)
    lib = ctypes.CDLL(p)
    P = ctypes.POINTER
    c_ctx = ctypes.c_void_p
    lib.phpfdtd_create.argtypes = [P(Grid), P(Material), ctypes.c_int, P(ctypes.c_int32), P(c_ctx)]
    lib.phpfdtd_set_sources.argtypes = [c_ctx, ctypes.c_int, P(Source)]
    lib.phpfdtd_set_monitor.argtypes = [c_ctx] + [ctypes.c_int] * 6
    lib.phpfdtd_run.argtypes = [c_ctx, ctypes.c_int]
    lib.phpfdtd_sync.argtypes = [c_ctx, P(Diag)]
    lib.phpfdtd_read_slice.argtypes = [c_ctx, ctypes.c_int, ctypes.c_char, ctypes.c_int, ctypes.c_int,
                                       P(ctypes.c_float)]
    lib.phpfdtd_read_monitor.argtypes = [c_ctx, P(ctypes.c_double), ctypes.c_int, P(ctypes.c_int)]
    lib.phpfdtd_read_point.argtypes = [c_ctx, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]
    lib.phpfdtd_read_ions.argtypes = [c_ctx, ctypes.c_int, ctypes.c_int, ctypes.c_int, P(ctypes.c_double),
                                      P(ctypes.c_double)]
    lib.phpfdtd_set_dft.argtypes = [c_ctx, ctypes.c_int, ctypes.c_char, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, ctypes.c_double, P(ctypes.c_int)]
    lib.phpfdtd_set_dft_noalloc.argtypes = lib.phpfdtd_set_dft.argtypes
    lib.phpfdtd_read_dft.argtypes = [c_ctx, P(ctypes.c_double), P(ctypes.c_double), P(ctypes.c_int)]
    c_b = ctypes.c_void_p
    lib.phpfdtd_batch_create.argtypes = [ctypes.c_int, P(c_b)]
    lib.phpfdtd_batch_add.argtypes = [c_b, c_ctx]
    lib.phpfdtd_batch_remove.argtypes = [c_b, c_ctx]
    lib.phpfdtd_batch_run.argtypes = [c_b, ctypes.c_int]
    lib.phpfdtd_batch_sync.argtypes = [c_b]
    lib.phpfdtd_batch_wait.argtypes = [c_b, ctypes.c_int]
    lib.phpfdtd_batch_size.argtypes = [c_b]
    lib.phpfdtd_batch_destroy.argtypes = [c_b]
    lib.phpfdtd_batch_destroy.restype = None
    lib.phpfdtd_read_field.argtypes = [c_ctx, ctypes.c_int, P(ctypes.c_double)]
    lib.phpfdtd_write_field.argtypes = [c_ctx, ctypes.c_int, P(ctypes.c_double)]
    lib.phpfdtd_count_cells.argtypes = [c_ctx, ctypes.c_int, P(ctypes.c_int64)]
    lib.phpfdtd_read_state.argtypes = [c_ctx, ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]
    lib.phpfdtd_write_state.argtypes = [c_ctx, ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]
    lib.phpfdtd_footprint_bytes.argtypes = [P(Grid), P(Material), ctypes.c_int, P(ctypes.c_int32),
                                            P(ctypes.c_size_t)]
    lib.phpfdtd_profile.argtypes = [c_ctx, ctypes.c_int, P(KTime), ctypes.c_int, P(ctypes.c_int)]
    lib.phpfdtd_time_steps.argtypes = [c_ctx, ctypes.c_int, P(ctypes.c_double)]
    lib.phpfdtd_source_value.argtypes = [P(Source), ctypes.c_double]
    lib.phpfdtd_source_value.restype = ctypes.c_double
    lib.phpfdtd_step.argtypes = [c_ctx]
    lib.phpfdtd_step.restype = ctypes.c_int64
    lib.phpfdtd_last_error.restype = ctypes.c_char_p
    lib.phpfdtd_abi_version.restype = ctypes.c_int
    lib.phpfdtd_destroy.argtypes = [c_ctx]
    lib.phpfdtd_destroy.restype = None
    lib.phpfdtd_struct_sizes.argtypes = [P(ctypes.c_int), ctypes.c_int]
    if path is None:
        _LIB = lib
    return lib


EXPORTED = [
    "phpfdtd_create", "phpfdtd_set_sources", "phpfdtd_set_monitor", "phpfdtd_run", "phpfdtd_sync",
    "phpfdtd_read_slice", "phpfdtd_read_monitor", "phpfdtd_read_point", "phpfdtd_read_ions",
    "phpfdtd_set_dft", "phpfdtd_set_dft_noalloc", "phpfdtd_read_dft", "phpfdtd_batch_create", "phpfdtd_batch_add", "phpfdtd_batch_remove",
    "phpfdtd_batch_run", "phpfdtd_batch_sync", "phpfdtd_batch_wait", "phpfdtd_batch_size", "phpfdtd_batch_destroy",
    "phpfdtd_read_field", "phpfdtd_write_field", "phpfdtd_count_cells", "phpfdtd_read_state",
    "phpfdtd_write_state", "phpfdtd_footprint_bytes", "phpfdtd_profile", "phpfdtd_time_steps",
    "phpfdtd_step", "phpfdtd_last_error", "phpfdtd_abi_version", "phpfdtd_destroy", "phpfdtd_source_value",
]


class PhpFdtdError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("phpfdtd error %d: %s" % (code, msg))
        self.code = code


def check(lib, rc):
    if rc < 0:
        raise PhpFdtdError(rc, lib.phpfdtd_last_error().decode(errors="replace"))
    return rc
