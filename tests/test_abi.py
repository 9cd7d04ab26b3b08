"""CPU checks of the C ABI library: it loads without a GPU, exports every
symbol include/phpfdtd.h declares, agrees with the ctypes struct layouts, and
rejects bad inputs before touching the device."""
import ctypes
import os
import re

import numpy as np
import pytest

import golden_cases
from php_fdtd_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "phpfdtd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(phpfdtd_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    return _abi.load_library()


def test_exports_every_declared_symbol(lib):
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(_abi.EXPORTED) <= set(names)


def test_struct_layouts(lib):
    sizes = (ctypes.c_int * 5)()
    assert lib.phpfdtd_struct_sizes(sizes, 5) == 5
    assert list(sizes) == [ctypes.sizeof(t) for t in (_abi.Grid, _abi.Material, _abi.Source, _abi.Diag, _abi.KTime)]
    assert lib.phpfdtd_abi_version() == 1


def _case():
    c = golden_cases.load("hbn")
    return c["spec"], c["medium"]


def test_footprint_without_gpu(lib):
    spec, med = _case()
    grid = _abi.make_grid(spec["grid"])
    mats = _abi.MaterialTable(spec["materials"])
    m = _abi.medium_array(med)
    out = ctypes.c_size_t()
    assert lib.phpfdtd_footprint_bytes(ctypes.byref(grid), mats.arr, mats.n, _abi.i32p(m), ctypes.byref(out)) == 0
    nx, ny, nz = med.shape
    nxp = (nx + 31) // 32 * 32
    assert out.value >= 9 * nxp * ny * nz * 8
    # FP32 fields take half the bytes
    grid32 = _abi.make_grid(spec["grid"], _abi.FP32)
    out32 = ctypes.c_size_t()
    lib.phpfdtd_footprint_bytes(ctypes.byref(grid32), mats.arr, mats.n, _abi.i32p(m), ctypes.byref(out32))
    assert out32.value < out.value


def test_rejects_bad_medium_before_device(lib):
    spec, med = _case()
    grid = _abi.make_grid(spec["grid"])
    mats = _abi.MaterialTable(spec["materials"])
    bad = med.copy()
    bad[1, 2, 3] = len(spec["materials"])  # outside the table (unchecked in fdtd.c:3408)
    h = ctypes.c_void_p()
    rc = lib.phpfdtd_create(ctypes.byref(grid), mats.arr, mats.n, _abi.i32p(_abi.medium_array(bad)), ctypes.byref(h))
    assert rc == _abi.EFORMAT
    assert b"outside the material table" in lib.phpfdtd_last_error()


@pytest.mark.parametrize("mutate,msg", [
    (lambda m: m.update(anglex=0.3, anglez=0.2), b"Can not rotate along X and Z"),
    (lambda m: m.update(use_ade=1, eps_re=[5.0, 0, 0, 0, 5.0, 0, 0, 0, 25.0]), b"EPS_INF is too small"),
    (lambda m: m.update(eps_re=[0.0] * 9), b"Singular dielectric tensor"),
])
def test_reference_input_errors(lib, mutate, msg):
    spec, med = _case()
    mats_spec = [dict(m) for m in spec["materials"]]
    mutate(mats_spec[2])
    grid = _abi.make_grid(spec["grid"])
    mats = _abi.MaterialTable(mats_spec)
    h = ctypes.c_void_p()
    rc = lib.phpfdtd_create(ctypes.byref(grid), mats.arr, mats.n, _abi.i32p(_abi.medium_array(med)), ctypes.byref(h))
    assert rc == _abi.EINVAL
    assert msg in lib.phpfdtd_last_error()
