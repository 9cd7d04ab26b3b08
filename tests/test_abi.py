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


def test_footprint_by_precision(lib):
    """FP64 > FP32 with FP64 r, v > FP32 with compensated r, v; the last two
    differ by the lattice state alone: 64-slot tiles of 3N (here 12) dofs,
    r and v, 8 bytes a value against a 1 KiB row of FP32 words per 4 dofs
    plus a 1 KiB row of 4-bit compensation codes per 8 such rows, where v
    keeps the codes of a last row of 1-4 groups in r's last row (4 atoms: 3
    groups, so v has no code row)."""
    spec, med = _case()
    mats = _abi.MaterialTable(spec["materials"])
    m = _abi.medium_array(med)
    fp = {}
    for p in (_abi.FP64, _abi.FP32_RV64, _abi.FP32):
        grid = _abi.make_grid(spec["grid"], p)
        out = ctypes.c_size_t()
        assert lib.phpfdtd_footprint_bytes(ctypes.byref(grid), mats.arr, mats.n, _abi.i32p(m), ctypes.byref(out)) == 0
        fp[p] = out.value
    assert fp[_abi.FP64] > fp[_abi.FP32_RV64] > fp[_abi.FP32]
    lat = [i for i, mat in enumerate(spec["materials"]) if mat["natom"]]
    assert len(lat) == 1
    n3 = 3 * spec["materials"][lat[0]]["natom"]
    ntile = (int((med == lat[0]).sum()) + 63) // 64
    n3_pairs, ng = (n3 + 1) // 2 * 2, (n3 + 3) // 4
    nb = (ng + 7) // 8
    nbv = nb - (1 if 1 <= ng % 8 <= 4 else 0)
    rows = (ng + nb) + (ng + nbv)
    assert fp[_abi.FP32_RV64] - fp[_abi.FP32] == ntile * (2 * 64 * n3_pairs * 8 - 1024 * rows)


def test_rejects_bad_precision_and_atoms(lib):
    spec, med = _case()
    mats = _abi.MaterialTable(spec["materials"])
    grid = _abi.make_grid(spec["grid"], 3)
    h = ctypes.c_void_p()
    rc = lib.phpfdtd_create(ctypes.byref(grid), mats.arr, mats.n, _abi.i32p(_abi.medium_array(med)), ctypes.byref(h))
    assert rc == _abi.EINVAL and b"precision" in lib.phpfdtd_last_error()
    b = ctypes.c_void_p()
    assert lib.phpfdtd_batch_create(3, ctypes.byref(b)) == _abi.EINVAL
    mats_spec = [dict(mt) for mt in spec["materials"]]
    lat = [i for i, mt in enumerate(mats_spec) if mt["natom"]][0]
    mats_spec[lat]["natom"] = 33
    mats = _abi.MaterialTable(mats_spec)
    grid = _abi.make_grid(spec["grid"], _abi.FP32)
    rc = lib.phpfdtd_create(ctypes.byref(grid), mats.arr, mats.n, _abi.i32p(_abi.medium_array(med)), ctypes.byref(h))
    assert rc == _abi.EINVAL
