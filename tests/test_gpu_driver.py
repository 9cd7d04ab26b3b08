"""The reference-compatible driver (php_fdtd_amd/phpfdtd_run) run on each
golden case directory reproduces the files the reference solver wrote there.

Bars: source.txt identical text; the monitor and full-monitor files (driver
run with --precise, the golden files hold the reference's values at %.17e)
within 1e-10 relative L2 per column group; every field slice file present
with the same header and values within 1 float32 ulp.  The precision form
through the driver: monitor within 1e-4 relative L2.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import golden_cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "php_fdtd_amd", "phpfdtd_run")
OUTPUTS = ("monitor_point.dat", "full_monitor.dat", "source.txt", "reference.log", "case.json")


def stage(name, tmp):
    src = os.path.join(golden_cases.GOLDEN, name)
    for f in os.listdir(src):
        if f in OUTPUTS or f.startswith("field_slice_"):
            continue
        shutil.copy(os.path.join(src, f), tmp)
    return src


def run_driver(tmp, *flags):
    if not os.path.exists(DRIVER):
        raise FileNotFoundError("driver not built: " + DRIVER)
    p = subprocess.run([DRIVER, "case.cfg", *flags], cwd=tmp, capture_output=True, text=True, timeout=300)
    print(p.stdout[-2000:], p.stderr[-2000:], flush=True)
    assert p.returncode == 0, p.stderr
    return p.stdout


def rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / n if n else np.linalg.norm(a)


def ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


@pytest.mark.parametrize("name", golden_cases.CASES)
def test_driver_reproduces_reference_outputs(name, tmp_path):
    print("driver case", name, flush=True)
    src = stage(name, str(tmp_path))
    out = run_driver(str(tmp_path), "--precise")
    assert "Simulation completed in" in out
    assert open(tmp_path / "source.txt").read() == open(os.path.join(src, "source.txt")).read()
    for f in ("monitor_point.dat", "full_monitor.dat"):
        ref_p = os.path.join(src, f)
        if not os.path.exists(ref_p):
            assert not os.path.exists(tmp_path / f), f
            continue
        ref = golden_cases.read_table(ref_p)
        got = golden_cases.read_table(str(tmp_path / f))
        assert open(tmp_path / f).readline() == open(ref_p).readline(), "header of " + f
        assert got.shape == ref.shape, (f, got.shape, ref.shape)
        assert np.array_equal(got[:, 0], ref[:, 0]), f
        assert rel(got[:, 1:10], ref[:, 1:10]) <= 1e-10, f
        if ref.shape[1] > 10:
            assert rel(got[:, 10:], ref[:, 10:]) <= 1e-10, f
    ref_slices = sorted(p for p in os.listdir(src) if p.startswith("field_slice_"))
    got_slices = sorted(p for p in os.listdir(tmp_path) if p.startswith("field_slice_"))
    assert got_slices == ref_slices
    for p in ref_slices:
        a = np.fromfile(tmp_path / p, dtype="<i4", count=2)
        b = np.fromfile(os.path.join(src, p), dtype="<i4", count=2)
        assert np.array_equal(a, b), p
        assert ulp_diff(golden_cases.read_slice(str(tmp_path / p)), golden_cases.read_slice(os.path.join(src, p))).max() <= 1, p


def test_driver_precision_form(tmp_path):
    src = stage("hbn", str(tmp_path))
    run_driver(str(tmp_path), "--fp32")
    ref = golden_cases.read_table(os.path.join(src, "monitor_point.dat"))
    got = golden_cases.read_table(str(tmp_path / "monitor_point.dat"))
    assert got.shape == ref.shape
    assert rel(got[:, 1:10], ref[:, 1:10]) <= 1e-4
