"""The reference-compatible driver (php_fdtd_amd/phpfdtd_run) run on each
golden case directory reproduces the files the reference solver wrote there
and what it printed.

Bars: source.txt identical text; the monitor and full-monitor files (driver
run with --precise, the golden files hold the reference's values at %.17e)
within 1e-10 relative L2 per column group; every field slice file present
with the same header and values within 1 float32 ulp; every progress line
(fdtd.c:3562-3594: step, percentage, time, |E|_max, r_max) as the reference
printed it, the numbers equal at their printed precision (one unit of the
last printed digit).  On the case the reference aborted at its divergence
guard: exit status 1, the same progress lines up to the abort, and the
divergence report of fdtd.c:3626-3658 (step, location, atom, direction
exactly; max r, v and a at %e precision).  The precision form through the
driver: monitor within 1e-4 relative L2.
"""
import os
import re
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


def run_driver(tmp, *flags, rc=0):
    if not os.path.exists(DRIVER):
        raise FileNotFoundError("driver not built: " + DRIVER)
    p = subprocess.run([DRIVER, "case.cfg", *flags], cwd=tmp, capture_output=True, text=True, timeout=300)
    print(p.stdout[-2000:], p.stderr[-2000:], flush=True)
    assert p.returncode == rc, p.stderr
    return p.stdout


PROGRESS = re.compile(r"^it (\d+)/(\d+) \((\S+)%\), t = (\S+) s,\|E\|_max = (\S+) V/m,r_max = (\S+) A,")
REPORT = re.compile(r"ERROR: r field too large at step (\d+)\nMax r value: (\S+)\n"
                    r"Location: i=(\d+), j=(\d+), k=(\d+) \(global k=(\d+)\)\nAtom: (\d+), Direction: (\d+)\n"
                    r"v value: (\S+)\na value: (\S+)\n")


def same_printed(a, b):
    """Two numbers printed with the same format agree to one unit of their
    last printed digit (FP64 results within 1e-10 can straddle a rounding
    boundary of the printed value)."""
    x, y = float(a), float(b)
    if x == y:
        return True
    mant = b.lower().split("e")[0]
    digits = len(mant.split(".")[1]) if "." in mant else 0
    exp = int(b.lower().split("e")[1]) if "e" in b.lower() else 0
    return abs(x - y) <= 1.0001 * 10.0 ** (exp - digits)


def progress(text):
    return [m.groups() for m in map(PROGRESS.match, text.splitlines()) if m]


def check_progress(out, ref_log):
    got, ref = progress(out), progress(ref_log)
    assert len(got) == len(ref) > 0, (len(got), len(ref))
    for g, r in zip(got, ref):
        assert g[:4] == r[:4], (g, r)  # step, steps, percentage, time: identical text
        assert same_printed(g[4], r[4]) and same_printed(g[5], r[5]), (g, r)


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
    check_progress(out, open(os.path.join(src, "reference.log")).read())
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


@pytest.mark.parametrize("name", golden_cases.ABORTED)
def test_driver_reproduces_reference_abort(name, tmp_path):
    print("driver abort case", name, flush=True)
    src = stage(name, str(tmp_path))
    out = run_driver(str(tmp_path), "--precise", rc=1)
    ref_log = open(os.path.join(src, "reference.log")).read()
    check_progress(out, ref_log)
    got, ref = REPORT.search(out), REPORT.search(ref_log)
    assert got and ref, out[-1500:]
    g, r = got.groups(), ref.groups()
    assert g[0] == r[0] and g[2:8] == r[2:8], (g, r)  # step, i, j, k, global k, atom, direction
    for q in (1, 8, 9):  # max r, v value, a value
        assert same_printed(g[q], r[q]), (q, g[q], r[q])
    assert "Simulation completed in" not in out
    # the slices written before the abort
    ref_slices = sorted(p for p in os.listdir(src) if p.startswith("field_slice_"))
    got_slices = sorted(p for p in os.listdir(tmp_path) if p.startswith("field_slice_"))
    assert got_slices == ref_slices and ref_slices
    for p in ref_slices:
        assert ulp_diff(golden_cases.read_slice(str(tmp_path / p)),
                        golden_cases.read_slice(os.path.join(src, p))).max() <= 1, p
