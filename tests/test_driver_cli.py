"""Host-side input handling of the reference-compatible driver, checked
without a GPU: every case below is rejected while reading the inputs, before
the first device call, with the reference's message (fdtd.c:718-959,
1491-1506, 2863-2866)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import golden_cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "php_fdtd_amd", "phpfdtd_run")


def stage(tmp):
    src = os.path.join(golden_cases.GOLDEN, "hbn")
    for f in os.listdir(src):
        if f.endswith((".cfg", ".bin", ".txt")) and not f.startswith(("field_slice_", "source")):
            shutil.copy(os.path.join(src, f), tmp)


def run(tmp):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    return subprocess.run([DRIVER, "case.cfg"], cwd=tmp, capture_output=True, text=True, timeout=60)


def edit_cfg(tmp, fn):
    p = os.path.join(tmp, "case.cfg")
    lines = open(p).read().splitlines()
    open(p, "w").write("\n".join(fn(lines)) + "\n")


def test_medium_dimension_mismatch(tmp_path):
    stage(str(tmp_path))
    p = tmp_path / "medium_map.bin"
    raw = np.fromfile(p, dtype="<i4")
    raw[0] += 1
    raw.tofile(p)
    r = run(str(tmp_path))
    assert r.returncode != 0
    assert "Medium file dimensions (21,20,16) do not match simulation (20,20,16)" in r.stderr


def test_switchxz_on_simple_material(tmp_path):
    stage(str(tmp_path))
    edit_cfg(str(tmp_path), lambda ls: ls + ["material id=5, name=plain, eps_xx=2.0, switchxz=1"])
    r = run(str(tmp_path))
    assert r.returncode != 0
    assert "Please switch X and Z by hand for this simple material" in r.stderr


def test_dipole_needs_one_frequency(tmp_path):
    stage(str(tmp_path))
    edit_cfg(str(tmp_path), lambda ls: [l.replace("wavenumber=1370.0", "wavenumber=1370.0, frequency=4e13")
                                        for l in ls])
    r = run(str(tmp_path))
    assert r.returncode != 0
    assert "Cannot specify both frequency and wavenumber for dipole source" in r.stderr
