"""The reference-side binding (php_fdtd_amd/binding/fdtd_binding.c) inside the
reference's own fdtd.c: oracle/_ref/fdtd_bound.x is the unmodified fdtd.c
(main() renamed at compile time) plus the binding and oracle/bound_main.c,
which loads a case with fdtd.c's read_config() and
read_medium_map_parallel() and hands it to libphpfdtd_hip through the
binding's two functions.

On every golden case whose materials are complete after read_config() (cfg
keys and the built-in library: no Born, QeZV/QeZM/DMM or dielectric-table
files, whose loading is inline code of main()), the bound binary must
reproduce the reference's monitor file within 1e-10 relative L2 and its
progress lines (step, percentage, time identical; |E|_max and r_max to one
unit of the printed digit).  The CPU test checks that the binding compiles
against fdtd.c's types wherever the reference sources are present."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import golden_cases
from test_gpu_driver import check_progress, rel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BOUND = os.path.join(ROOT, "oracle", "_ref", "fdtd_bound.x")
REF = "/root/reference"
KEY_CASES = ["vacuum_cpml", "aniso_loss", "builtins_nopml"]


def test_binding_compiles_against_fdtd_types():
    if not os.path.exists(os.path.join(REF, "fdtd.c")):
        pytest.skip("reference sources absent")
    p = subprocess.run(["gcc", "-fsyntax-only", "-w", "-I" + REF, "-I" + os.path.join(ROOT, "php_fdtd_amd", "binding"),
                        "-I" + os.path.join(ROOT, "include"), "-I/opt/conda/include",
                        os.path.join(ROOT, "oracle", "bound_main.c")], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-3000:]


@pytest.mark.gpu
@pytest.mark.parametrize("name", KEY_CASES)
def test_bound_reference_reproduces_monitor(name, tmp_path):
    print("bound case", name, flush=True)
    if not os.path.exists(BOUND):
        pytest.skip("oracle/_ref/fdtd_bound.x not built (needs the reference sources at build time)")
    src = os.path.join(golden_cases.GOLDEN, name)
    for f in ("case.cfg", "medium_map.bin"):
        shutil.copy(os.path.join(src, f), tmp_path)
    p = subprocess.run([BOUND, "case.cfg"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(p.stdout[-1500:], p.stderr[-1500:], flush=True)
    assert p.returncode == 0, p.stderr[-2000:]
    check_progress(p.stdout, open(os.path.join(src, "reference.log")).read())
    ref = golden_cases.read_table(os.path.join(src, "monitor_point.dat"))
    got = golden_cases.read_table(str(tmp_path / "bound_monitor.dat"))
    assert got.shape[0] == ref.shape[0] and got.shape[1] == 10, (got.shape, ref.shape)
    assert np.array_equal(got[:, 0], ref[:, 0])
    assert rel(got[:, 1:10], ref[:, 1:10]) <= 1e-10
