"""bench.py's link to the committed profile: every step kernel of the plan
maps to a device symbol that has rows in the kernel trace and the PMC traffic
table of profiles/step_kernels_tip_slab_1024x1024x640.md, and the traffic it
reads back for the lattice kernel equals the bytes its state layout moves
(FP64: the compulsory bytes; compensated FP32: the words and the 1 KiB rows
of 4-bit codes, v's last codes in the free half of r's last row)."""
import os
import re
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

PLAN = {"fp32": ["yee_update[r8]", "ade_update[m5]", "ion_update[m4,N16]", "step_end"],
        "fp64": ["yee_update[r8]", "ade_update[m5]", "ion_update[m4,N16]", "step_end"]}


def profile_symbols(section):
    text = open(bench.PROFILE).read()
    part = text.split(section, 1)[1].split("\n## ", 1)[0]
    return set(re.findall(r"^\| `([^`]+)` \|", part, re.M))


def test_plan_kernels_have_profile_rows():
    trace = profile_symbols("## Kernel trace")
    pmc = profile_symbols("## HBM traffic")
    for form in ("fp32", "fp64"):
        for name in PLAN[form]:
            sym = bench.kernel_symbol(name, form)
            assert sym in trace, (form, name, sym)
            assert sym in pmc, (form, name, sym)


def test_lattice_kernel_traffic_is_its_compulsory_bytes():
    # 50,331,648 lattice slots: 6w + 4 bytes each (D, E, cell index), and r and v
    # read and written in 64-slot tiles of 1 KiB rows (N = 16): r 12 word rows + 2
    # code rows, v 12 + 1
    slots = 1024 * 1024 * 48
    moved = slots * (6 * 4 + 4) + slots // 64 * 2 * 1024 * ((12 + 2) + (12 + 1))
    t = bench.profile_traffic(bench.kernel_symbol("ion_update[m4,N16]", "fp32"))
    assert t is not None and abs(t / moved - 1) < 0.002, (t, moved)
    t64 = bench.profile_traffic(bench.kernel_symbol("ion_update[m4,N16]", "fp64"))
    algorithmic64 = slots * (6 * 8 + 4 + 12 * 16 * 8)
    assert t64 is not None and abs(t64 / algorithmic64 - 1) < 0.002, (t64, algorithmic64)
