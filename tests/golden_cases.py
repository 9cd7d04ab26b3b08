"""Loader for the golden fixtures written by tests/golden/make_golden.py."""
import glob
import json
import os
import re

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALL_CASES = sorted(d for d in os.listdir(GOLDEN) if os.path.exists(os.path.join(GOLDEN, d, "case.json")))


def _aborts(name):
    return bool(json.load(open(os.path.join(GOLDEN, name, "case.json"))).get("aborts"))


# cases the reference ran to completion (monitor, full monitor and every slice
# present); ABORTED: runs the reference stopped at its divergence guard
# (fdtd.c:3626-3667), which keep only the log and the slices written before it
CASES = [c for c in ALL_CASES if not _aborts(c)]
ABORTED = [c for c in ALL_CASES if _aborts(c)]


def read_medium(path):
    raw = np.fromfile(path, dtype="<i4")
    nx, ny, nz = (int(v) for v in raw[:3])
    return raw[3:].reshape(nx, ny, nz).astype(np.int32)


def read_table(path):
    rows = []
    for ln in open(path):
        if ln.startswith("#") or not ln.strip():
            continue
        rows.append([float(x) for x in ln.split()])
    return np.array(rows, dtype=np.float64)


def read_slice(path):
    raw = np.fromfile(path, dtype="<i4", count=2)
    data = np.fromfile(path, dtype="<f4", offset=8)
    return data.reshape(int(raw[0]), int(raw[1]))


def load(name):
    d = os.path.join(GOLDEN, name)
    spec = json.load(open(os.path.join(d, "case.json")))
    medium = read_medium(os.path.join(d, "medium_map.bin"))
    mon_p, full_p = os.path.join(d, "monitor_point.dat"), os.path.join(d, "full_monitor.dat")
    mon = read_table(mon_p) if os.path.exists(mon_p) else None
    full = read_table(full_p) if os.path.exists(full_p) else None
    slices = {}
    for p in glob.glob(os.path.join(d, "field_slice_*.bin")):
        step = int(re.search(r"_(\d{8})\.bin$", p).group(1))
        slices[step] = read_slice(p)
    return dict(dir=d, spec=spec, medium=medium, monitor=mon, full=full, slices=slices)
