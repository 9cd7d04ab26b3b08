"""Size-independent properties at the full 1024x1024x640 size (tip_slab workload).

The update is linear in the source amplitude and scaling by 2 is exact in
binary floating point, so doubling every source amplitude doubles every
field, monitor value, slice value and lattice state bit for bit (FP64; the
FP32 form within 1e-6, where subnormal tails may round differently).  A run
without sources stays exactly zero.
"""
import numpy as np
import pytest

from php_fdtd_amd import _abi, workloads
from php_fdtd_amd.sim import Simulation

pytestmark = pytest.mark.gpu

STEPS = 24


def sample(spec, med, precision, scale):
    sp = workloads.with_steps(spec, STEPS + 8)
    sp = dict(sp, sources=[dict(s, amplitude=s["amplitude"] * scale) for s in sp["sources"]])
    sim = Simulation(sp, med, precision)
    m = spec["monitor"]
    sim.set_monitor(m["x"], m["y"], m["z"], m["atom"], m["dir"], ref_compat=True)
    sim.run(STEPS)
    src = sp["sources"][0]
    out = {"rows": sim.monitor()}
    for comp in ("Ez", "Hx", "Dy"):
        out["slice_z_" + comp] = sim.slice(comp, "z", src["z"])
        out["slice_x_" + comp] = sim.slice(comp, "x", src["x"])
    r, v = sim.ions(m["x"], m["y"], m["z"])
    out["r"], out["v"] = r, v
    for dz in (-3, -1, 0, 2, 6):
        out["point%d" % dz] = sim.point(src["x"], src["y"], src["z"] + dz)
    d = sim.sync()
    out["max_e"] = np.array([d.max_e, d.max_ex, d.max_r])
    sim.close()
    return out


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32])
def test_full_size_linearity(precision):
    spec, med = workloads.tip_slab()
    print("tip_slab 1024x1024x640, precision", precision, flush=True)
    a = sample(spec, med, precision, 1.0)
    b = sample(spec, med, precision, 2.0)
    assert np.abs(a["slice_z_Ez"]).max() > 0 and np.abs(a["max_e"]).max() > 0
    for k in a:
        x, y = a[k], b[k]
        if k == "rows":
            x, y = x[:, 1:], y[:, 1:]  # the time column does not scale
            assert np.array_equal(a[k][:, 0], b[k][:, 0])
        if precision == _abi.FP64 and np.asarray(x).dtype == np.float32:
            # float32 slices of FP64 fields: the cast rounds subnormals, which
            # does not commute with scaling; exact above the normal range
            x32, y32 = np.asarray(x, np.float64), np.asarray(y, np.float64)
            normal = np.abs(y32) >= 2.0 ** -125
            assert np.array_equal(2 * x32[normal], y32[normal]), k
            assert np.abs(2 * x32[~normal] - y32[~normal]).max(initial=0) <= 2.0 ** -148, k
        elif precision == _abi.FP64:
            assert np.array_equal(2 * np.asarray(x, np.float64), np.asarray(y, np.float64)), k
        else:
            n = np.linalg.norm(y)
            assert n == 0 or np.linalg.norm(2 * np.asarray(x, np.float64) - y) / n <= 1e-6, k


def test_no_sources_stays_zero():
    spec, med = workloads.lattice4_slab()
    spec = dict(workloads.with_steps(spec, 12), sources=[])
    for precision in (_abi.FP64, _abi.FP32):
        sim = Simulation(spec, med, precision)
        sim.run(12)
        d = sim.sync()
        assert d.max_e == 0 and d.max_ex == 0 and d.max_r == 0
        assert not np.any(sim.slice("E", "z", 60))
        sim.close()
