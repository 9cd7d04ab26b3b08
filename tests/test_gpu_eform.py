"""The precision forms' E-form update of plain cells (E^{n+1} = E^n + M . dD,
D not stored; fdtd.c:3238-3462 restated) against the oracle, which runs the
reference's D-form.

Bars (stated here, checked below):
  * D of plain cells, which these forms derive as M^-1 E when asked, matches
    the oracle's D through every reader: whole fields, points, slices of each
    axis, monitor rows and a D-component near-field DFT, within 1e-4
    relative L2 after a driven run (the FP32 bar of the other parity tests).
  * Written into an FP32 context, a reachable state's D comes back within
    1e-6 (FP32 rounding of E, then M^-1 E in FP64): the derivation matrix is
    the inverse of the one the update uses.
  * A plain material whose M has no inverse (eps_re = I, eps_im = diag(-1,
    -1, 1): the reference's pseudo-complex inverse, fdtd.c:555-625, has real
    part diag(0, 0, 1)/1) keeps D as its state in these forms; its cells'
    fields match the oracle within 1e-4 as well.
"""
import numpy as np
import pytest

import golden_cases
import states as reach_states
from oracle import orc
from php_fdtd_amd import _abi, workloads
from php_fdtd_amd.sim import Simulation

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-4


def rel_l2(got, ref):
    ref = np.asarray(ref, np.float64)
    nr = np.linalg.norm(ref)
    d = np.linalg.norm(np.asarray(got, np.float64) - ref)
    return d / nr if nr else d


def singular_case(steps=160):
    """24x20x18 grid, 3-cell CPML: a rotated lossy anisotropic block (plain,
    E-form), the singular-M block (keeps D), a Drude ADE block (slot cells)
    and two dipoles, so every kind of cell lies next to the others."""
    f = 3e13
    g = workloads._grid((24, 20, 18), 20e-9, steps, 3, f)
    aniso = workloads.vacuum()
    aniso.update(eps_re=[4.0, 0.3, 0.0, 0.2, 3.0, 0.1, 0.0, 0.4, 2.5], eps_im=workloads.diag(0.2, 0.1, 0.3),
                 anglez=0.5)
    sing = workloads.vacuum()
    sing.update(eps_re=workloads.diag(1, 1, 1), eps_im=workloads.diag(-1, -1, 1))
    mats = workloads.builtin_table(f) + [aniso, sing, workloads.material_lorentz(f)]
    med = np.zeros((24, 20, 18), np.int32)
    med[4:12, 3:17, 5:10] = 4
    med[12:20, 3:17, 5:10] = 5
    med[6:18, 5:15, 11:14] = 6
    src = [workloads._source("dipole", (10, 10, 3), "z", f, tsteps=steps, dt=g["dt"]),
           workloads._source("gaussian", (15, 8, 15), "x", f, tsteps=steps, dt=g["dt"])]
    spec = dict(name="eform_singular", grid=g, materials=mats, sources=src,
                monitor=dict(x=8, y=10, z=7, atom=0, dir=0))
    return spec, med


@pytest.mark.parametrize("prec", [_abi.FP32, _abi.FP32_RV64])
def test_plain_d_readers_vs_oracle(prec):
    spec, med = singular_case()
    steps = spec["grid"]["tsteps"]
    m = spec["monitor"]
    print("E-form readers, precision", prec, flush=True)
    o = orc.Oracle(spec, med)
    sim = Simulation(spec, med, prec)
    sim.set_monitor(m["x"], m["y"], m["z"])  # plane z+1, as the reference samples
    o.step(steps)
    sim.run(steps)
    for comp in range(9):
        e = rel_l2(sim.field(comp), o.field(comp))
        assert e <= FP32_TOL, "component %d rel err %g" % (comp, e)
    # cells of each material alone: the E-form block, the singular-M block (D kept)
    for mat in (4, 5):
        sel = med == mat
        for comp in range(3, 6):
            e = rel_l2(sim.field(comp)[sel], o.field(comp)[sel])
            assert e <= FP32_TOL, "material %d D component %d rel err %g" % (mat, comp, e)
    for axis, pos in (("z", 7), ("x", 8), ("y", 10)):
        for comp in (3, 4, 5):
            e = rel_l2(sim.slice(comp, axis, pos), o.slice(comp, axis, pos))  # z: plane pos+1
            assert e <= FP32_TOL, "slice %s%d component %d rel err %g" % (axis, pos, comp, e)
    for cell in ((8, 10, 7), (15, 10, 7), (1, 1, 1)):
        ref = np.array([o.field(q)[cell] for q in range(9)])
        got = sim.point(*cell)
        assert rel_l2(got[3:6], ref[3:6]) <= FP32_TOL, cell
    rows = sim.monitor()
    assert len(rows) == steps
    last = o.monitor_row(m["x"], m["y"], m["z"])
    assert rel_l2(rows[-1][4:7], last[4:7]) <= FP32_TOL
    sim.close()
    o.close()


def test_monitor_d_trace_vs_fp64():
    """The whole monitor trace of D at a plain cell (derived every step inside
    the step epilogue) against the FP64 form's stored D."""
    spec, med = singular_case()
    steps = spec["grid"]["tsteps"]
    m = spec["monitor"]
    rows = {}
    for prec in (_abi.FP64, _abi.FP32):
        sim = Simulation(spec, med, prec)
        sim.set_monitor(m["x"], m["y"], m["z"])
        sim.run(steps)
        rows[prec] = sim.monitor()
        sim.close()
    e = rel_l2(rows[_abi.FP32][:, 4:7], rows[_abi.FP64][:, 4:7])
    print("monitor D trace FP32 vs FP64 rel L2", e, flush=True)
    assert np.abs(rows[_abi.FP64][:, 4:7]).max() > 0
    assert e <= FP32_TOL


def test_d_component_dft_vs_fp64():
    spec, med = singular_case()
    steps = spec["grid"]["tsteps"]
    N = steps // 2
    res = {}
    for prec in (_abi.FP64, _abi.FP32):
        sim = Simulation(spec, med, prec)
        sim.set_dft("Dz", "z", 7, 2, N, spec["grid"]["f_source"], ref_compat=False)
        sim.run(steps)
        res[prec], n = sim.dft()
        assert n == N
        sim.close()
    amp = rel_l2(np.abs(res[_abi.FP32]), np.abs(res[_abi.FP64]))
    print("Dz DFT amplitude FP32 vs FP64 rel L2", amp, flush=True)
    assert np.abs(res[_abi.FP64]).max() > 0
    assert amp <= 1e-3


def test_written_state_d_comes_back():
    c = golden_cases.load("aniso_loss")
    spec, med = c["spec"], c["medium"]
    rng = np.random.default_rng(17)
    fields = [rng.standard_normal(med.shape) * (8.854e-12 if 3 <= q < 6 else (1 / 377.0 if q >= 6 else 1.0))
              for q in range(9)]
    fields, _ = reach_states.reachable(spec, med, fields, {})
    sim = Simulation(spec, med, _abi.FP32)
    for q, f in enumerate(fields):
        sim.set_field(q, f)
    for q in range(3, 6):
        e = rel_l2(sim.field(q), fields[q])
        assert e <= 1e-6, "D component %d rel err %g" % (q, e)
    sim.close()
