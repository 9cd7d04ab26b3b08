"""Parity of the HIP path with the oracle on the benchmark workloads at their
full grids, on full-tensor ADE materials, and of the divergence guard.

Bars (stated here, checked below):
  * lorentz_slab (100^3), lattice4_slab (160x160x120) and lattice16_slab
    (384x384x256) at full size from a random state with every term active
    (fields, CPML psi reached through the fields, ADE P and P_old, lattice
    r and v), advanced one oracle step so that it is one a run reaches (E =
    M . D in plain cells, which the FP32 form assumes), 3 steps: FP64 within 1e-12 relative L2 of the oracle per
    component and state, diagnostics and argmax equal; the two FP32 forms
    within 1e-4.
  * Full, non-symmetric 3x3 Drude tensors through the ADE (the coupling the
    reference leaves undefined, fdtd.c:3465-3475; the oracle's recurrence is
    pinned by numpy's linear solve in tests/test_oracle_pins.py), rotated by
    anglez and by anglex: 5 steps from a random state, same bars.  A driven
    full-tensor ADE block over 10,000 steps: FP64 within 1e-8 of the oracle
    (rounding accumulated over 10^4 steps), the FP32 form within 1e-4 of FP64.
  * The divergence guard (fdtd.c:3626-3667): PHPFDTD_EDIVERGED exactly when
    max |r| > 1e-10 m, with the first argmax in (i, j, k, atom, dir) order;
    on the golden case the reference aborted ("diverge"), no output step
    before the reference's abort step diverges and the abort step reports the
    reference's location, atom, direction and max |r|.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import golden_cases
import states as reach_states
from oracle import orc
from php_fdtd_amd import _abi, workloads
from php_fdtd_amd.sim import Simulation

pytestmark = pytest.mark.gpu

RANDOM_TOL = 1e-12
FP32_TOL = 1e-4
FORMS = ((_abi.FP64, RANDOM_TOL), (_abi.FP32, FP32_TOL), (_abi.FP32_RV64, FP32_TOL))


def rel_l2(got, ref):
    ref = np.asarray(ref, np.float64)
    nr = np.linalg.norm(ref)
    d = np.linalg.norm(np.asarray(got, np.float64) - ref)
    return d / nr if nr else d


def random_state(spec, medium, seed):
    """Fields and every ADE / lattice state at the scales a run reaches
    (D ~ eps0 E, H ~ E / 377, r ~ 1e-13 m, v ~ 1e-3 m/s, P ~ 1e-11)."""
    rng = np.random.default_rng(seed)
    fields = []
    for comp in range(9):
        f = rng.standard_normal(medium.shape)
        f *= 8.854e-12 if 3 <= comp < 6 else (1 / 377.0 if comp >= 6 else 1.0)
        fields.append(f)
    states = {}
    counts = np.bincount(medium.ravel(), minlength=len(spec["materials"]))
    for m, mat in enumerate(spec["materials"]):
        if counts[m] == 0:
            continue
        if mat["natom"]:
            for kind, scale in ((0, 1e-13), (1, 1e-3)):
                states[(m, kind)] = rng.standard_normal(counts[m] * 3 * mat["natom"]) * scale
        if mat["use_ade"]:
            for kind in (2, 3):
                states[(m, kind)] = rng.standard_normal(counts[m] * 3) * 1e-11
    return fields, states


def apply_state(target, state):
    fields, states = state
    for comp, f in enumerate(fields):
        target.set_field(comp, f)
    for (m, kind), a in states.items():
        target.set_state(m, kind, a)


def check_forms_vs_oracle(spec, med, nstep, seed, label):
    # a reachable state (E = M . D in plain cells), which the FP32 form's
    # E-form update of those cells assumes
    state = reach_states.reachable(spec, med, *random_state(spec, med, seed))
    o = orc.Oracle(spec, med)
    apply_state(o, state)
    o.step(nstep)
    ref_f = [o.field(c) for c in range(9)]
    ref_s = {k: o.state(*k) for k in state[1]}
    ref_d = o.diag()
    o.close()
    assert any(k[1] in (0, 2) for k in ref_s), "no ADE or lattice state in " + label
    for prec, tol in FORMS:
        print(label, "precision", prec, flush=True)
        sim = Simulation(spec, med, prec)
        apply_state(sim, state)
        sim.run(nstep)
        for comp in range(9):
            e = rel_l2(sim.field(comp), ref_f[comp])
            assert e <= tol, "%s precision %d component %d rel err %g" % (label, prec, comp, e)
        for k, ref in ref_s.items():
            e = rel_l2(sim.state(*k), ref)
            assert e <= tol, "%s precision %d material %d kind %d rel err %g" % (label, prec, k[0], k[1], e)
        if prec == _abi.FP64:
            d = sim.sync()
            v, idx = ref_d
            assert abs(d.max_e - v[0]) <= 1e-12 * v[0]
            assert abs(d.max_ex - v[1]) <= 1e-12 * v[1]
            assert abs(d.max_r - v[2]) <= 1e-12 * max(v[2], 1e-300)
            if v[2] > 0:
                assert [d.r_i, d.r_j, d.r_k, d.r_atom, d.r_dir] == idx
        sim.close()


@pytest.mark.timeout(600)
@pytest.mark.parametrize("name", ["lorentz_slab", "lattice4_slab", "lattice16_slab"])
def test_config_workload_vs_oracle(name):
    spec, med = workloads.WORKLOADS[name]()
    spec = workloads.with_steps(spec, 8)
    print("workload", spec["name"], flush=True)
    check_forms_vs_oracle(spec, med, 3, 2024, spec["name"])


def full_tensor_mat(rng, f, **kw):
    """Drude-type material (Re eps - EPS_INF < 0 on the diagonal) with full,
    non-symmetric real and imaginary tensors, through the ADE."""
    er = np.diag(rng.uniform(-60.0, -25.0, 3)) + rng.uniform(-4.0, 4.0, (3, 3)) * (1 - np.eye(3))
    ei = np.diag(rng.uniform(2.0, 9.0, 3)) + rng.uniform(-1.0, 1.0, (3, 3)) * (1 - np.eye(3))
    m = workloads.vacuum()
    m.update(eps_re=er.ravel().tolist(), eps_im=ei.ravel().tolist(), use_ade=1, **kw)
    return m


def full_tensor_case(shape=(28, 20, 16), L=3, steps=40, seed=31):
    f = 3e13
    rng = np.random.default_rng(seed)
    g = workloads._grid(shape, 20e-9, steps, L, f)
    mats = workloads.builtin_table(f) + [full_tensor_mat(rng, f, anglez=0.4), full_tensor_mat(rng, f, anglex=0.6),
                                         full_tensor_mat(rng, f)]
    nx, ny, nz = shape
    med = np.zeros(shape, np.int32)
    med[2:nx // 2, 3:ny - 3, 4:9] = 4
    med[nx // 2:nx - 2, 3:ny - 3, 4:9] = 5
    med[4:nx - 4, 2:ny - 2, 10:13] = 6
    src = [workloads._source("dipole", (nx // 2, ny // 2, 2), "z", f, tsteps=steps, dt=g["dt"]),
           workloads._source("gaussian", (nx // 3, ny // 3, 14), "x", f, tsteps=steps, dt=g["dt"])]
    spec = dict(name="full_tensor_ade%dx%dx%d" % shape, grid=g, materials=mats, sources=src,
                monitor=dict(x=nx // 2, y=ny // 2, z=7, atom=0, dir=0))
    return spec, med


def test_full_tensor_ade_vs_oracle():
    spec, med = full_tensor_case()
    for m in (4, 5, 6):
        er = np.asarray(spec["materials"][m]["eps_re"]).reshape(3, 3)
        assert np.abs(er - er.T).max() > 0.1  # non-symmetric, non-diagonal
    check_forms_vs_oracle(spec, med, 5, 77, spec["name"])


@pytest.mark.timeout(300)
def test_full_tensor_ade_long_run():
    steps = 10000
    spec, med = full_tensor_case(shape=(24, 24, 20), L=4, steps=steps, seed=37)
    m = spec["monitor"]
    o = orc.Oracle(spec, med)
    o.step(steps)
    ref = [o.field(c) for c in range(9)]
    ref_p = {k: o.state(k, 2) for k in (4, 5, 6)}
    o.close()
    rows = {}
    for prec in (_abi.FP64, _abi.FP32):
        sim = Simulation(spec, med, prec)
        sim.set_monitor(m["x"], m["y"], m["z"], m["atom"], m["dir"], ref_compat=True)
        sim.run(steps)
        rows[prec] = sim.monitor(steps)
        if prec == _abi.FP64:
            for c in range(9):
                e = rel_l2(sim.field(c), ref[c])
                print("FP64 vs oracle after %d steps, component %d: %.3g" % (steps, c, e), flush=True)
                assert e <= 1e-8, (c, e)
            for k, p in ref_p.items():
                assert rel_l2(sim.state(k, 2), p) <= 1e-8, k
        sim.close()
    assert np.abs(rows[_abi.FP64][:, 1:10]).max() > 0
    err = rel_l2(rows[_abi.FP32][:, 1:10], rows[_abi.FP64][:, 1:10])
    print("full-tensor ADE, %d steps, FP32 vs FP64 monitor rel L2 %.3g" % (steps, err), flush=True)
    assert err <= 1e-4


def numpy_argmax(sim, spec, med):
    """max |r| and its first argmax in (i, j, k, atom, dir) order from the
    lattice state read back through the ABI."""
    flat = med.ravel()
    table = np.full((flat.size, 96), -1.0)
    for m, mat in enumerate(spec["materials"]):
        n = mat["natom"]
        if n and (flat == m).any():
            table[np.flatnonzero(flat == m), :3 * n] = np.abs(sim.state(m, 0)).reshape(-1, 3 * n)
    mx = float(table.max())
    cell, dof = divmod(int(np.argmax(table.ravel())), 96)
    return mx, list(np.unravel_index(cell, med.shape)) + [dof // 3, dof % 3]


def raw_sync(sim):
    d = _abi.Diag()
    rc = sim.lib.phpfdtd_sync(sim.h, ctypes.byref(d))
    return rc, d


@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32, _abi.FP32_RV64])
def test_divergence_guard(precision):
    c = golden_cases.load("moo3")
    spec, med = c["spec"], c["medium"]
    sim = Simulation(spec, med, precision)
    sim.run(10)
    rc, d = raw_sync(sim)
    assert rc == 0 and d.diverged == 0 and 0 < d.max_r <= 1e-10
    n = sim.count_cells(4)
    rng = np.random.default_rng(5)
    r = rng.standard_normal(n * 48) * 1e-12
    # a dof just above the guard; then the same magnitude at a dof earlier in
    # loop order (a tie: the first in (i, j, k, atom, dir) order wins); then a
    # smaller value at the very first dof, which must not move the argmax.
    # max |r| is the stored value itself in every form (compensated FP32 state
    # included), read back exactly by phpfdtd_read_state.
    for pos, val in ((n * 48 // 2 + 17, 1.5e-10), (n * 48 // 3 + 5, -1.5e-10), (0, 0.9e-10)):
        r[pos] = val
        sim.set_state(4, 0, r)
        rc, d = raw_sync(sim)
        mx, idx = numpy_argmax(sim, spec, med)
        assert d.max_r == mx
        assert [d.r_i, d.r_j, d.r_k, d.r_atom, d.r_dir] == idx, (pos, idx)
        assert rc == _abi.EDIVERGED and d.diverged == 1
        assert "r field too large" in sim.lib.phpfdtd_last_error().decode()
    r[:] = 0.9e-10
    sim.set_state(4, 0, r)
    rc, d = raw_sync(sim)
    assert rc == 0 and d.diverged == 0
    assert [d.r_i, d.r_j, d.r_k, d.r_atom, d.r_dir] == numpy_argmax(sim, spec, med)[1]
    sim.close()


def test_diverge_case_matches_reference_abort():
    c = golden_cases.load("diverge")
    spec, med = c["spec"], c["medium"]
    txt = open(os.path.join(c["dir"], "reference.log")).read()
    m = re.search(r"ERROR: r field too large at step (\d+)\nMax r value: (\S+)\n"
                  r"Location: i=(\d+), j=(\d+), k=(\d+) \(global k=\d+\)\nAtom: (\d+), Direction: (\d+)\n", txt)
    assert m
    abort = int(m.group(1))
    max_r = float(m.group(2))
    loc = [int(m.group(q)) for q in (3, 4, 5, 6, 7)]
    interval = spec["output"]["interval"]
    sim = Simulation(spec, med, _abi.FP64)
    for t in range(0, abort + 1, interval):
        sim.run(1)  # step t
        rc, d = raw_sync(sim)
        if t < abort:
            assert rc == 0 and not d.diverged, t
            sim.run(interval - 1)
    assert rc == _abi.EDIVERGED and d.diverged
    assert [d.r_i, d.r_j, d.r_k, d.r_atom, d.r_dir] == loc
    assert abs(d.max_r - max_r) <= 1e-6 * max_r  # printed with %e: 7 significant digits
    sim.close()
