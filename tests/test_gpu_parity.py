"""Parity of the HIP path (through the C ABI) with the reference and the oracle.

Bars (stated here, checked below):
  * FP64 plain form vs the reference's golden runs: monitor traces and
    full-monitor rows within 1e-10 relative L2, float32 slices within 1 ulp.
  * FP64 plain form vs the CPU oracle from random states (every term active):
    every field, lattice and ADE state within 1e-12 relative L2 per component.
  * FP32 precision form vs the golden FP64 traces: relative L2 <= 1e-4.
  * Random states of the FP32 form are reachable ones (tests/states.py): the
    form carries E alone for plain cells, so a state whose E is not M . D
    there is not one it can hold.
"""
import os

import numpy as np
import pytest

import golden_cases
import states as reach_states
from oracle import orc
from php_fdtd_amd import _abi, workloads
from php_fdtd_amd.sim import Simulation

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
RANDOM_TOL = 1e-12
FP32_TOL = 1e-4


def rel_l2(got, ref):
    ref = np.asarray(ref, np.float64)
    got = np.asarray(got, np.float64)
    nr = np.linalg.norm(ref)
    if nr == 0:
        return np.linalg.norm(got)
    return np.linalg.norm(got - ref) / nr


def ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


def run_case(c, precision):
    spec = c["spec"]
    print("case", spec["name"], "precision", precision, flush=True)
    sim = Simulation(spec, c["medium"], precision)
    mon = spec["monitor"]
    out = spec["output"]
    sim.set_monitor(mon["x"], mon["y"], mon["z"], mon["atom"], mon["dir"], ref_compat=True)
    T = spec["grid"]["tsteps"]
    slices, full = {}, []
    t = 0
    while t < T:
        if t % out["interval"] == 0:
            sim.run(1)
            slices[t] = sim.slice(out["component"], out["axis"], out["pos"])
            row = [t * spec["grid"]["dt"]] + list(sim.point(mon["x"], mon["y"], min(mon["z"] + 1, spec["grid"]["nz"])))
            r, v = sim.ions(mon["x"], mon["y"], mon["z"])
            for a in range(len(r) // 3):
                row += list(r[3 * a:3 * a + 3]) + list(v[3 * a:3 * a + 3])
            full.append(row)
            t += 1
        else:
            n = min(out["interval"] - t % out["interval"], T - t)
            sim.run(n)
            t += n
    rows = sim.monitor()
    d = sim.sync()
    sim.close()
    return rows, slices, np.array(full), d


@pytest.fixture(scope="module", params=golden_cases.CASES)
def golden(request):
    return golden_cases.load(request.param)


def test_golden_fp64(golden):
    c = golden
    rows, slices, full, d = run_case(c, _abi.FP64)
    ref = c["monitor"]
    assert rows.shape[0] == ref.shape[0]
    np.testing.assert_array_equal(rows[:, 0], ref[:, 0])
    assert rel_l2(rows[:, 1:10], ref[:, 1:10]) <= FP64_TOL
    if ref.shape[1] > 10:
        assert rel_l2(rows[:, 10:12], ref[:, 10:12]) <= FP64_TOL
    for t, s in c["slices"].items():
        assert ulp_diff(slices[t], s).max() <= 1, "slice at step %d" % t
    assert full.shape == c["full"].shape
    assert rel_l2(full[:, 1:], c["full"][:, 1:]) <= FP64_TOL
    assert d.diverged == 0


def test_golden_fp32(golden):
    c = golden
    rows, slices, full, d = run_case(c, _abi.FP32)
    ref = c["monitor"]
    if np.abs(ref[:, 1:10]).max() > 0:
        err = rel_l2(rows[:, 1:10], ref[:, 1:10])
        print("fp32 monitor rel L2", err)
        assert err <= FP32_TOL
    if ref.shape[1] > 10:
        assert rel_l2(rows[:, 10:12], ref[:, 10:12]) <= FP32_TOL


def randomize(spec, medium, seed, sim, o, reach=False):
    """Write one random state into sim and o: fields at the scales a run
    reaches (D ~ eps0 E, H ~ E / 377) and every lattice and ADE state; with
    reach, that state advanced one oracle step (E = M . D in plain cells)."""
    rng = np.random.default_rng(seed)
    shape = medium.shape
    fields = []
    for comp in range(9):
        f = rng.standard_normal(shape)
        if comp >= 3 and comp < 6:
            f *= 8.854e-12  # D ~ eps0 E
        elif comp >= 6:
            f /= 377.0
        fields.append(f)
    states = {}
    for m, mat in enumerate(spec["materials"]):
        n = sim.count_cells(m)
        if n == 0:
            continue
        if mat["natom"] > 0:
            for kind, scale in ((0, 1e-13), (1, 1e-3)):
                states[(m, kind)] = rng.standard_normal(n * 3 * mat["natom"]) * scale
        if mat["use_ade"]:
            for kind in (2, 3):
                states[(m, kind)] = rng.standard_normal(n * 3) * 1e-11
    if reach:
        fields, states = reach_states.reachable(spec, medium, fields, states)
    for comp, f in enumerate(fields):
        sim.set_field(comp, f)
        o.set_field(comp, f)
    for (m, kind), a in states.items():
        sim.set_state(m, kind, a)
        o.set_state(m, kind, a)


@pytest.mark.parametrize("name", golden_cases.CASES)
def test_random_state_vs_oracle(name):
    c = golden_cases.load(name)
    spec = c["spec"]
    print("random-state case", name, flush=True)
    sim = Simulation(spec, c["medium"], _abi.FP64)
    o = orc.Oracle(spec, c["medium"])
    randomize(spec, c["medium"], 1234, sim, o)
    nstep = 5
    sim.run(nstep)
    o.step(nstep)
    for comp in range(9):
        e = rel_l2(sim.field(comp), o.field(comp))
        assert e <= RANDOM_TOL, "component %d rel err %g" % (comp, e)
    for m, mat in enumerate(spec["materials"]):
        if sim.count_cells(m) == 0:
            continue
        kinds = ([0, 1] if mat["natom"] else []) + ([2, 3] if mat["use_ade"] else [])
        for kind in kinds:
            e = rel_l2(sim.state(m, kind), o.state(m, kind))
            assert e <= RANDOM_TOL, "material %d kind %d rel err %g" % (m, kind, e)
    d = sim.sync()
    v, idx = o.diag()
    assert abs(d.max_e - v[0]) <= 1e-12 * v[0]
    assert abs(d.max_ex - v[1]) <= 1e-12 * v[1]
    assert abs(d.max_r - v[2]) <= 1e-12 * max(v[2], 1e-300)
    if v[2] > 0:
        assert [d.r_i, d.r_j, d.r_k, d.r_atom, d.r_dir] == idx
    sim.close()


def small_lattice_case(natom):
    """16x16x12 grid, 3-cell CPML, a rotated slab of a synthetic natom-atom
    lattice (atom counts outside the kernels' exact instantiations take the
    generic path)."""
    f = 1000.0 * workloads.WN_TO_HZ
    g = workloads._grid((16, 16, 12), 20e-9, 40, 3, f)
    lat = workloads.lattice(natom, [20.0 + 3 * a for a in range(natom)], 40.0, 500.0, 1500.0, 2.0, 100 + natom)
    mat = workloads.vacuum()
    mat.update(eps_re=workloads.diag(3.0, 3.5, 2.5), anglez=0.3, natom=natom, qezv=lat["qezv"], qezm=lat["qezm"],
               dmm=lat["dmm"])
    med = np.zeros((16, 16, 12), np.int32)
    med[:, :, 4:8] = 4
    src = [workloads._source("dipole", (8, 8, 2), "z", f, tsteps=40, dt=g["dt"])]
    spec = dict(name="lattice%d" % natom, grid=g, materials=workloads.builtin_table(f) + [mat], sources=src)
    return spec, med


@pytest.mark.parametrize("natom", [1, 3, 5, 20, 32])
def test_lattice_atom_counts(natom):
    """N = 1, 3 take exact instantiations; 5, 20 and 32 (the largest count the
    library accepts) take the generic runtime-N kernels, 20 and 32 their
    64-lane work-groups with up to 96 dofs."""
    spec, med = small_lattice_case(natom)
    print("lattice natom", natom, flush=True)
    forms = (_abi.FP64, _abi.FP32, _abi.FP32_RV64)
    sims = {p: Simulation(spec, med, p) for p in forms}
    o = orc.Oracle(spec, med)
    randomize(spec, med, 99, sims[_abi.FP64], o, reach=True)
    for p in forms[1:]:
        randomize(spec, med, 99, sims[p], _Sink(), reach=True)
    for s in sims.values():
        s.run(5)
    o.step(5)
    comps = list(range(9))
    for comp in comps:
        assert rel_l2(sims[_abi.FP64].field(comp), o.field(comp)) <= RANDOM_TOL, comp
        for p in forms[1:]:
            assert rel_l2(sims[p].field(comp), o.field(comp)) <= FP32_TOL, (p, comp)
    for kind in (0, 1):
        ref = o.state(4, kind)
        assert rel_l2(sims[_abi.FP64].state(4, kind), ref) <= RANDOM_TOL, kind
        for p in forms[1:]:
            assert rel_l2(sims[p].state(4, kind), ref) <= FP32_TOL, (p, kind)
    for s in sims.values():
        s.close()


def ragged_case(shape, L=3):
    """A grid whose x, y and z extents end inside a wavefront tile, a y block
    and a z chunk of both forms (FP32: 256-cell x tiles, 4-plane chunks;
    FP64: 128-cell x tiles, 2-row waves, 16-plane chunks), with a 3-cell CPML
    on every face, a rotated 4-atom lattice slab and a Drude ADE block that
    straddle the x tile seams, a lossy dielectric block, and Gaussian sources
    (peak at t = 0, so they inject in the first steps) on both sides of each
    seam.  With L = 3 the x-PML lanes mix PML and interior cells (per-cell
    psi path); with L = 4 and nx a multiple of 4 every x-PML lane holds four
    consecutive slab cells (vector psi path)."""
    nx, ny, nz = shape
    f = 1000.0 * workloads.WN_TO_HZ
    g = workloads._grid(shape, 20e-9, 40, L, f)
    lossy = dict(workloads.material_lorentz(f), use_ade=0)
    mats = workloads.builtin_table(f) + [workloads.material_lattice4(), workloads.material_drude_tip(f), lossy]
    med = np.zeros(shape, np.int32)
    med[nx // 3:nx - 5, 2:ny - 2, 8:12] = 4
    med[100:nx - 10, 1:ny - 1, 14:18] = 5
    med[4:60, 3:ny - 3, 2:7] = 6
    amp = 1e-11 / g["dt"]  # D changes by ~1e-11, the size of the random D state
    src = [workloads._source("gaussian", (x, ny // 2, z), c, f, t0=0.0, amplitude=amp, tsteps=40, dt=g["dt"])
           for x, z, c in ((127, 5, "x"), (128, 13, "y"), (255, 19, "z"), (256, 10, "x"), (nx - 2, 16, "z"))]
    spec = dict(name="ragged%dx%dx%d" % shape, grid=g, materials=mats, sources=src)
    return spec, med


@pytest.mark.parametrize("plan", ["fused", "split"])
@pytest.mark.parametrize("shape,L", [((300, 13, 21), 3), ((257, 9, 18), 3), ((264, 12, 20), 4)])
def test_ragged_tiles_vs_oracle(monkeypatch, shape, L, plan):
    """Every field, lattice and ADE state of every form against the oracle
    from a random (reachable) state on grids with ragged x tiles, y blocks and
    z chunks, under both Yee plans (one-pass, and H then D/E)."""
    monkeypatch.setenv("PHPFDTD_YEE", plan)
    spec, med = ragged_case(shape, L)
    print("ragged case", shape, "L", L, "plan", plan, flush=True)
    o = orc.Oracle(spec, med)
    for prec, tol in ((_abi.FP64, RANDOM_TOL), (_abi.FP32, FP32_TOL), (_abi.FP32_RV64, FP32_TOL)):
        sim = Simulation(spec, med, prec)
        randomize(spec, med, 7, sim, o if prec == _abi.FP64 else _Sink(), reach=True)
        if prec == _abi.FP64:
            o.step(6)
        sim.run(6)
        for comp in range(9):
            e = rel_l2(sim.field(comp), o.field(comp))
            assert e <= tol, "precision %d component %d rel err %g" % (prec, comp, e)
        for m, kinds in ((4, (0, 1)), (5, (2, 3))):
            assert sim.count_cells(m) == int((med == m).sum())
            for kind in kinds:
                e = rel_l2(sim.state(m, kind), o.state(m, kind))
                assert e <= tol, "precision %d material %d kind %d rel err %g" % (prec, m, kind, e)
        if prec == _abi.FP64:
            d = sim.sync()
            v, idx = o.diag()
            assert abs(d.max_e - v[0]) <= 1e-12 * v[0]
            assert abs(d.max_r - v[2]) <= 1e-12 * v[2]
            assert [d.r_i, d.r_j, d.r_k, d.r_atom, d.r_dir] == idx
        sim.close()


def test_compensated_state():
    """Lattice r, v of the FP32 form are stored as an FP32 word plus a 4-bit
    compensation code: any double written comes back within 2^-27 of itself
    (2^-28 but for the clamped last codes), values below the FP32 range within
    2^-153, and storing what was read changes nothing."""
    c = golden_cases.load("moo3")
    spec = c["spec"]
    sim = Simulation(spec, c["medium"], _abi.FP32)
    rng = np.random.default_rng(11)
    n = sim.count_cells(4)
    for kind, scale in ((0, 1e-12), (1, 1e2)):
        a = rng.standard_normal(n * 48) * scale * np.exp(rng.uniform(-20, 20, n * 48))
        a[:5] = [0.0, -0.0, scale, -scale, 2.0 ** -130]
        sim.set_state(4, kind, a)
        b = sim.state(4, kind)
        big = np.abs(a) >= 2.0 ** -126
        assert np.all(np.abs(b[big] - a[big]) <= 2.0 ** -27 * np.abs(a[big])), kind
        assert np.all(np.abs(b[~big] - a[~big]) <= 2.0 ** -153), kind
        sim.set_state(4, kind, b)
        np.testing.assert_array_equal(sim.state(4, kind), b)
    sim.close()


@pytest.mark.parametrize("name", ["hbn", "moo3"])
def test_compensated_vs_fp64_state(name):
    """The FP32 form with compensated FP32 r, v against the same form with
    FP64 r, v over the whole golden run: the storage adds no error visible
    beside the FP32 fields (monitor, full monitor and diagnostics agree to
    1e-5 relative L2)."""
    c = golden_cases.load(name)
    print("compensated vs fp64 storage", name, flush=True)
    res = {}
    for p in (_abi.FP32, _abi.FP32_RV64):
        rows, slices, full, d = run_case(c, p)
        res[p] = (rows, full, d)
    a, b = res[_abi.FP32], res[_abi.FP32_RV64]
    assert rel_l2(a[0][:, 1:], b[0][:, 1:]) <= 1e-5
    if a[1] is not None and len(a[1]):
        assert rel_l2(a[1], b[1]) <= 1e-5
    assert abs(a[2].max_r - b[2].max_r) <= 1e-5 * max(b[2].max_r, 1e-300)


@pytest.mark.parametrize("plan", ["fused", "split"])
@pytest.mark.parametrize("case", ["hbn", "tip_slab/8", "tip_slab/8-air"])
def test_graph_matches_stream_launches(monkeypatch, case, plan):
    """16-step graphs give the same bits as plain sequential launches, with
    and without an ADE material beside the lattice, under both Yee plans and
    in both forms: every field, lattice and ADE state and every monitor row
    (the monitor in the lattice slab, and away from it)."""
    monkeypatch.setenv("PHPFDTD_YEE", plan)
    if case == "hbn":
        c = golden_cases.load("hbn")
        spec, med = c["spec"], c["medium"]
    else:  # 1024x1024x640 layout at 1/8 scale: lattice slab + Drude ADE cone
        spec, med = workloads.tip_slab(8)
        spec = workloads.with_steps(spec, 40)
        if case.endswith("-air"):  # monitor in the planes the early launch updates
            g = spec["grid"]
            spec = dict(spec, monitor=dict(x=g["nx"] // 3, y=g["ny"] // 3, z=3, atom=0, dir=0))
    mon = spec.get("monitor")
    for prec in (_abi.FP64, _abi.FP32):
        res = []
        for ng in ("0", "1"):
            monkeypatch.setenv("PHPFDTD_NO_GRAPH", ng)
            sim = Simulation(spec, med, prec)
            if mon:
                sim.set_monitor(mon["x"], mon["y"], mon["z"], mon.get("atom", 0), mon.get("dir", 0))
            sim.run(37)  # 2 x 16-step graph + 5 x 1-step graph
            snap = [sim.field(q) for q in range(9)]
            for m, mat in enumerate(spec["materials"]):
                if sim.count_cells(m) == 0:
                    continue
                for kind in ([0, 1] if mat["natom"] else []) + ([2, 3] if mat["use_ade"] else []):
                    snap.append(sim.state(m, kind))
            if mon:
                snap.append(sim.monitor())
            res.append(snap)
            sim.close()
        assert len(res[0]) == len(res[1])
        for a, b in zip(*res):
            np.testing.assert_array_equal(a, b)


def test_errors_and_bounds():
    c = golden_cases.load("vacuum_cpml")
    spec = c["spec"]
    sim = Simulation(spec, c["medium"], _abi.FP64)
    with pytest.raises(_abi.PhpFdtdError) as ei:
        sim.run(spec["grid"]["tsteps"] + 1)
    assert ei.value.code == _abi.EINVAL
    with pytest.raises(_abi.PhpFdtdError):
        sim.slice("Ex", "q", 0)
    with pytest.raises(_abi.PhpFdtdError):
        sim.slice("Ex", "x", spec["grid"]["nx"])
    sim.run(3)
    assert sim.t == 3
    # the z slice of the top plane reads the zero halo, as the reference does
    top = sim.slice("Ez", "z", spec["grid"]["nz"] - 1)
    assert np.all(top == 0)
    sim.close()


def test_state_roundtrip():
    c = golden_cases.load("moo3")
    spec = c["spec"]
    sim = Simulation(spec, c["medium"], _abi.FP64)
    rng = np.random.default_rng(7)
    n = sim.count_cells(4)
    a = rng.standard_normal(n * 48)
    sim.set_state(4, 0, a)
    np.testing.assert_array_equal(sim.state(4, 0), a)
    f = rng.standard_normal(sim.shape)
    sim.set_field(5, f)
    np.testing.assert_array_equal(sim.field(5), f)
    sim.close()


def test_fp32_ade_state_exchange():
    """FP32 form ADE state (stored as P and J = P - P_old): P and P_old round
    to FP32 on the way in, come back as written in either write order."""
    c = golden_cases.load("ade")
    spec = c["spec"]
    m = [i for i, mt in enumerate(spec["materials"]) if mt["use_ade"] and i >= 4][0]
    sim = Simulation(spec, c["medium"], _abi.FP32)
    n = sim.count_cells(m)
    rng = np.random.default_rng(5)
    p, po = rng.standard_normal(3 * n) * 1e-11, rng.standard_normal(3 * n) * 1e-11
    for order in ((2, 3), (3, 2)):
        for kind in order:
            sim.set_state(m, kind, p if kind == 2 else po)
        np.testing.assert_allclose(sim.state(m, 2), p, rtol=2 ** -23)
        np.testing.assert_allclose(sim.state(m, 3), po, rtol=0, atol=4 * 2 ** -24 * np.abs(p).max())
    sim.close()


def test_fp32_ade_long_run():
    """A driven ADE slab over 10,000 steps: the FP32 form stays within 1e-4 of
    the FP64 form (the increment form of the ADE recurrence; the reference's
    two-level form in FP32 drifts to ~1e-2 here)."""
    d = dict(id=0, shape=(64, 64, 64), kind="lorentz", wn=1244.8, steps=10000)
    spec, med = workloads.serving_case(d)
    assert spec["materials"][4]["use_ade"] == 1
    rows = {}
    for p in (_abi.FP64, _abi.FP32):
        sim = Simulation(spec, med, p)
        m = spec["monitor"]
        sim.set_monitor(m["x"], m["y"], m["z"], m["atom"], m["dir"], ref_compat=True)
        sim.run(d["steps"])
        rows[p] = sim.monitor(d["steps"])
        sim.close()
    err = rel_l2(rows[_abi.FP32][:, 1:10], rows[_abi.FP64][:, 1:10])
    print("ADE 10000 steps FP32 vs FP64 rel L2", err, flush=True)
    assert err <= 1e-4


@pytest.mark.parametrize("case,rows", [("hbn", 4), ("aniso_loss", 4)] +
                         [(c, r) for c in ("ragged", "tip_slab/8") for r in (4, 6, 8, 12, 16)])
def test_fused_plan_matches_split(monkeypatch, case, rows):
    """The one-pass Yee kernel passes E and the new H between the rows of a
    work-group through LDS and recomputes the new H of the cells around the
    work-group (row j0-1, plane k0-1, cell x0-1) with the code that computes
    them in their owner, so it gives the same bits as writing H and reading it
    back (PHPFDTD_YEE=split): every field, lattice and ADE state, from a
    random state and then driven by the sources, in every form, for every
    work-group height (PHPFDTD_YEE_ROWS), including work-groups whose last
    rows lie past the grid."""
    if case == "ragged":
        spec, med = ragged_case((300, 13, 21), 3)
    elif case == "tip_slab/8":
        spec, med = workloads.tip_slab(8)
        spec = workloads.with_steps(spec, 40)
    else:
        c = golden_cases.load(case)
        spec, med = c["spec"], c["medium"]
    print("fused vs split", case, "rows", rows, flush=True)
    monkeypatch.setenv("PHPFDTD_YEE_ROWS", str(rows))
    for prec in (_abi.FP64, _abi.FP32, _abi.FP32_RV64):
        res = []
        for plan in ("fused", "split"):
            monkeypatch.setenv("PHPFDTD_YEE", plan)
            sim = Simulation(spec, med, prec)
            randomize(spec, med, 3, sim, orc.Oracle(spec, med) if case != "tip_slab/8" else _Sink())
            sim.run(21)
            snap = [sim.field(q) for q in range(9)]
            for m, mat in enumerate(spec["materials"]):
                if sim.count_cells(m) == 0:
                    continue
                for kind in ([0, 1] if mat["natom"] else []) + ([2, 3] if mat["use_ade"] else []):
                    snap.append(sim.state(m, kind))
            res.append(snap)
            sim.close()
        for a, b in zip(*res):
            np.testing.assert_array_equal(a, b)


class _Sink:
    """Stands in for the oracle where only the GPU state is randomised."""

    def set_field(self, *a):
        pass

    def set_state(self, *a):
        pass
