"""Batched stepping (phpfdtd_batch_*) and the continuous-batching server.

Bar: a simulation computes exactly the same numbers in a batch as alone -
every field component, the lattice / ADE state and the monitor trace are
bit-identical (tolerance 0) in both forms, across membership changes and
through the serving engine with admission queueing, under both Yee plans
of a batch (PHPFDTD_BATCH_YEE: the one-pass kernel, or the split pair).
"""
import numpy as np
import pytest

import golden_cases
from php_fdtd_amd import _abi, serve, workloads
from php_fdtd_amd.sim import Batch, Simulation

pytestmark = pytest.mark.gpu

CASES = ["hbn", "moo3", "ade", "aniso_loss"]


def snapshot(sim, spec):
    out = {c: sim.field(c) for c in ("Ex", "Ey", "Ez", "Dx", "Dy", "Dz", "Hx", "Hy", "Hz")}
    for m, mat in enumerate(spec["materials"]):
        if sim.count_cells(m) == 0:
            continue
        if mat.get("natom", 0):
            out["r%d" % m] = sim.state(m, 0)
            out["v%d" % m] = sim.state(m, 1)
        if mat.get("use_ade", 0):
            out["P%d" % m] = sim.state(m, 2)
            out["Po%d" % m] = sim.state(m, 3)
    out["monitor"] = sim.monitor()
    return out


def assert_same(a, b, label):
    assert a.keys() == b.keys(), label
    for k in a:
        assert a[k].shape == b[k].shape, (label, k)
        assert np.array_equal(a[k], b[k]), (label, k, np.abs(a[k] - b[k]).max())


def make(case, precision):
    spec = case["spec"]
    sim = Simulation(spec, case["medium"], precision)
    m = spec["monitor"]
    sim.set_monitor(m["x"], m["y"], m["z"], m["atom"], m["dir"], ref_compat=True)
    return sim


def alone(case, precision, steps):
    sim = make(case, precision)
    sim.run(steps)
    snap = snapshot(sim, case["spec"])
    sim.close()
    return snap


@pytest.mark.parametrize("plan", ["fused", "split"])
@pytest.mark.parametrize("precision", [_abi.FP64, _abi.FP32, _abi.FP32_RV64])
def test_batch_members_match_alone(monkeypatch, precision, plan):
    monkeypatch.setenv("PHPFDTD_BATCH_YEE", plan)
    cases = [golden_cases.load(n) for n in CASES]
    T = 100
    ref = [alone(c, precision, T) for c in cases]
    print("batch of", CASES, "precision", precision, flush=True)
    b = Batch(precision)
    sims = [make(c, precision) for c in cases]
    for s in sims:
        b.add(s)
    assert len(b) == len(cases)
    b.run(7)
    b.run(T - 7)
    b.sync()
    for c, s, r in zip(cases, sims, ref):
        assert_same(snapshot(s, c["spec"]), r, c["spec"]["name"])
        s.close()
    b.close()


@pytest.mark.parametrize("plan", ["fused", "split"])
def test_batch_membership_changes(monkeypatch, plan):
    monkeypatch.setenv("PHPFDTD_BATCH_YEE", plan)
    prec = _abi.FP32
    ca, cb = golden_cases.load("hbn"), golden_cases.load("moo3")
    ra, rb = alone(ca, prec, 60), alone(cb, prec, 60)
    a, bsim = make(ca, prec), make(cb, prec)
    a.run(10)                    # alone first (graph replays)
    b = Batch(prec)
    b.add(a)
    b.run(5)                     # batch of one
    b.add(bsim)                  # b joins at step 0, a at step 15
    b.run(20)
    b.remove(a)
    a.run(25)                    # a leaves and finishes alone: 10 + 5 + 20 + 25
    b.run(40)                    # b finishes in the batch: 20 + 40
    b.sync()
    assert_same(snapshot(a, ca["spec"]), ra, "a")
    assert_same(snapshot(bsim, cb["spec"]), rb, "b")
    a.close()
    bsim.close()
    b.close()


def test_batch_errors():
    c = golden_cases.load("hbn")
    s64 = make(c, _abi.FP64)
    b = Batch(_abi.FP32)
    with pytest.raises(_abi.PhpFdtdError) as e:
        b.add(s64)
    assert e.value.code == _abi.EINVAL
    s32 = make(c, _abi.FP32)
    b.add(s32)
    with pytest.raises(_abi.PhpFdtdError) as e:
        b.add(s32)
    assert e.value.code == _abi.ESTATE
    with pytest.raises(_abi.PhpFdtdError) as e:
        s32.run(1)
    assert e.value.code == _abi.ESTATE
    with pytest.raises(_abi.PhpFdtdError) as e:
        b.run(c["spec"]["grid"]["tsteps"] + 1)
    assert e.value.code == _abi.EINVAL
    with pytest.raises(_abi.PhpFdtdError):
        b.remove(s64)
    s32.close()                  # destroying a member removes it
    assert len(b) == 0
    b.run(3)                     # empty batch: nothing to do
    b.close()
    s64.close()


def small_stream(n, seed):
    descs = workloads.serving_stream(n, seed=seed, step_scale=0.008)
    reqs = []
    for d in descs:
        d = dict(d, shape=workloads.SERVE_GRIDS[d["id"] % 2])   # 64^3 and 96^3 only
        spec, med = workloads.serving_case(d)
        reqs.append(serve.Request(d["id"], spec, med, arrival=0.002 * d["id"]))
    return reqs


@pytest.mark.parametrize("precision", [_abi.FP32, _abi.FP64])
def test_serving_results_match_alone(precision):
    reqs = small_stream(6, seed=11)
    print("serving", [(r.spec["name"], r.steps) for r in reqs], flush=True)
    srv = serve.Server(precision, max_members=3, max_chunk=32)
    srv.serve(reqs)
    assert max(m for _, m, _ in srv.trace) <= 3   # slots bound the batch
    for r in reqs:
        assert r.error is None, r.error
        rows, _ = serve.run_alone(r, precision)
        assert r.rows.shape == (r.steps, 12)
        assert np.array_equal(r.rows, rows), r.spec["name"]
    st = serve.summarize(reqs)
    assert st["completed"] == len(reqs)


def test_serving_admission_by_footprint():
    reqs = []
    for i in range(3):  # same shape, material kind and steps: equal footprints
        d = dict(id=i, shape=(64, 64, 64), kind="lattice4", wn=1000.0 + 100 * i, steps=48, gap=0.0)
        spec, med = workloads.serving_case(d)
        reqs.append(serve.Request(i, spec, med))
    fps = [serve.footprint_bytes(r.spec, r.medium, _abi.FP32) for r in reqs]
    assert max(fps) == min(fps)
    fp = fps[0]
    srv = serve.Server(_abi.FP32, hbm_budget=1.5 * fp, max_members=8, max_chunk=32)
    srv.serve(reqs)
    assert max(m for _, m, _ in srv.trace) == 1   # only one request fits at a time
    for r in reqs:
        rows, _ = serve.run_alone(r, _abi.FP32)
        assert np.array_equal(r.rows, rows)


@pytest.mark.timeout(300)
def test_serving_every_grid_and_material():
    """One request of every serving grid (64^3 .. 256x256x160) with every
    material kind (Lorentz ADE, 4-atom and 16-atom lattices): 15 requests of
    different sizes and step counts served together in the FP32 form, each
    result bit-identical to the same request run alone.  Step counts of 150 to
    230 let the source's field reach every monitor (up to 36 cells away; the
    stencil carries it one cell per step)."""
    reqs = []
    rng = np.random.default_rng(8)
    for gi, shape in enumerate(workloads.SERVE_GRIDS):
        for ki, kind in enumerate(workloads.SERVE_KINDS):
            rid = 3 * gi + ki
            d = dict(id=rid, shape=shape, kind=kind, wn=float(rng.uniform(600.0, 1600.0)),
                     steps=int(rng.integers(150, 230)), gap=0.0)
            spec, med = workloads.serving_case(d)
            reqs.append(serve.Request(rid, spec, med, arrival=0.001 * rid))
    print("serving all grids", [(r.spec["name"], r.steps) for r in reqs], flush=True)
    srv = serve.Server(_abi.FP32, max_members=6, max_chunk=16)
    srv.serve(reqs)
    assert max(m for _, m, _ in srv.trace) > 1   # requests overlapped on the device
    for r in reqs:
        assert r.error is None, r.error
        rows, _ = serve.run_alone(r, _abi.FP32)
        assert r.rows.shape == (r.steps, 12)
        assert np.abs(r.rows[:, 1:10]).max() > 0, r.spec["name"]
        assert np.array_equal(r.rows, rows), r.spec["name"]
