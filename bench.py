#!/usr/bin/env python3
"""Benchmark of the per-timestep update on one MI355X.

Workload: the 1024x1024x640 configuration (10 nm cells, 20-cell CPML, 48-cell
16-atom lattice slab, Drude-metal cone tip, Ricker source), synthetic seeded
materials, all state resident in HBM before timing.  A "step" is one full time
step (H, D/E, ADE, lattice, monitor epilogue) of the whole domain.

Reported: Gcell-updates/s of the FP32 precision form (value) and of the FP64
plain form, the HBM roofline of the dominant kernel (HIP events on the
library's stream, algorithmic bytes), the whole-step byte model, the FP32 vs
FP64 monitor error over the timed run, and the reference solver on the host
cores on a bounded sample of the same workload (cpu_baseline).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload tip_slab]
"""
import argparse
import json
import math
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
PROFILE = os.path.join(ROOT, "profiles", "step_kernels_tip_slab_1024x1024x640.md")
MFMA_FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32 matrix peak

METRIC = "Gcell-updates/s + HBM roofline fraction; solves/s, p50/p99, FP32 vs FP64"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def copy_ceiling(gib=4.0, iters=5):
    """Measured streaming ceilings of this device (php_fdtd_amd/bwprobe): the
    non-temporal copy and the two-buffer in-place update, GB/s."""
    exe = os.path.join(ROOT, "php_fdtd_amd", "bwprobe")
    out = {}
    if not os.path.exists(exe):
        return None
    for v in ("nt", "ip"):
        p = subprocess.run([exe, str(gib), str(iters), v], capture_output=True, text=True, timeout=120)
        for ln in p.stdout.splitlines():
            if ln.startswith("{"):
                out[v] = json.loads(ln)["gbps"]
    return {"copy_nt_gbs": out.get("nt"), "inplace_2buf_gbs": out.get("ip"), "buffer_gib": gib,
            "source": "php_fdtd_amd/bwprobe %g %d nt|ip" % (gib, iters)}


def profile_traffic(symbol):
    """HBM bytes per launch of a kernel from the committed rocprofv3 PMC summary
    (FETCH_SIZE + WRITE_SIZE per dispatch, scaled by the copy-probe factors),
    or None when the summary has no row for it."""
    if not os.path.exists(PROFILE):
        return None
    text = open(PROFILE).read()
    m = re.search(r"over the (\d+) steps of each form", text)
    if not m:
        return None
    steps = int(m.group(1))  # steps each form ran in the counter runs
    in_pmc = False
    for ln in text.splitlines():
        if ln.startswith("## "):
            in_pmc = ln.startswith("## HBM traffic")
        cells = [c.strip() for c in ln.strip().strip("|").split("|")]
        if in_pmc and len(cells) == 7 and cells[0] == "`%s`" % symbol:
            try:
                launches, traffic = int(cells[1]), float(cells[4])
            except ValueError:
                return None
            per_step = max(1, launches // steps)
            return traffic * 1e9 / per_step
    return None


def kernel_symbol(name, form):
    """Device symbol of a plan kernel name of the given form."""
    real = "double" if form == "fp64" else "float"
    if name.startswith("ion_update"):
        n = int(name.split(",N")[1].rstrip("]"))
        if form == "fp64":
            return "phpfdtd::ion_update_f64<%d>" % n
        return "phpfdtd::ion_update_mfma<%d, %s>" % (n, "true" if form == "fp32" else "false")
    base = name.split("[")[0]
    if base in ("h_update", "de_update"):  # default Yee tiles: 1 row per wavefront in FP32, 2 in FP64
        return "phpfdtd::%s<%s, %d>" % (base, real, 2 if form == "fp64" else 1)
    return "phpfdtd::%s<%s>" % (base, real)


def run_form(spec, med, precision, warmup, steps, prof_steps):
    from php_fdtd_amd import _abi
    from php_fdtd_amd.sim import Simulation

    t0 = time.time()
    sim = Simulation(spec, med, precision)
    mon = spec["monitor"]
    sim.set_monitor(mon["x"], mon["y"], mon["z"], mon["atom"], mon["dir"], ref_compat=True)
    log("  context ready in %.1f s" % (time.time() - t0))
    sim.run(warmup)
    sim.sync(diag=False)
    wall0 = time.perf_counter()
    ms = sim.time_steps(steps)  # HIP events on the library stream, synced on both sides
    wall = time.perf_counter() - wall0
    prof = sim.profile(prof_steps)
    # the same steps with output sampling: a z slice and the monitor rows read
    # back every `interval` steps, as a reference run with output_interval does
    interval, chunks = 10, max(1, steps // 10)
    sim.sync(diag=False)
    t0 = time.perf_counter()
    for _ in range(chunks):
        sim.run(interval)
        sim.slice("Ez", "z", spec["sources"][0]["z"])
        sim.monitor()
    sim.sync(diag=False)
    sampled_ms = (time.perf_counter() - t0) * 1e3 / (chunks * interval)
    rows = sim.monitor()
    d = sim.sync()
    sim.close()
    return dict(ms=ms, wall_ms=wall * 1e3 / steps, prof=prof, rows=rows, max_e=d.max_e, max_r=d.max_r,
                sampled_ms=sampled_ms)


def central(a):
    """Central 25%-75% region of a plane, strict bounds (frequency_analysis.c:300-303)."""
    n0, n1 = a.shape
    i = np.arange(n0)[:, None]
    j = np.arange(n1)[None, :]
    return a[(i > n0 * 0.25) & (i < n0 * 0.75) & (j > n1 * 0.25) & (j < n1 * 0.75)]


def accuracy(name, interval=10, steps=None):
    """FP32 precision form vs FP64 plain form over a full-length run: monitor
    trace relative L2 (fields; lattice R, V) and the near-field single-bin DFT
    at the source frequency of the z plane one cell above the slab, central
    25-75% region (frequency_analysis.c semantics)."""
    from php_fdtd_amd import _abi, workloads
    from php_fdtd_amd.sim import Simulation

    spec, med = workloads.WORKLOADS[name]()
    if steps:
        spec = workloads.with_steps(spec, steps)
    g = spec["grid"]
    T = g["tsteps"]
    N = T // interval
    zs = np.nonzero((med != 0).any(axis=(0, 1)))[0]
    zplane = int(zs.max()) + 1
    f = spec["sources"][-1]["frequency"]
    mon = spec["monitor"]
    res = {}
    names = {_abi.FP32: "fp32", _abi.FP32_RV64: "fp32 (r,v fp64)", _abi.FP64: "fp64"}
    for prec in (_abi.FP32, _abi.FP32_RV64, _abi.FP64):
        sim = Simulation(spec, med, prec)
        sim.set_monitor(mon["x"], mon["y"], mon["z"], mon["atom"], mon["dir"], ref_compat=True)
        k = sim.set_dft("Ez", "z", zplane, interval, N, f, ref_compat=False)
        t0 = time.perf_counter()
        sim.run(T)
        sim.sync(diag=False)
        wall = time.perf_counter() - t0
        rows = sim.monitor(T)
        x, ns = sim.dft()
        d = sim.sync()
        sim.close()
        res[prec] = dict(rows=rows, x=x, k=k, wall=wall, max_r=d.max_r)
        log("  accuracy %s %s: %d steps in %.2f s" % (name, names[prec], T, wall))
    a, b, c = res[_abi.FP32], res[_abi.FP64], res[_abi.FP32_RV64]
    x32, x64 = central(a["x"]), central(b["x"])
    amp32, amp64 = np.abs(x32), np.abs(x64)
    ph = np.angle(x32 * np.conj(x64))
    strong = amp64 >= 0.1 * amp64.max()
    fs = 1.0 / (interval * g["dt"])

    def rl2(u, v):
        n = np.linalg.norm(v)
        return float(np.linalg.norm(u - v) / n) if n else None

    return {
        "workload": "%s%s: %dx%dx%d cells, %d steps, FP32 form vs FP64 plain form" % (
            name, "" if steps else " full length", g["nx"], g["ny"], g["nz"], T),
        "monitor_rel_l2": rl2(a["rows"][:, 1:10], b["rows"][:, 1:10]),
        "monitor_rv_rel_l2": rl2(a["rows"][:, 10:12], b["rows"][:, 10:12]),
        "rv_fp64_storage": {"monitor_rel_l2": rl2(c["rows"][:, 1:10], b["rows"][:, 1:10]),
                            "monitor_rv_rel_l2": rl2(c["rows"][:, 10:12], b["rows"][:, 10:12]),
                            "vs_compensated_monitor_rel_l2": rl2(a["rows"][:, 1:12], c["rows"][:, 1:12])},
        "nearfield": {"component": "Ez", "plane": "z=%d" % zplane, "samples": N, "interval": interval,
                      "bin": b["k"], "bin_hz": b["k"] * fs / N, "source_hz": f,
                      "amp_rel_l2": rl2(amp32, amp64),
                      "phase_rms_rad": float(np.sqrt(np.sum(amp64 ** 2 * ph ** 2) / np.sum(amp64 ** 2))),
                      "phase_max_rad_strong": float(np.abs(ph[strong]).max())},
        "solve_s": {"fp32": round(a["wall"], 3), "fp32_rv64": round(c["wall"], 3), "fp64": round(b["wall"], 3)},
    }


def serving(n, step_scale, seed=4, load=0.8):
    """Serving mix (grids 64^3..256x256x160, Lorentz / 4-atom / 16-atom
    materials, source U(600, 1600) cm^-1, steps U(5000, 20000) x step_scale,
    seeded): the FP64 plain form one request at a time (end to end: build,
    run, read, destroy), then the FP32 form through the continuous-batching
    server with every request queued at t=0 (capacity), then with Poisson
    arrivals at `load` x that capacity (time-to-result)."""
    from php_fdtd_amd import _abi, serve, workloads

    descs = workloads.serving_stream(n, seed=seed, step_scale=step_scale)
    cases = [workloads.serving_case(d) for d in descs]

    def requests(arrivals):
        return [serve.Request(d["id"], sp, med, arrival=a) for d, (sp, med), a in zip(descs, cases, arrivals)]

    work = float(sum(med.size * sp["grid"]["tsteps"] for sp, med in cases))
    seq_rows, seq_t = [], 0.0
    last = time.perf_counter()
    for q, r in enumerate(requests([0.0] * n)):
        rows, t = serve.run_alone(r, _abi.FP64)
        seq_rows.append(rows)
        seq_t += t
        if time.perf_counter() - last >= 30:
            last = time.perf_counter()
            log("  serving: FP64 one at a time, %d of %d done, %.0f s" % (q + 1, n, seq_t))
    log("  serving: FP64 one at a time %.2f s" % seq_t)
    srv = serve.Server(_abi.FP32, progress=30)
    sat = srv.serve(requests([0.0] * n))
    st_sat = serve.summarize(sat)
    members = [m for _, m, _ in srv.trace]
    log("  serving: FP32 batched, all queued: %s" % st_sat)
    rate = load * st_sat["solves_per_s"]
    arr = np.cumsum([d["gap"] for d in descs]) / rate
    arr -= arr[0]
    poi = serve.Server(_abi.FP32, progress=30).serve(requests(list(arr)))
    st_poi = serve.summarize(poi)
    log("  serving: FP32 batched, Poisson %.2f/s: %s" % (rate, st_poi))
    errs, ids = [], []
    by_id = {d["id"]: rows for d, rows in zip(descs, seq_rows)}
    for r in sat:
        ref = by_id[r.id][:, 1:10]
        den = np.linalg.norm(ref)
        errs.append(float(np.linalg.norm(r.rows[:, 1:10] - ref) / den) if den else 0.0)
        ids.append(r.id)
    worst = {d["id"]: d for d in descs}[ids[int(np.argmax(errs))]]
    seq_sps = n / seq_t
    return {
        "workload": "serving mix, %d requests (seed %d), steps U(5000, 20000) x %g, FP32 form batched vs FP64 plain "
                    "one at a time" % (n, seed, step_scale),
        "cell_updates": work,
        "fp64_one_at_a_time": {"solves_per_s": round(seq_sps, 4), "gcell_updates_per_s": round(work / seq_t / 1e9, 3),
                               "total_s": round(seq_t, 3)},
        "fp32_batched_saturated": dict(st_sat, mean_members=round(float(np.mean(members)), 2),
                                       max_members=int(max(members))),
        "fp32_batched_poisson": dict(st_poi, offered_per_s=round(rate, 4), load=load),
        "solves_ratio_vs_fp64_one_at_a_time": round(st_sat["solves_per_s"] / seq_sps, 3),
        "per_request_monitor_rel_l2_vs_fp64": {
            "median": float(np.median(errs)), "max": float(np.max(errs)),
            "max_request": {k: worst[k] for k in ("id", "kind", "shape", "wn", "steps")}},
    }


def cpu_baseline(budget_s=25.0):
    """Reference fdtd.c (oracle/_ref/fdtd.x, -Ofast -flto -funroll-loops) under
    mpirun on the host cores, on the same workload scaled down 4x per axis
    (256x256x160 cells, same slab/cone/CPML layout), for a step count sized to
    ~budget_s of CPU time.  Falls back to the single-thread C oracle
    (oracle/liborc.so) when the reference binary is absent."""
    from php_fdtd_amd import workloads

    ref = os.path.join(ROOT, "oracle", "_ref", "fdtd.x")
    mpirun = "/opt/conda/bin/mpirun"
    try:
        ncpu = len(os.sched_getaffinity(0))
    except Exception:
        ncpu = os.cpu_count() or 1
    cores = max(1, min(16, ncpu))
    spec, med = workloads.tip_slab(scale=4)
    cells = med.size
    if os.path.exists(ref) and os.path.exists(mpirun):
        # probe 2 steps to size the sample
        tmp = tempfile.mkdtemp(prefix="phpfdtd_cpu_")
        try:
            def run(nsteps):
                sp = workloads.with_steps(spec, nsteps)
                cfg = workloads.write_reference_case(tmp, sp, med)
                env = dict(os.environ, OMP_NUM_THREADS="1")
                p = subprocess.run([mpirun, "-np", str(cores), ref, cfg], cwd=tmp, capture_output=True, text=True,
                                   timeout=600, env=env)
                out = p.stdout
                for ln in out.splitlines():
                    if "Simulation completed in" in ln:
                        return float(ln.split("in")[1].split()[0])
                raise RuntimeError("reference run failed rc=%d: %s" % (p.returncode, (out + p.stderr)[-800:]))
            t2 = run(2)
            n = int(max(3, min(400, budget_s / max(t2 / 2, 1e-3))))
            t = run(n)
            return dict(value=cells * n / t / 1e9, unit="Gcell-updates/s", cores=cores, kind="reference",
                        sample="reference fdtd.c (-Ofast -flto -funroll-loops) under mpirun -np %d, "
                               "256x256x160 (the 1024x1024x640 workload scaled 1/4 per axis), %d steps, loop time %.2f s" % (cores, n, t))
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
    from oracle import orc

    spec, med = workloads.tip_slab(scale=8)
    spec = workloads.with_steps(spec, 1000)
    o = orc.Oracle(spec, med)
    t0 = time.time()
    n = 0
    while time.time() - t0 < budget_s / 2 and n < 1000:
        o.step(1)
        n += 1
    t = time.time() - t0
    return dict(value=med.size * n / t / 1e9, unit="Gcell-updates/s", cores=1, kind="port",
                sample="C oracle (FP64 restatement), 1 thread, 128x128x80 (the 1024x1024x640 workload scaled 1/8), %d steps" % n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", "--config", dest="workload", default="tip_slab",
                    help="tip_slab, lattice16_slab, lattice4_slab or lorentz_slab")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp32,fp64", help="comma list of fp32, fp32rv64, fp64")
    ap.add_argument("--accuracy", default="lattice4_slab",
                    help="workload of the FP32-vs-FP64 accuracy run ('' = skip)")
    ap.add_argument("--accuracy-steps", type=int, default=0, help="steps of the accuracy run (0 = the workload's own)")
    ap.add_argument("--serving", type=int, default=64, help="requests of the serving mix (0 = skip)")
    ap.add_argument("--serving-step-scale", type=float, default=0.05)
    ap.add_argument("--no-copy-probe", dest="copy_probe", action="store_false",
                    help="skip the measured copy-bandwidth ceilings")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving block and print it as the JSON line (for profiling the batched path)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; the domain is resident on a single MI355X and no "
              "multi-GPU path is built (--gpus %d not supported)" % args.gpus, file=sys.stderr)
        return 2

    from php_fdtd_amd import _abi, workloads

    if args.serving_only:
        print(json.dumps({"serving": serving(args.serving, args.serving_step_scale)}))
        return 0
    prof_steps = 3
    t0 = time.time()
    spec, med = workloads.WORKLOADS[args.workload]()
    spec = workloads.with_steps(spec, args.warmup + args.steps + prof_steps + max(1, args.steps // 10) * 10)
    ceiling = copy_ceiling() if args.copy_probe else None
    if ceiling:
        log("copy ceilings: %s" % ceiling)
    log("workload %s: %s cells, built in %.1f s" % (spec["name"], med.size, time.time() - t0))
    cells = float(med.size)
    res = {}
    precs = {"fp32": _abi.FP32, "fp32rv64": _abi.FP32_RV64, "fp64": _abi.FP64}
    rv_bytes = {"fp32": 5, "fp32rv64": 8, "fp64": 8}
    for form in args.forms.split(","):
        prec = precs[form]
        log("form %s" % form)
        res[form] = run_form(spec, med, prec, args.warmup, args.steps, prof_steps)
        log("  %.3f ms/step (events), %.3f ms/step (wall), %.3f ms/step with a slice + monitor read every 10 steps" % (
            res[form]["ms"], res[form]["wall_ms"], res[form]["sampled_ms"]))
        for k in res[form]["prof"]:
            log("    %-24s %9.3f ms  %8.1f GB/s" % (k["name"], k["ms"], k["bytes"] / k["ms"] / 1e6 if k["ms"] else 0))

    head = [f for f in ("fp32", "fp32rv64", "fp64") if f in res][0]
    r = res[head]
    word = 8 if head == "fp64" else 4
    gcell = cells / (r["ms"] * 1e-3) / 1e9
    dom = max(r["prof"], key=lambda k: k["ms"])
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
    dom_sym = kernel_symbol(dom["name"], head)
    traffic = profile_traffic(dom_sym)
    sb = workloads.step_bytes(spec, med, word, rv_bytes[head])
    rv_desc = {"fp32": "compensated f32 (f64 integration)", "fp32rv64": "f64", "fp64": "f64"}[head]
    out = {
        "metric": METRIC,
        "value": round(gcell, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if head == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": "%s: 16-atom lattice slab + Drude cone, %s form (fields/CPML/ADE %s, lattice r,v "
                               "%s)" % (spec["name"], head.upper(), "f64" if head == "fp64" else "f32", rv_desc),
                   "cells": int(cells), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom["name"], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic) if traffic else None,
                     "algorithmic_bytes": round(dom["bytes"]), "symbol": dom_sym,
                     "traffic_source": "profiles/step_kernels_tip_slab_1024x1024x640.md: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per "
                                       "dispatch, copy-probe calibrated" if traffic else None},
        "step_roofline": {"bytes_per_step": sb, "achieved": round(sb / (r["ms"] * 1e-3) / 1e9, 1),
                          "frac": round(sb / (r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "model": "C(18w+1) + 8w*CPML incidences + 9w*ADE cells + 12Nb*lattice cells, b = %d "
                                   "bytes per stored r or v value" % rv_bytes[head]},
        "kernels_ms": {k["name"]: round(k["ms"], 4) for k in r["prof"]},
        "ms_per_step_with_sampling": round(r["sampled_ms"], 4),
    }
    if ceiling:
        out["measured_ceiling"] = dict(ceiling, dominant_frac_of_inplace=round(
            achieved / ceiling["inplace_2buf_gbs"], 4) if ceiling.get("inplace_2buf_gbs") else None,
            step_frac_of_copy=round(sb / (r["ms"] * 1e-3) / 1e9 / ceiling["copy_nt_gbs"], 4)
            if ceiling.get("copy_nt_gbs") else None)
    # MFMA work of the DMM . r step (precision form): 2 (3N)^2 flops per lattice cell
    if head != "fp64":
        counts, _ = workloads.cell_counts(spec, med)
        for m, mat in enumerate(spec["materials"]):
            n = mat.get("natom", 0)
            if n and counts[m]:
                name = "ion_update[m%d,N%d]" % (m, n)
                kt = [k for k in r["prof"] if k["name"] == name]
                if kt:
                    fl = 2.0 * (3 * n) ** 2 * counts[m]
                    tf = fl / (kt[0]["ms"] * 1e-3) / 1e12
                    out["mfma"] = {"kernel": name, "op": "v_mfma_f32_16x16x4f32, -DMM . r", "flops": fl,
                                   "achieved": round(tf, 2), "peak": MFMA_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                                   "frac": round(tf / MFMA_FP32_PEAK_TFLOPS, 4)}
    if "fp64" in res and head != "fp64":
        r64 = res["fp64"]
        out["fp64"] = {"value": round(cells / (r64["ms"] * 1e-3) / 1e9, 4), "ms_per_step": round(r64["ms"], 4),
                       "step_frac": round(workloads.step_bytes(spec, med, 8) / (r64["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                       "kernels_ms": {k["name"]: round(k["ms"], 4) for k in r64["prof"]}}
        out["speedup_fp32_vs_fp64"] = round(r64["ms"] / r["ms"], 3)
        if "fp32rv64" in res and head != "fp32rv64":
            rr = res["fp32rv64"]
            out["fp32_rv64"] = {"value": round(cells / (rr["ms"] * 1e-3) / 1e9, 4), "ms_per_step": round(rr["ms"], 4),
                                "kernels_ms": {k["name"]: round(k["ms"], 4) for k in rr["prof"]}}
        a, b = r["rows"], r64["rows"]
        n = min(len(a), len(b))
        if n:
            num = np.linalg.norm(a[:n, 1:10] - b[:n, 1:10])
            den = np.linalg.norm(b[:n, 1:10])
            out["fp32_monitor_rel_l2_vs_fp64"] = float(num / den) if den else None
    if args.accuracy:
        out["accuracy"] = accuracy(args.accuracy, steps=args.accuracy_steps or None)
    if args.serving:
        out["serving"] = serving(args.serving, args.serving_step_scale)
    if not args.no_cpu:
        try:
            t0 = time.time()
            out["cpu_baseline"] = cpu_baseline()
            log("cpu baseline %.1f s: %s" % (time.time() - t0, out["cpu_baseline"]))
        except Exception as e:  # reported, not fatal
            out["cpu_baseline"] = {"value": None, "error": str(e)[:300]}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
