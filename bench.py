#!/usr/bin/env python3
"""Benchmark of the per-timestep update on one MI355X.

Workload: the 1024x1024x640 configuration (10 nm cells, 20-cell CPML, 48-cell
16-atom lattice slab, Drude-metal cone tip, Ricker source), synthetic seeded
materials, all state resident in HBM before timing.  A "step" is one full time
step (H, D/E, ADE, lattice, monitor epilogue) of the whole domain.

Reported: Gcell-updates/s of the FP32 precision form (value) and of the FP64
plain form, the HBM roofline of the dominant kernel (HIP events on the
library's stream, algorithmic bytes), the whole-step byte model, the time of
one phpfdtd_sync with diagnostics, the FP32 vs FP64 error on the same
workload over 5000 steps (monitor trace and near-field DFT at the source
frequency), the serving mix, and the reference solver on the host cores
(cpu_baseline: the workload's own grid when host RAM holds it).

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
PROFILE_WORKLOAD = "tip_slab"  # the workload whose PMC traffic the profile holds
GEOMETRY = {"lorentz_slab": "isotropic Lorentz slab", "lattice4_slab": "4-atom lattice slab",
            "lattice16_slab": "16-atom lattice slab", "tip_slab": "16-atom lattice slab + Drude cone"}
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
    if base == "yee_update":  # one-pass kernel, instantiated per rows per work-group
        return "phpfdtd::yee_update<%s, %d>" % (real, int(name.split("[r")[1].rstrip("]")))
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
    rows = [sim.monitor()]  # every row since the start (the ring drains on read)
    sim.sync(diag=False)
    t0 = time.perf_counter()
    for _ in range(chunks):
        sim.run(interval)
        sim.slice("Ez", "z", spec["sources"][0]["z"])
        rows.append(sim.monitor())
    sim.sync(diag=False)
    sampled_ms = (time.perf_counter() - t0) * 1e3 / (chunks * interval)
    rows.append(sim.monitor())
    # phpfdtd_sync with diagnostics (max |E|, max |Ex|, max |r| and its argmax)
    t0 = time.perf_counter()
    for _ in range(3):
        d = sim.sync()
    sync_ms = (time.perf_counter() - t0) * 1e3 / 3
    sim.close()
    return dict(ms=ms, wall_ms=wall * 1e3 / steps, prof=prof, rows=np.concatenate(rows), max_e=d.max_e,
                max_r=d.max_r, sampled_ms=sampled_ms, sync_ms=sync_ms)


def central(a):
    """Central 25%-75% region of a plane, strict bounds (frequency_analysis.c:300-303)."""
    n0, n1 = a.shape
    i = np.arange(n0)[:, None]
    j = np.arange(n1)[None, :]
    return a[(i > n0 * 0.25) & (i < n0 * 0.75) & (j > n1 * 0.25) & (j < n1 * 0.75)]


def accuracy(name, interval=10, steps=None, rv64=True):
    """FP32 precision form vs FP64 plain form over a run of the workload:
    monitor trace relative L2 (fields; lattice R, V) and the near-field
    single-bin DFT at the source frequency of the z plane one cell above the
    slab (material 4), central 25-75% region (frequency_analysis.c
    semantics).  rv64 adds the FP32 form with FP64 lattice r, v."""
    from php_fdtd_amd import _abi, workloads
    from php_fdtd_amd.sim import Simulation

    spec, med = workloads.WORKLOADS[name]()
    if steps:
        spec = workloads.with_steps(spec, steps)
    g = spec["grid"]
    T = g["tsteps"]
    N = T // interval
    zs = np.nonzero((med == 4).any(axis=(0, 1)))[0]
    zplane = int(zs.max()) + 1
    f = spec["sources"][-1]["frequency"]
    mon = spec["monitor"]
    res = {}
    names = {_abi.FP32: "fp32", _abi.FP32_RV64: "fp32 (r,v fp64)", _abi.FP64: "fp64"}
    for prec in ((_abi.FP32, _abi.FP32_RV64, _abi.FP64) if rv64 else (_abi.FP32, _abi.FP64)):
        sim = Simulation(spec, med, prec)
        sim.set_monitor(mon["x"], mon["y"], mon["z"], mon["atom"], mon["dir"], ref_compat=True)
        k = sim.set_dft("Ez", "z", zplane, interval, N, f, ref_compat=False)
        t0 = time.perf_counter()
        done = 0
        while done < T:  # in chunks, so the run reports its progress
            n = min(1000, T - done)
            sim.run(n)
            sim.sync(diag=False)
            done += n
            if done < T:
                log("    %s: %d of %d steps, %.1f s" % (names[prec], done, T, time.perf_counter() - t0))
        wall = time.perf_counter() - t0
        rows = sim.monitor(T)
        x, ns = sim.dft()
        d = sim.sync()
        sim.close()
        res[prec] = dict(rows=rows, x=x, k=k, wall=wall, max_r=d.max_r)
        log("  accuracy %s %s: %d steps in %.2f s" % (name, names[prec], T, wall))
    a, b, c = res[_abi.FP32], res[_abi.FP64], res.get(_abi.FP32_RV64)
    x32, x64 = central(a["x"]), central(b["x"])
    amp32, amp64 = np.abs(x32), np.abs(x64)
    ph = np.angle(x32 * np.conj(x64))
    strong = amp64 >= 0.1 * amp64.max()
    fs = 1.0 / (interval * g["dt"])

    def rl2(u, v):
        n = np.linalg.norm(v)
        return float(np.linalg.norm(u - v) / n) if n else None

    src = spec["sources"][0]
    out = {
        "workload": "%s%s: %dx%dx%d cells, %d steps, FP32 form vs FP64 plain form" % (
            name, "" if steps else " full length", g["nx"], g["ny"], g["nz"], T),
        "monitor": "x=%d y=%d z=%d atom %d dir %d (fields at z+1)" % (mon["x"], mon["y"], mon["z"], mon["atom"],
                                                                     mon["dir"]),
        "monitor_rel_l2": rl2(a["rows"][:, 1:10], b["rows"][:, 1:10]),
        "monitor_rv_rel_l2": rl2(a["rows"][:, 10:12], b["rows"][:, 10:12]),
        # the first steps only: the source has injected a negligible fraction
        # of its pulse (a Ricker wavelet peaks at t0), so the monitor holds the
        # far tail of the numerical field, where FP32 rounding dominates
        "monitor_rel_l2_first_70_steps": rl2(a["rows"][:70, 1:10], b["rows"][:70, 1:10]),
        "source_peak_step": round(src["t0"] / g["dt"], 1) if src["type"] in (0, 2, 3) else None,
        "nearfield": {"component": "Ez", "plane": "z=%d" % zplane, "samples": N, "interval": interval,
                      "bin": b["k"], "bin_hz": b["k"] * fs / N, "source_hz": f,
                      "amp_rel_l2": rl2(amp32, amp64),
                      # (None where the plane holds no field at the bin)
                      "phase_rms_rad": float(np.sqrt(np.sum(amp64 ** 2 * ph ** 2) / np.sum(amp64 ** 2)))
                      if amp64.max() > 0 else None,
                      "phase_max_rad_strong": float(np.abs(ph[strong]).max()) if amp64.max() > 0 else None},
        "solve_s": {"fp32": round(a["wall"], 3), "fp64": round(b["wall"], 3)},
    }
    if c:
        out["rv_fp64_storage"] = {"monitor_rel_l2": rl2(c["rows"][:, 1:10], b["rows"][:, 1:10]),
                                  "monitor_rv_rel_l2": rl2(c["rows"][:, 10:12], b["rows"][:, 10:12]),
                                  "vs_compensated_monitor_rel_l2": rl2(a["rows"][:, 1:12], c["rows"][:, 1:12])}
        out["solve_s"]["fp32_rv64"] = round(c["wall"], 3)
    return out


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


def serving_batched(n, step_scale, seed=4):
    """The FP32 batched half of the serving block alone: every request of the
    serving mix queued at t=0 and stepped through phpfdtd_batch_run (batched
    launches only, no per-context graphs), with the step-model bytes of all
    requests, so a kernel trace of this run prices the batched kernels."""
    from php_fdtd_amd import _abi, serve, workloads

    descs = workloads.serving_stream(n, seed=seed, step_scale=step_scale)
    cases = [workloads.serving_case(d) for d in descs]
    reqs = [serve.Request(d["id"], sp, med, arrival=0.0) for d, (sp, med) in zip(descs, cases)]
    srv = serve.Server(_abi.FP32, progress=30)
    st = serve.summarize(srv.serve(reqs))
    members = [m for _, m, _ in srv.trace]
    return dict(st, workload="serving mix, %d requests (seed %d), steps U(5000, 20000) x %g, FP32 form batched, "
                             "all queued at t=0" % (n, seed, step_scale),
                cell_updates=float(sum(med.size * sp["grid"]["tsteps"] for sp, med in cases)),
                step_model_bytes=float(sum(sp["grid"]["tsteps"] * workloads.step_bytes(sp, med, 4, 4.5)
                                           for sp, med in cases)),
                mean_members=round(float(np.mean(members)), 2), max_members=int(max(members)))


def host_info():
    """CPU model, physical cores, the CPU share of this process, host RAM and
    the MPI implementation, for the CPU baseline record."""
    info = {"logical_cpus": os.cpu_count()}
    try:
        info["cpu_share"] = len(os.sched_getaffinity(0))
    except Exception:
        info["cpu_share"] = os.cpu_count()
    try:
        cores, model = set(), None
        phys = core = None
        for ln in open("/proc/cpuinfo"):
            k, _, v = ln.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model is None:
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None:
                cores.add((phys, core))
                phys = core = None
        if phys is not None:
            cores.add((phys, core))
        info["cpu_model"] = model
        info["physical_cores"] = len(cores) or None
    except Exception:
        pass
    try:
        mem = {ln.split(":")[0]: int(ln.split()[1]) for ln in open("/proc/meminfo")}
        info["mem_total_gb"] = round(mem["MemTotal"] / 1e6, 1)
        info["mem_available_gb"] = round(mem["MemAvailable"] / 1e6, 1)
    except Exception:
        pass
    try:
        v = subprocess.run(["/opt/conda/bin/mpirun", "--version"], capture_output=True, text=True, timeout=30)
        info["mpi"] = " ".join((v.stdout + v.stderr).split()[:6])
    except Exception:
        pass
    return info


def reference_host_bytes(spec, med):
    """Host bytes the reference solver holds for a workload, from its data
    layout: nine FP64 field arrays plus |E|, int32 medium ids and per-cell r/v
    pointers (fdtd.c:261-284, 2959-3049), one heap array of 3N doubles for each
    of r and v per lattice cell, and the CPML slabs (z slabs full-volume on the
    ranks holding them), scaled by 1.3: the reference's own [MEM] totals were
    1.26x (256x256x160) and 1.42x (128x128x80) of this layout count for the
    1024x1024x640 workload scaled down, the difference being allocator
    overhead and buffers."""
    from php_fdtd_amd import workloads

    counts, pml = workloads.cell_counts(spec, med)
    b = med.size * (10 * 8 + 4 + 16) + 8 * 8 * pml
    for m, mat in enumerate(spec["materials"]):
        if mat.get("natom", 0) and counts[m]:
            b += counts[m] * 2 * (24 * mat["natom"] + 16)
    return 1.3 * float(b)


def run_reference(spec, med, nsteps, ranks, timeout, log_path=None):
    """Reference fdtd.c (oracle/_ref/fdtd.x, -Ofast -flto -funroll-loops,
    README.md:122) under mpirun on `ranks` host cores; output_interval past the
    run, so only the t = 0 output falls inside the loop timer.  Returns the
    loop time of "Simulation completed in X seconds" (fdtd.c:3917), the wall
    time with set-up, and the reference's own [MEM] total.  With log_path the
    reference's output also streams to that file as it runs."""
    from php_fdtd_amd import workloads

    ref = os.path.join(ROOT, "oracle", "_ref", "fdtd.x")
    tmp = tempfile.mkdtemp(prefix="phpfdtd_cpu_")
    try:
        sp = workloads.with_steps(spec, nsteps)
        cfg = workloads.write_reference_case(tmp, sp, med)
        env = dict(os.environ, OMP_NUM_THREADS="1")
        out_path = log_path or os.path.join(tmp, "reference.out")
        t0 = time.time()
        with open(out_path, "w") as f:
            p = subprocess.run(["/opt/conda/bin/mpirun", "-np", str(ranks), ref, cfg], cwd=tmp, stdout=f,
                               stderr=subprocess.STDOUT, text=True, timeout=timeout, env=env)
        wall = time.time() - t0
        text = open(out_path).read()
        loop = None
        for ln in text.splitlines():
            if "Simulation completed in" in ln:
                loop = float(ln.split("in")[1].split()[0])
        if loop is None:
            raise RuntimeError("reference run failed rc=%d: %s" % (p.returncode, text[-800:]))
        mem = re.findall(r"\[MEM\] Before time loop: per-rank = \S+ GB, total \(\d+ ranks\) = (\S+) GB", text)
        return loop, wall, float(mem[-1]) if mem else None
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def cpu_baseline(workload="tip_slab", full_steps=1, prefix_steps=10, timeout=1200, log_path=None):
    """The reference solver on the host cores of this box, built with the
    reference's README.md:122 flags and run under mpirun with its loop timer
    (fdtd.c:3917): the workload's own grid for `full_steps` steps when host RAM holds the
    reference's layout of it; otherwise the workload is recorded as not
    runnable and a labelled prefix of the 384x384x256 16-atom configuration
    is timed instead.  Ranks: the CPU share of this process (the box allots
    a fixed share of host cores to each GPU), at most nz.  Without the
    reference binary the single-process C oracle is timed on a 1/8-scale
    sample instead (kind "port")."""
    from php_fdtd_amd import workloads

    info = host_info()
    ranks = info["cpu_share"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():  # the box's per-GPU CPU share
        ranks = min(ranks, int(os.environ["OMP_NUM_THREADS"]))
    ref = os.path.join(ROOT, "oracle", "_ref", "fdtd.x")
    if os.path.exists(ref) and os.path.exists("/opt/conda/bin/mpirun"):
        spec, med = workloads.WORKLOADS[workload]()
        need = reference_host_bytes(spec, med)
        avail = info.get("mem_available_gb", 0) * 1e9
        note = None
        if need * 1.2 > avail:
            note = "%s not runnable: the reference's layout needs ~%.0f GB, host RAM available %.0f GB" % (
                spec["name"], need / 1e9, avail / 1e9)
            spec, med = workloads.lattice16_slab()
            nsteps = prefix_steps
        else:
            nsteps = full_steps
        ranks = max(1, min(ranks, spec["grid"]["nz"]))
        loop, wall, mem = run_reference(spec, med, nsteps, ranks, timeout, log_path)
        out = dict(value=med.size * nsteps / loop / 1e9, unit="Gcell-updates/s", cores=ranks, kind="reference",
                   sample="reference fdtd.c (-Ofast -flto -funroll-loops) under mpirun -np %d on %s, %d step(s) "
                          "(output_interval past the run), loop time %.2f s (fdtd.c:3917), %.1f s with set-up"
                          % (ranks, spec["name"], nsteps, loop, wall),
                   host=info, reference_mem_gb=mem, reference_mem_estimate_gb=round(need / 1e9, 1))
        if note:
            out["note"] = note
        return out
    from oracle import orc

    spec, med = workloads.tip_slab(scale=8)
    spec = workloads.with_steps(spec, 1000)
    o = orc.Oracle(spec, med)
    t0 = time.time()
    n = 0
    while time.time() - t0 < 12.0 and n < 1000:
        o.step(1)
        n += 1
    t = time.time() - t0
    return dict(value=med.size * n / t / 1e9, unit="Gcell-updates/s", cores=ranks, kind="port", host=info,
                sample="C oracle (FP64 restatement, OpenMP over %d threads), 128x128x80 (the 1024x1024x640 "
                       "workload scaled 1/8), %d steps; reference binary absent" % (ranks, n))


def _cpu_baseline_child(path, workload, log_path=None):
    """Entry of the CPU-baseline subprocess: writes its result as JSON."""
    try:
        res = cpu_baseline(workload, log_path=log_path)
    except Exception as e:  # reported, not fatal
        res = {"value": None, "error": str(e)[:400]}
    with open(path, "w") as f:
        json.dump(res, f)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", "--config", dest="workload", default="tip_slab",
                    help="tip_slab, lattice16_slab, lattice4_slab or lorentz_slab")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp32,fp64", help="comma list of fp32, fp32rv64, fp64")
    ap.add_argument("--accuracy", default=None,
                    help="workload of the FP32-vs-FP64 accuracy run (default: the benchmarked workload; '' = skip)")
    ap.add_argument("--accuracy-steps", type=int, default=5000,
                    help="steps of the accuracy run (0 = the workload's own; 5000 carries the 1024x1024x640 "
                         "workload's Ricker pulse, peak at step 2624, past the monitor)")
    ap.add_argument("--accuracy-rv64", action="store_true", help="also run the FP32 form with FP64 r, v")
    ap.add_argument("--cpu-only", action="store_true", help="run only the CPU baseline and print it as the JSON line")
    ap.add_argument("--cpu-log", default=None, help="file the reference's own output streams to as it runs")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--serving", type=int, default=64, help="requests of the serving mix (0 = skip)")
    ap.add_argument("--serving-step-scale", type=float, default=0.05)
    ap.add_argument("--no-copy-probe", dest="copy_probe", action="store_false",
                    help="skip the measured copy-bandwidth ceilings")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving block and print it as the JSON line")
    ap.add_argument("--serving-batched-only", action="store_true",
                    help="run only the FP32 batched serving run (all requests queued) and print it as the JSON "
                         "line, with the step-model bytes of the requests (for profiling the batched kernels)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU only; the domain is resident on a single MI355X and no "
              "multi-GPU path is built (--gpus %d not supported)" % args.gpus, file=sys.stderr)
        return 2

    if args.cpu_child:  # the CPU-baseline subprocess (no GPU use)
        _cpu_baseline_child(args.cpu_child, args.workload, args.cpu_log)
        return 0
    if args.cpu_only:
        print(json.dumps({"cpu_baseline": cpu_baseline(args.workload, log_path=args.cpu_log)}))
        return 0

    from php_fdtd_amd import _abi, workloads

    if args.serving_only:
        print(json.dumps({"serving": serving(args.serving, args.serving_step_scale)}))
        return 0
    if args.serving_batched_only:
        print(json.dumps({"serving_batched": serving_batched(args.serving, args.serving_step_scale)}))
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
    rv_bytes = {"fp32": 4.5, "fp32rv64": 8, "fp64": 8}
    for form in args.forms.split(","):
        prec = precs[form]
        log("form %s" % form)
        res[form] = run_form(spec, med, prec, args.warmup, args.steps, prof_steps)
        log("  %.3f ms/step (events), %.3f ms/step (wall), %.3f ms/step with a slice + monitor read every 10 steps, "
            "%.3f ms per sync with diagnostics" % (res[form]["ms"], res[form]["wall_ms"], res[form]["sampled_ms"],
                                                  res[form]["sync_ms"]))
        for k in res[form]["prof"]:
            log("    %-24s %9.3f ms  %8.1f GB/s" % (k["name"], k["ms"], k["bytes"] / k["ms"] / 1e6 if k["ms"] else 0))

    head = [f for f in ("fp32", "fp32rv64", "fp64") if f in res][0]
    r = res[head]
    word = 8 if head == "fp64" else 4
    gcell = cells / (r["ms"] * 1e-3) / 1e9
    dom = max(r["prof"], key=lambda k: k["ms"])
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
    dom_sym = kernel_symbol(dom["name"], head)
    traffic = profile_traffic(dom_sym) if args.workload == PROFILE_WORKLOAD else None
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
        "config": {"workload": "%s: %s, %s form (fields/CPML/ADE %s, lattice r,v %s)"
                               % (spec["name"], GEOMETRY[args.workload], head.upper(),
                                  "f64" if head == "fp64" else "f32", rv_desc),
                   "cells": int(cells), "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom["name"], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": round(traffic) if traffic else None,
                     "algorithmic_bytes": round(dom["bytes"]), "symbol": dom_sym,
                     "traffic_source": "profiles/step_kernels_tip_slab_1024x1024x640.md: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per "
                                       "dispatch, copy-probe calibrated" if traffic else None},
        "step_roofline": {"bytes_per_step": sb, "achieved": round(sb / (r["ms"] * 1e-3) / 1e9, 1),
                          "frac": round(sb / (r["ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                          "model": ("C(18w+1) - 6w*plain cells (E-form: their D is not stored)" if head != "fp64"
                                    else "C(18w+1)") + " + 8w*CPML incidences + 9w*ADE cells + 12Nb*lattice cells, "
                                   "b = %g bytes per stored r or v value" % rv_bytes[head]},
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
        out["fp64"]["sync_ms"] = round(r64["sync_ms"], 3)
    out["sync_ms"] = round(r["sync_ms"], 3)
    # the reference on the host cores runs in a child process beside the GPU
    # accuracy leg (it uses the host cores only); the serving leg, whose
    # context builds use the host cores too, waits for it
    cpu_proc, cpu_path, cpu_t0 = None, None, time.time()
    if not args.no_cpu:
        fd, cpu_path = tempfile.mkstemp(prefix="phpfdtd_cpu_", suffix=".json")
        os.close(fd)
        cpu_proc = subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-child", cpu_path,
                                     "--workload", args.workload] + (["--cpu-log", args.cpu_log] if args.cpu_log else []))
    acc = args.workload if args.accuracy is None else args.accuracy
    if acc:
        out["accuracy"] = accuracy(acc, steps=args.accuracy_steps or None, rv64=args.accuracy_rv64)
    if cpu_proc is not None:
        cpu_proc.wait()
        try:
            out["cpu_baseline"] = json.load(open(cpu_path))
        except Exception as e:  # reported, not fatal
            out["cpu_baseline"] = {"value": None, "error": "no result from the CPU baseline process: %s" % e}
        os.remove(cpu_path)
        log("cpu baseline %.1f s: %s" % (time.time() - cpu_t0, out["cpu_baseline"]))
    if args.serving:
        out["serving"] = serving(args.serving, args.serving_step_scale)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
