#!/usr/bin/env python3
"""Generate the golden fixtures under tests/golden/<case>/.

For every case this script writes the reference's own inputs (case.cfg,
medium_map.bin, material text files), runs the reference solver built from
fdtd.c by oracle/Makefile (target ``ref``: ``oracle/_ref/fdtd_precise.x``, whose
monitor rows print at %.17e), and keeps its outputs as fixtures:

  monitor_point.dat, full_monitor.dat   per-step / per-interval FP64 samples
  field_slice_*.bin                     float32 slices (int32 dims + data)
  source.txt                            source values at output steps

It also writes case.json: the exact inputs the reference ends up using after
its host loading (grid, dt, unique source frequency, resolved material tensors
with switchxz applied, sources with defaults filled), in the form the C ABI
(include/phpfdtd.h) and the oracle take.  If this JSON misstated what the
reference loads, the oracle-vs-golden tests would fail, so the reference run
pins it.

Run here (needs the reference tree and the built binary):
    make -C oracle ref && python3 tests/golden/make_golden.py
"""
import json
import math
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("PHPFDTD_REFERENCE", "/root/reference")
REF_BIN = os.path.join(REPO, "oracle", "_ref", "fdtd_precise.x")

C0 = 3e8
EPS0 = 8.854e-12
MU0 = math.pi * 4e-7


def cfl_dt(dx, dy, dz):
    """fdtd.c:1906-1910: 0.99 * min(d) / (c*sqrt(3))."""
    m = min(dx, dy, dz)
    return m / (C0 * math.sqrt(3.0)) * 0.99


def switchxz(t):
    """x<->z similarity permutation of a row-major 3x3 (fdtd.c:2270-2286)."""
    p = [2, 1, 0]
    return [t[p[r] * 3 + p[c]] for r in range(3) for c in range(3)]


def switchxz_blocks(a, natom, dmm=False):
    """Apply switchxz to every 3x3 block of QeZV/QeZM ([natom][3][3]) or DMM
    ([natom][3][natom][3]) as fdtd.c:2333-2349, 2405-2423 do."""
    a = list(a)
    if not dmm:
        out = []
        for ia in range(natom):
            out += switchxz(a[ia * 9:(ia + 1) * 9])
        return out
    n3 = 3 * natom
    out = list(a)
    for ia in range(natom):
        for ja in range(natom):
            blk = [a[(ia * 3 + r) * n3 + ja * 3 + c] for r in range(3) for c in range(3)]
            blk = switchxz(blk)
            for r in range(3):
                for c in range(3):
                    out[(ia * 3 + r) * n3 + ja * 3 + c] = blk[r * 3 + c]
    return out


def sio2(f):
    """Built-in material 1 (fdtd.c:649-667) at source frequency f."""
    w = f / 0.03e12
    e0, s1, g1, w1, s2, g2, w2 = 1.5, 231, 69, 806, 866, 75, 1063
    t1 = w1 * w1 - w * w
    t2 = w2 * w2 - w * w
    er = e0 + s1 * s1 * t1 / (t1 * t1 + w * w * g1 * g1) + s2 * s2 * t2 / (t2 * t2 + w * w * g2 * g2)
    ei = s1 * s1 * w * g1 / (t1 * t1 + w * w * g1 * g1) + s2 * s2 * w * g2 / (t2 * t2 + w * w * g2 * g2)
    return er, ei


def diag(a, b, c):
    return [a, 0.0, 0.0, 0.0, b, 0.0, 0.0, 0.0, c]


def vacuum_mat():
    return dict(eps_re=diag(1, 1, 1), eps_im=[0.0] * 9, sigma=0.0, anglez=0.0, anglex=0.0,
                use_ade=0, natom=0, qezv=[], qezm=[], dmm=[])


def builtins(f):
    er, ei = sio2(f)
    m1 = vacuum_mat()
    m1.update(eps_re=diag(er, er, er), eps_im=diag(ei, ei, ei), use_ade=1 if er <= 0 else 0)
    m2 = vacuum_mat()
    m2.update(eps_re=diag(8.0, 7.0, 5.0))
    m3 = vacuum_mat()
    m3.update(eps_re=diag(9.0, 9.0, 9.0), sigma=0.1)
    return [vacuum_mat(), m1, m2, m3]


def read_numbers(path, n):
    vals = open(path).read().split()
    return [float(x) for x in vals[:n]]


def outcar_eps(path):
    """Static dielectric block of an OUTCAR fragment (fdtd.c:423-552)."""
    lines = open(path).read().splitlines()
    out = []
    on = False
    for ln in lines:
        if "MACROSCOPIC STATIC DIELECTRIC TENSOR" in ln:
            on = True
            continue
        if on:
            if "---" in ln:
                continue
            parts = ln.split()
            if len(parts) == 3:
                out += [float(x) for x in parts]
            if len(out) == 9:
                return out
    raise ValueError("no dielectric block in " + path)


def write_medium(path, ids):
    nx, ny, nz = ids.shape
    with open(path, "wb") as f:
        f.write(np.array([nx, ny, nz], dtype="<i4").tobytes())
        f.write(np.ascontiguousarray(ids, dtype="<i4").tobytes())


def cfg_source(s):
    names = {0: "gaussian", 1: "sinusoidal", 2: "ricker", 3: "dipole"}
    parts = ["type=%s" % names[s["type"]], "x=%d" % s["x"], "y=%d" % s["y"], "z=%d" % s["z"],
             "amplitude=%r" % s["amplitude"], "component=%s" % s["component"]]
    if s.get("wavenumber") is not None:
        parts.append("wavenumber=%r" % s["wavenumber"])
    else:
        parts.append("frequency=%r" % s["frequency"])
    for k in ("t0", "tau", "start_time", "end_time"):
        if s.get(k) is not None:
            parts.append("%s=%r" % (k, s[k]))
    return "source " + ", ".join(parts)


def resolve_source(s, tsteps, dt):
    """Source after parse_source (fdtd.c:1403-1510) and read_config's end_time
    default (fdtd.c:1913-1917).  Every case sets t0/tau/position explicitly."""
    f = s["frequency"] if s.get("wavenumber") is None else s["wavenumber"] * 0.03e12
    end = s.get("end_time")
    if end is None or end <= 0:
        end = tsteps * dt
    # parse-time defaults: t0 = 40*dt and tau = 15*dt with dt still -1 (fdtd.c:1424-1425, 1755)
    return dict(type=s["type"], x=s["x"], y=s["y"], z=s["z"], component=s["component"],
                amplitude=s["amplitude"], frequency=f, t0=s.get("t0", -40.0), tau=s.get("tau", -15.0),
                start_time=s.get("start_time", 0.0) or 0.0, end_time=end)


# --------------------------------------------------------------------------
# Cases
# --------------------------------------------------------------------------
def case_vacuum_cpml():
    nx, ny, nz = 22, 20, 18
    g = dict(nx=nx, ny=ny, nz=nz, dx=20e-9, dy=25e-9, dz=15e-9, L=4, R=1e-6, tsteps=150, interval=25)
    ids = np.zeros((nx, ny, nz), np.int32)
    srcs = [dict(type=2, x=11, y=10, z=9, amplitude=1.0, frequency=2e14, t0=2e-15, component="z")]
    out = dict(component="Ez", axis="z", pos=9)
    mon = dict(x=11, y=10, z=12, atom=0, dir="x")
    return g, ids, [], srcs, out, mon, {}


def case_aniso_loss():
    nx, ny, nz = 24, 22, 20
    g = dict(nx=nx, ny=ny, nz=nz, dx=20e-9, dy=20e-9, dz=20e-9, L=4, R=1e-6, tsteps=160, interval=40)
    ids = np.zeros((nx, ny, nz), np.int32)
    ids[:, :, 8:14] = 4
    ids[3:9, 4:10, 14:18] = 5
    m4 = ("material id=4, name=aniso, eps_xx=4.0, eps_xy=0.3, eps_xz=0.1, eps_yx=0.3, eps_yy=5.0, "
          "eps_yz=-0.2, eps_zx=0.1, eps_zy=-0.2, eps_zz=3.0, eps_xxi=0.5, eps_xyi=0.05, eps_xzi=0.0, "
          "eps_yxi=0.05, eps_yyi=0.4, eps_yzi=0.01, eps_zxi=0.0, eps_zyi=0.01, eps_zzi=0.3, "
          "sigma=2000.0, anglez=0.4")
    m5 = "material id=5, name=rotx, eps_xx=2.0, eps_yy=3.0, eps_zz=4.0, anglex=0.7, sigma=500.0"
    mats = {4: dict(eps_re=[4.0, 0.3, 0.1, 0.3, 5.0, -0.2, 0.1, -0.2, 3.0],
                    eps_im=[0.5, 0.05, 0.0, 0.05, 0.4, 0.01, 0.0, 0.01, 0.3],
                    sigma=2000.0, anglez=0.4, anglex=0.0, use_ade=0, natom=0, qezv=[], qezm=[], dmm=[]),
            5: dict(eps_re=diag(2.0, 3.0, 4.0), eps_im=[0.0] * 9, sigma=500.0, anglez=0.0, anglex=0.7,
                    use_ade=0, natom=0, qezv=[], qezm=[], dmm=[])}
    srcs = [dict(type=0, x=12, y=11, z=5, amplitude=1.0, frequency=3e13, t0=1e-15, tau=5e-16, component="x"),
            dict(type=1, x=6, y=7, z=15, amplitude=0.5, frequency=1e14, component="y",
                 start_time=1e-16, end_time=2e-15)]
    out = dict(component="Ey", axis="y", pos=11)
    mon = dict(x=12, y=11, z=10, atom=0, dir="x")
    return g, ids, [m4, m5], srcs, out, mon, mats


def case_ade():
    nx, ny, nz = 24, 24, 20
    g = dict(nx=nx, ny=ny, nz=nz, dx=20e-9, dy=20e-9, dz=20e-9, L=4, R=1e-6, tsteps=160, interval=40)
    ids = np.zeros((nx, ny, nz), np.int32)
    ids[:, :, 11:15] = 4
    ids[5:19, 5:19, 4:7] = 5
    # dielectric table (15 columns, fdtd.c:346-401); rows around the source frequency
    f_src = 1000 * 0.03e12
    table = []
    for fr, er, ei in ((2.9e13, (-4.0, -7.0, 3.5), (1.8, 0.9, 0.4)),
                       (3.0e13, (-5.0, -8.0, 3.0), (2.0, 1.0, 0.5)),
                       (3.2e13, (-6.0, -9.0, 2.5), (2.2, 1.1, 0.6))):
        table.append("%r %r %r %r 0 0 %r 0 %r %r 0 0 %r 0 %r" % (
            fr, C0 / fr, fr / 0.03e12, er[0], er[1], er[2], ei[0], ei[1], ei[2]))
    files = {"eps_table.txt": "# freq wavelength wavenumber xx xy xz yy yz zz (imag ...)\n" + "\n".join(table) + "\n"}
    m4 = ("material id=4, name=drude, use_ADE=1, eps_xx=-30.0, eps_yy=-30.0, eps_zz=-30.0, "
          "eps_xxi=8.0, eps_yyi=8.0, eps_zzi=8.0")
    m5 = "material id=5, name=table, dielectric=eps_table.txt, use_ADE=1, anglex=0.5, switchxz=1"
    e5r = switchxz(diag(-5.0, -8.0, 3.0))
    e5i = switchxz(diag(2.0, 1.0, 0.5))
    mats = {4: dict(eps_re=diag(-30.0, -30.0, -30.0), eps_im=diag(8.0, 8.0, 8.0), sigma=0.0, anglez=0.0,
                    anglex=0.0, use_ade=1, natom=0, qezv=[], qezm=[], dmm=[]),
            5: dict(eps_re=e5r, eps_im=e5i, sigma=0.0, anglez=0.0, anglex=0.5, use_ade=1, natom=0,
                    qezv=[], qezm=[], dmm=[])}
    srcs = [dict(type=3, x=12, y=12, z=9, amplitude=1.0, wavenumber=1000.0, t0=1.5e-15, tau=8e-16,
                 component="z")]
    out = dict(component="E", axis="x", pos=12)
    mon = dict(x=12, y=12, z=12, atom=0, dir="x")
    assert abs(srcs[0]["wavenumber"] * 0.03e12 - f_src) < 1.0
    return g, ids, [m4, m5], srcs, out, mon, mats, files


def lattice_material(name, natom, datadir, files, sw, anglez=0.0, anglex=0.0, qezm_name="QeZM_clean.txt"):
    """Copy the reference's material data (fixtures) and resolve them as the
    reference loads them (fdtd.c:2248-2431)."""
    qv = read_numbers(os.path.join(REF, datadir, "QeZV_clean.txt"), 9 * natom)
    qm = read_numbers(os.path.join(REF, datadir, qezm_name), 9 * natom)
    dm = read_numbers(os.path.join(REF, datadir, "DMM_clean.txt"), 9 * natom * natom)
    eps = outcar_eps(os.path.join(REF, datadir, "OUTCAR.txt"))
    for fn in ("QeZV_clean.txt", qezm_name, "DMM_clean.txt", "OUTCAR.txt"):
        files[name + "_" + fn] = open(os.path.join(REF, datadir, fn)).read()
    if sw:
        qv = switchxz_blocks(qv, natom)
        qm = switchxz_blocks(qm, natom)
        dm = switchxz_blocks(dm, natom, dmm=True)
        eps = switchxz(eps)
    mat = dict(eps_re=eps, eps_im=[0.0] * 9, sigma=0.0, anglez=anglez, anglex=anglex, use_ade=0,
               natom=natom, qezv=qv, qezm=qm, dmm=dm)
    line = ("material id=4, name=%s, natom=%d, born=%s_OUTCAR.txt, QeZV=%s_QeZV_clean.txt, "
            "QeZM=%s_%s, DMM=%s_DMM_clean.txt" % (name, natom, name, name, name, qezm_name, name))
    if sw:
        line += ", switchxz=1"
    if anglez:
        line += ", anglez=%r" % anglez
    if anglex:
        line += ", anglex=%r" % anglex
    return line, mat


def case_hbn():
    nx, ny, nz = 20, 20, 16
    g = dict(nx=nx, ny=ny, nz=nz, dx=20e-9, dy=20e-9, dz=20e-9, L=3, R=1e-6, tsteps=200, interval=20)
    ids = np.zeros((nx, ny, nz), np.int32)
    ids[:, :, 6:10] = 4
    files = {}
    line, mat = lattice_material("hBN", 4, "Material_data/hBN", files, sw=False, anglez=0.5235987755982988)
    srcs = [dict(type=3, x=10, y=10, z=4, amplitude=1.0, wavenumber=1370.0, t0=2e-15, tau=1e-15, component="z")]
    out = dict(component="Ez", axis="z", pos=7)
    mon = dict(x=10, y=10, z=7, atom=1, dir="x")
    return g, ids, [line], srcs, out, mon, {4: mat}, files


def case_moo3():
    nx, ny, nz = 18, 18, 14
    g = dict(nx=nx, ny=ny, nz=nz, dx=20e-9, dy=20e-9, dz=20e-9, L=3, R=1e-6, tsteps=120, interval=30)
    ids = np.zeros((nx, ny, nz), np.int32)
    ids[2:16, 3:15, 5:9] = 4
    files = {}
    line, mat = lattice_material("MoO3", 16, "Material_data/MoO3", files, sw=True, anglex=0.3,
                                 qezm_name="QeZm_clean.txt")
    srcs = [dict(type=3, x=9, y=9, z=3, amplitude=1.0, wavenumber=800.0, t0=2e-15, tau=1e-15, component="x")]
    out = dict(component="Dz", axis="y", pos=9)
    mon = dict(x=9, y=9, z=6, atom=3, dir="z")
    return g, ids, [line], srcs, out, mon, {4: mat}, files


def case_builtins_nopml():
    nx, ny, nz = 20, 18, 16
    g = dict(nx=nx, ny=ny, nz=nz, dx=20e-9, dy=20e-9, dz=20e-9, L=0, R=1e-6, tsteps=120, interval=30)
    ids = np.zeros((nx, ny, nz), np.int32)
    ids[2:8, :, 3:8] = 1
    ids[10:18, 2:9, 8:13] = 2
    ids[10:18, 9:16, 2:15] = 3
    srcs = [dict(type=0, x=9, y=9, z=8, amplitude=1.0, frequency=2.5e13, t0=1e-15, tau=5e-16, component="z"),
            dict(type=2, x=4, y=12, z=10, amplitude=2.0, frequency=1.5e14, t0=1.2e-15, component="y")]
    out = dict(component="Dx", axis="z", pos=nz - 3)
    mon = dict(x=14, y=12, z=nz - 1, atom=0, dir="y")  # monitor at the top plane: fields from the halo
    return g, ids, [], srcs, out, mon, {}


def case_diverge():
    """The hBN lattice driven ~1e19 times harder from an off-centre dipole, so
    that max |r| passes the reference's 1e-10 m guard part way through the
    run: the reference prints its divergence report and aborts
    (fdtd.c:3626-3667).  Every equation is linear, so nothing overflows; the
    off-centre source keeps the argmax cell free of mirror-image ties."""
    nx, ny, nz = 20, 20, 16
    g = dict(nx=nx, ny=ny, nz=nz, dx=20e-9, dy=20e-9, dz=20e-9, L=3, R=1e-6, tsteps=200, interval=20)
    ids = np.zeros((nx, ny, nz), np.int32)
    ids[:, :, 6:10] = 4
    files = {}
    line, mat = lattice_material("hBN", 4, "Material_data/hBN", files, sw=False, anglez=0.5235987755982988)
    srcs = [dict(type=3, x=7, y=12, z=4, amplitude=5e18, wavenumber=1370.0, t0=2e-15, tau=1e-15, component="z")]
    out = dict(component="Ez", axis="x", pos=8)
    mon = dict(x=8, y=11, z=7, atom=2, dir="y")
    return g, ids, [line], srcs, out, mon, {4: mat}, files, True


CASES = {
    "vacuum_cpml": case_vacuum_cpml,
    "aniso_loss": case_aniso_loss,
    "ade": case_ade,
    "hbn": case_hbn,
    "moo3": case_moo3,
    "builtins_nopml": case_builtins_nopml,
    "diverge": case_diverge,
}


def build_case(name):
    res = CASES[name]()
    g, ids, mat_lines, srcs, out, mon, mats = res[:7]
    files = res[7] if len(res) > 7 else {}
    aborts = res[8] if len(res) > 8 else False
    d = os.path.join(HERE, name)
    if os.path.isdir(d):
        shutil.rmtree(d)
    os.makedirs(d)
    for fn, text in files.items():
        with open(os.path.join(d, fn), "w") as f:
            f.write(text)
    write_medium(os.path.join(d, "medium_map.bin"), ids)
    lines = ["# golden case %s (generated by tests/golden/make_golden.py)" % name,
             "nx = %d" % g["nx"], "ny = %d" % g["ny"], "nz = %d" % g["nz"], "tsteps = %d" % g["tsteps"],
             "delta_x = %r" % g["dx"], "delta_y = %r" % g["dy"], "delta_z = %r" % g["dz"],
             "pml_thickness = %d" % g["L"], "pml_reflection = %r" % g["R"],
             "output_interval = %d" % g["interval"], "output_component = %s" % out["component"],
             "output_slice_axis = %s" % out["axis"], "output_slice_position = %d" % out["pos"]]
    lines += [cfg_source(s) for s in srcs]
    lines += mat_lines
    lines += ["medium_file = medium_map.bin",
              "monitor_point = x=%d,y=%d,z=%d,atom=%d,dir=%s" % (mon["x"], mon["y"], mon["z"], mon["atom"], mon["dir"]),
              "monitor_output_file = monitor_point.dat", "full_monitor = 1",
              "full_monitor_output_file = full_monitor.dat"]
    with open(os.path.join(d, "case.cfg"), "w") as f:
        f.write("\n".join(lines) + "\n")

    dt = cfl_dt(g["dx"], g["dy"], g["dz"])
    rsrcs = [resolve_source(s, g["tsteps"], dt) for s in srcs]
    f_src = rsrcs[0]["frequency"]
    nmat = max(4, int(ids.max()) + 1, max(mats.keys(), default=0) + 1)
    table = builtins(f_src) + [vacuum_mat() for _ in range(nmat - 4)]
    for k, m in mats.items():
        table[k] = m
    spec = dict(
        name=name,
        grid=dict(nx=g["nx"], ny=g["ny"], nz=g["nz"], tsteps=g["tsteps"], dx=g["dx"], dy=g["dy"], dz=g["dz"],
                  dt=dt, eps0=EPS0, mu0=MU0, pml_thickness=g["L"], pml_reflection=g["R"], pml_order=3.0,
                  kappa_max=7.0, alpha_max=0.05, f_source=f_src),
        materials=table, sources=rsrcs, output=dict(out, interval=g["interval"]),
        monitor=dict(mon, dir={"x": 0, "y": 1, "z": 2}[mon["dir"]]),
    )
    if aborts:
        spec["aborts"] = True
    with open(os.path.join(d, "case.json"), "w") as f:
        json.dump(spec, f, indent=1)

    log = subprocess.run([REF_BIN, "case.cfg"], cwd=d, capture_output=True, text=True, timeout=600)
    with open(os.path.join(d, "reference.log"), "w") as f:
        f.write(log.stdout[-20000:])
        f.write(log.stderr[-5000:])
    if aborts:
        # the reference stops at MPI_Abort without flushing its monitor files:
        # only the log (stdout, flushed) and the closed slice files are kept
        if log.returncode == 0 or "ERROR: r field too large" not in log.stdout:
            raise SystemExit("reference did not abort on %s" % name)
        for fn in ("monitor_point.dat", "full_monitor.dat"):
            p = os.path.join(d, fn)
            if os.path.exists(p):
                os.remove(p)
    elif log.returncode != 0:
        raise SystemExit("reference failed on %s: rc=%d\n%s" % (name, log.returncode, log.stdout[-2000:]))
    print("%-16s ok  slices=%d" % (name, len([x for x in os.listdir(d) if x.startswith("field_slice_")])))


def main():
    if not os.path.exists(REF_BIN):
        raise SystemExit("build the reference first: make -C oracle ref")
    names = sys.argv[1:] or list(CASES)
    for n in names:
        build_case(n)


if __name__ == "__main__":
    main()
