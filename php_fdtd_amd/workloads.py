"""Seeded synthetic workloads (grids, media, materials, sources) for the
benchmark configurations, in the plain-dict spec form of ``php_fdtd_amd.sim``,
plus writers of the reference's own input formats (cfg keys, medium_map.bin,
QeZV/QeZM/DMM text, OUTCAR fragment, 15-column dielectric table) so the same
workload can be run by the reference solver.

Workloads (key in WORKLOADS, spec name with its grid size):
  lorentz_slab    100^3, 20 nm, 2000 steps, 10-cell CPML (kappa_max 7); isotropic Lorentz slab
                  (eps_inf 2.5, w0 1370 cm^-1, d_eps 2.0, gamma 5 cm^-1) taken
                  at the source wavenumber 650 cm^-1 through the reference's
                  Drude-fit ADE.
  lattice4_slab   160x160x120, 20 nm, 20000 steps, 12-cell CPML; 20-cell slab
                  of a 4-atom lattice (10.8, 10.8, 14.0, 14.0 amu) rotated
                  anglez = 30 deg.
  lattice16_slab  384x384x256, 20 nm, 30000 steps, 16-cell CPML; 32-cell slab
                  of a 16-atom lattice (4 x 95.95 + 12 x 16.0 amu), anglex =
                  90 deg + switchxz, full complex eps tensor.
  tip_slab        1024x1024x640, 10 nm, 10000 steps, 20-cell CPML; 48-cell slab
                  of the 16-atom lattice material plus a Drude-metal cone (apex
                  radius 10 cells, 30 deg half-angle) 3 cells above the slab,
                  Ricker source under the apex.
  serving mix     grids {64^3, 96^3, 128^3, 192x192x128, 256x256x160},
                  materials {Lorentz, 4-atom lattice, 16-atom lattice}, source
                  U(600, 1600) cm^-1, steps U(5000, 20000), 10-cell CPML,
                  Poisson arrivals.

Every medium is drawn by ``php_fdtd_amd.medium`` from a ``[medium]`` geometry
(cubes, plus a cone for tip_slab) kept in ``spec["medium"]``; the cfg written by
``write_reference_case`` carries it, so the same map can be redrawn from the
cfg alone.
"""
import math
import os

import numpy as np

from . import medium as _medium

C0 = 3e8
EPS0 = 8.854e-12
MU0 = math.pi * 4e-7
E_CHARGE = 1.602176634e-19
AMU = 1.66053906660e-27
WN_TO_HZ = 0.03e12  # README.md:344: f = wavenumber * 0.03e12


def cfl_dt(d):
    """fdtd.c:1906-1910 automatic time step for cell size d (m)."""
    return d / (C0 * math.sqrt(3.0)) * 0.99


def diag(a, b, c):
    return [float(a), 0.0, 0.0, 0.0, float(b), 0.0, 0.0, 0.0, float(c)]


def vacuum():
    return dict(eps_re=diag(1, 1, 1), eps_im=[0.0] * 9, sigma=0.0, anglez=0.0, anglex=0.0, use_ade=0, natom=0,
                qezv=[], qezm=[], dmm=[])


def builtin_table(f):
    """Materials 0-3 that the reference always defines (fdtd.c:628-715)."""
    w = f / WN_TO_HZ
    e0, s1, g1, w1, s2, g2, w2 = 1.5, 231, 69, 806, 866, 75, 1063
    t1 = w1 * w1 - w * w
    t2 = w2 * w2 - w * w
    er = e0 + s1 * s1 * t1 / (t1 * t1 + w * w * g1 * g1) + s2 * s2 * t2 / (t2 * t2 + w * w * g2 * g2)
    ei = s1 * s1 * w * g1 / (t1 * t1 + w * w * g1 * g1) + s2 * s2 * w * g2 / (t2 * t2 + w * w * g2 * g2)
    m1 = vacuum()
    m1.update(eps_re=diag(er, er, er), eps_im=diag(ei, ei, ei), use_ade=1 if er <= 0 else 0)
    m2 = vacuum()
    m2.update(eps_re=diag(8, 7, 5))
    m3 = vacuum()
    m3.update(eps_re=diag(9, 9, 9), sigma=0.1)
    return [vacuum(), m1, m2, m3]


def switchxz3(t):
    p = [2, 1, 0]
    return [t[p[r] * 3 + p[c]] for r in range(3) for c in range(3)]


def switchxz_lattice(qezv, qezm, dmm, natom):
    """x<->z permutation of every 3x3 block, as the reference applies it on
    load (fdtd.c:2333-2349, 2369-2384, 2405-2423)."""
    p = np.array([2, 1, 0])
    qv = np.asarray(qezv).reshape(natom, 3, 3)[:, p][:, :, p]
    qm = np.asarray(qezm).reshape(natom, 3, 3)[:, p][:, :, p]
    d = np.asarray(dmm).reshape(natom, 3, natom, 3)[:, p][:, :, :, p]
    return qv.ravel().tolist(), qm.ravel().tolist(), d.ravel().tolist()


def lattice(natom, masses, volume_A3, wn_lo, wn_hi, z_diag, seed):
    """Synthetic phonon lattice: DMM = V diag(0,0,0, w_1^2..) V^T with the three
    zero modes the uniform translations and the optical frequencies uniform in
    [wn_lo, wn_hi] cm^-1; Born charges random 3x3 per atom (diagonal +-z_diag,
    off-diagonal U(-0.1, 0.1)) with the acoustic sum rule sum_a Z*_a = 0;
    QeZV = Z* e / Omega, QeZM = Z* e / m (README.md:537-581 units)."""
    rng = np.random.default_rng(seed)
    n3 = 3 * natom
    trans = np.zeros((n3, 3))
    for d in range(3):
        trans[d::3, d] = 1.0 / math.sqrt(natom)
    g = rng.standard_normal((n3, n3 - 3))
    g -= trans @ (trans.T @ g)
    q, _ = np.linalg.qr(g)
    V = np.concatenate([trans, q], axis=1)
    wn = np.sort(rng.uniform(wn_lo, wn_hi, n3 - 3))
    w2 = np.concatenate([np.zeros(3), (2 * math.pi * wn * WN_TO_HZ) ** 2])
    dmm = (V * w2) @ V.T
    dmm = 0.5 * (dmm + dmm.T)
    Z = rng.uniform(-0.1, 0.1, (natom, 3, 3))
    signs = np.where(np.arange(natom) % 2 == 0, 1.0, -1.0)
    for a in range(natom):
        for d in range(3):
            Z[a, d, d] = signs[a] * z_diag
    Z -= Z.mean(axis=0, keepdims=True)
    omega = volume_A3 * 1e-30
    m = np.asarray(masses, dtype=np.float64) * AMU
    qezv = (Z * E_CHARGE / omega).ravel()
    qezm = (Z * E_CHARGE / m[:, None, None]).ravel()
    return dict(qezv=qezv.tolist(), qezm=qezm.tolist(), dmm=dmm.ravel().tolist(), born=Z, wn=wn)


def lorentz_eps(wn, eps_inf, w0, deps, gamma):
    """eps(w) = eps_inf + deps w0^2 / (w0^2 - w^2 - i gamma w), w in cm^-1."""
    e = eps_inf + deps * w0 * w0 / (w0 * w0 - wn * wn - 1j * gamma * wn)
    return e.real, e.imag


def drude_eps(wn, wp, gamma):
    e = 1 - wp * wp / (wn * wn + 1j * gamma * wn)
    return e.real, e.imag


def _grid(n, d, steps, L, f):
    nx, ny, nz = n
    return dict(nx=nx, ny=ny, nz=nz, tsteps=steps, dx=d, dy=d, dz=d, dt=cfl_dt(d), eps0=EPS0, mu0=MU0,
                pml_thickness=L, pml_reflection=1e-6, pml_order=3.0, kappa_max=7.0, alpha_max=0.05, f_source=f)


def _source(kind, pos, comp, f, t0=None, tau=None, amplitude=1.0, tsteps=0, dt=0.0):
    types = {"gaussian": 0, "sinusoidal": 1, "ricker": 2, "dipole": 3}
    period = 1.0 / f
    return dict(type=types[kind], x=pos[0], y=pos[1], z=pos[2], component=comp, amplitude=amplitude, frequency=f,
                t0=t0 if t0 is not None else 1.5 * period, tau=tau if tau is not None else 0.5 * period,
                start_time=0.0, end_time=tsteps * dt)


def _slab(n, z0, z1, material=4):
    """Cube covering the whole x-y extent over z0 <= z < z1."""
    return dict(xrange=(0, n[0] - 1), yrange=(0, n[1] - 1), zrange=(z0, z1 - 1), material=material)


def _draw(spec, geometry):
    """Attach the [medium] geometry to spec and draw it on the spec's grid."""
    g = spec["grid"]
    spec["medium"] = geometry
    return _medium.build((g["nx"], g["ny"], g["nz"]), (g["dx"], g["dy"], g["dz"]), geometry)


def material_lorentz(f):
    er, ei = lorentz_eps(f / WN_TO_HZ, 2.5, 1370.0, 2.0, 5.0)
    m = vacuum()
    m.update(eps_re=diag(er, er, er), eps_im=diag(ei, ei, ei), use_ade=1, name="lorentz650")
    return m


def material_lattice4(seed=1):
    lat = lattice(4, [10.8, 10.8, 14.0, 14.0], 34.626, 750.0, 1650.0, 2.7, seed)
    m = vacuum()
    m.update(eps_re=diag(4.9, 4.9, 2.9), anglez=math.radians(30.0), natom=4, qezv=lat["qezv"], qezm=lat["qezm"],
             dmm=lat["dmm"], name="lattice4")
    m["born"] = lat["born"]
    return m


def material_lattice16(seed=2):
    lat = lattice(16, [95.95] * 4 + [16.0] * 12, 200.536, 200.0, 1050.0, 2.5, seed)
    rng = np.random.default_rng(seed + 100)
    re = np.diag(rng.uniform(2, 6, 3))
    off = rng.uniform(-0.3, 0.3, 3)
    re[0, 1] = re[1, 0] = off[0]
    re[0, 2] = re[2, 0] = off[1]
    re[1, 2] = re[2, 1] = off[2]
    im = 0.05 * re
    qv, qm, dm = switchxz_lattice(lat["qezv"], lat["qezm"], lat["dmm"], 16)
    m = vacuum()
    m.update(eps_re=re.ravel().tolist(), eps_im=im.ravel().tolist(), anglex=math.radians(90.0), natom=16, qezv=qv,
             qezm=qm, dmm=dm, name="lattice16")
    m["raw"] = dict(qezv=lat["qezv"], qezm=lat["qezm"], dmm=lat["dmm"], born=lat["born"])
    return m


def lorentz_slab():
    f = 650.0 * WN_TO_HZ
    # kappa_max stays at the reference's hard-coded 7 (fdtd.c:976), not 1:
    # the reference's CPML uses the old psi in the field update and advances
    # it afterwards (fdtd.c:3152-3155), and with kappa = 1 that recursion is
    # unstable - the field grows ~1e42-fold per 100 steps from the PML
    # corners, in vacuum, in the oracle and in both forms alike
    # (tests/test_oracle_pins.py::test_lagged_cpml_needs_kappa)
    g = _grid((100, 100, 100), 20e-9, 2000, 10, f)
    mats = builtin_table(f) + [material_lorentz(f)]
    src = [_source("ricker", (50, 50, 64), "z", f, t0=5.128e-14, tsteps=2000, dt=g["dt"])]
    spec = dict(name="lorentz_slab_100x100x100", grid=g, materials=mats, sources=src, monitor=dict(x=50, y=50, z=69, atom=0, dir=2))
    return spec, _draw(spec, dict(cubes=[_slab((100, 100, 100), 40, 60)]))


def lattice4_slab():
    f = 1400.0 * WN_TO_HZ
    g = _grid((160, 160, 120), 20e-9, 20000, 12, f)
    mats = builtin_table(f) + [material_lattice4()]
    src = [_source("dipole", (80, 80, 45), "z", f, tsteps=20000, dt=g["dt"])]
    spec = dict(name="lattice4_slab_160x160x120", grid=g, materials=mats, sources=src, monitor=dict(x=80, y=80, z=60, atom=0, dir=0))
    return spec, _draw(spec, dict(cubes=[_slab((160, 160, 120), 50, 70)]))


def lattice16_slab():
    f = 800.0 * WN_TO_HZ
    g = _grid((384, 384, 256), 20e-9, 30000, 16, f)
    mats = builtin_table(f) + [material_lattice16()]
    src = [_source("dipole", (192, 192, 100), "z", f, tsteps=30000, dt=g["dt"])]
    spec = dict(name="lattice16_slab_384x384x256", grid=g, materials=mats, sources=src, monitor=dict(x=192, y=192, z=128, atom=0, dir=0))
    return spec, _draw(spec, dict(cubes=[_slab((384, 384, 256), 112, 144)]))


def tip_cone(n, apex_z, apex_r, half_angle_deg, top_z, d, material=5, center=None):
    """Cone entry of the [medium] geometry: axis along +z from (cx, cy, apex_z)
    to (cx, cy, top_z - 1), radius apex_r cells at the apex growing by
    tan(half angle) per cell, in metres for cell size d."""
    cx, cy = center if center else (n[0] // 2, n[1] // 2)
    t = math.tan(math.radians(half_angle_deg))
    return dict(p1=(cx, cy, apex_z), p2=(cx, cy, top_z - 1), r1=apex_r * d, r2=(apex_r + (top_z - 1 - apex_z) * t) * d,
                material=material)


def material_drude_tip(f):
    er, ei = drude_eps(f / WN_TO_HZ, 7e4, 300.0)
    m = vacuum()
    m.update(eps_re=diag(er, er, er), eps_im=diag(ei, ei, ei), use_ade=1, name="drude_tip")
    return m


def tip_slab(scale=1):
    """scale > 1 shrinks every length by that factor (same physics layout);
    scale=1 is the full 1024x1024x640 configuration."""
    f = 2.7e13
    n = (1024 // scale, 1024 // scale, 640 // scale)
    L = max(2, 20 // scale)
    g = _grid(n, 10e-9 * scale, 10000, L, f)
    slab0 = n[2] * 200 // 640
    slab1 = slab0 + max(2, 48 // scale)
    apex = slab1 + max(1, 3 // scale)
    mats = builtin_table(f) + [material_lattice16(), material_drude_tip(f)]
    src_z = slab1 + max(0, (apex - slab1) // 2)
    src = [_source("ricker", (n[0] // 2, n[1] // 2, src_z), "z", f, t0=5e-14, tsteps=10000, dt=g["dt"])]
    spec = dict(name="tip_slab_%dx%dx%d" % n, grid=g, materials=mats, sources=src,
                monitor=dict(x=n[0] // 2, y=n[1] // 2, z=slab0 + (slab1 - slab0) // 2, atom=0, dir=2))
    geometry = dict(cubes=[_slab(n, slab0, slab1)],
                    cones=[tip_cone(n, apex, max(1, 10 // scale), 30.0, n[2] - L, g["dx"])])
    return spec, _draw(spec, geometry)


WORKLOADS = {"lorentz_slab": lorentz_slab, "lattice4_slab": lattice4_slab, "lattice16_slab": lattice16_slab,
             "tip_slab": tip_slab}

# serving mix
SERVE_GRIDS = [(64, 64, 64), (96, 96, 96), (128, 128, 128), (192, 192, 128), (256, 256, 160)]
SERVE_KINDS = ("lorentz", "lattice4", "lattice16")


def serving_stream(n, seed=4, step_scale=1.0):
    """n seeded request descriptors of the serving mix: grid, material kind,
    source wavenumber U(600, 1600) cm^-1, steps U(5000, 20000) x step_scale,
    and a unit-rate exponential inter-arrival gap (a Poisson stream once
    divided by the offered rate)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        out.append(dict(id=i, shape=SERVE_GRIDS[int(rng.integers(len(SERVE_GRIDS)))],
                        kind=SERVE_KINDS[int(rng.integers(len(SERVE_KINDS)))], wn=float(rng.uniform(600.0, 1600.0)),
                        steps=max(1, int(round(rng.uniform(5000, 20000) * step_scale))),
                        gap=float(rng.exponential(1.0))))
    return out


_LATTICE_CACHE = {}


def serving_case(req):
    """(spec, medium) of one serving request: 20 nm cells, 10-cell CPML, a slab
    of the request's material over the middle sixth of z, a z-dipole at the
    source wavenumber 5 cells under it, the monitor 3 cells above it.  The
    Lorentz material uses the ADE where the reference allows it (Re eps < 20)
    and the lossy-dielectric form elsewhere."""
    nx, ny, nz = req["shape"]
    f = req["wn"] * WN_TO_HZ
    g = _grid((nx, ny, nz), 20e-9, req["steps"], 10, f)
    s0 = nz * 5 // 12
    s1 = nz * 7 // 12
    kind = req["kind"]
    if kind == "lorentz":
        mat = material_lorentz(f)
        # the reference's Drude-fit ADE needs Re eps < EPS_INF = 20
        # (fdtd.c:2478-2485); nearer the 1370 cm^-1 pole the slab runs as a
        # lossy dielectric (eps_im loss term), as a reference user would set it
        if mat["eps_re"][0] >= 20.0:
            mat = dict(mat, use_ade=0)
    else:
        if kind not in _LATTICE_CACHE:
            _LATTICE_CACHE[kind] = material_lattice4() if kind == "lattice4" else material_lattice16()
        mat = _LATTICE_CACHE[kind]
    src = [_source("dipole", (nx // 2, ny // 2, s0 - 5), "z", f, tsteps=req["steps"], dt=g["dt"])]
    spec = dict(name="req%d/%s/%dx%dx%d" % (req["id"], kind, nx, ny, nz), grid=g, materials=builtin_table(f) + [mat],
                sources=src, monitor=dict(x=nx // 2, y=ny // 2, z=s1 + 3, atom=0, dir=2))
    return spec, _draw(spec, dict(cubes=[_slab((nx, ny, nz), s0, s1)]))


def with_steps(spec, steps):
    spec = dict(spec)
    g = dict(spec["grid"])
    g["tsteps"] = int(steps)
    spec["grid"] = g
    spec["sources"] = [dict(s, end_time=max(s["end_time"], steps * g["dt"])) for s in spec["sources"]]
    return spec


def cell_counts(spec, med):
    """Cells per material and CPML incidences, for the algorithmic byte model."""
    counts = np.bincount(med.ravel(), minlength=len(spec["materials"]))
    g = spec["grid"]
    nx, ny, nz, L = g["nx"], g["ny"], g["nz"], g["pml_thickness"]
    pml = min(2 * L, nx) * ny * nz + min(2 * L, ny) * nx * nz + min(2 * L, nz) * nx * ny
    return counts, pml


def step_bytes(spec, med, word, rv_bytes=8, eform=None):
    """Algorithmic HBM bytes of one time step: each state word read and written
    once.  C (18w + 1) + 8w * CPML incidences + 9w * ADE cells
    + 12 N b * lattice cells, b the bytes per stored r or v value (8 for FP64,
    4.5 for compensated FP32: a word and a 4-bit code): 96 N or 54 N.  In E-form (the FP32 forms; the
    default for w = 4) plain cells store E and H alone: 12w + 1 instead of
    18w + 1 (their D is not stored)."""
    if eform is None:
        eform = word == 4
    counts, pml = cell_counts(spec, med)
    C = med.size
    b = C * (18 * word + 1) + 8 * word * pml
    for m, mat in enumerate(spec["materials"]):
        if counts[m] == 0:
            continue
        if eform and not mat.get("use_ade") and not mat.get("natom", 0):
            b -= 6 * word * counts[m]  # no D read or write
        if mat.get("use_ade"):
            b += 9 * word * counts[m]
        if mat.get("natom", 0):
            b += 12 * rv_bytes * mat["natom"] * counts[m]
    return float(b)


# --------------------------------------------------------------------------
# reference-format writers (for running the same workload through fdtd.c)
# --------------------------------------------------------------------------
def write_medium(path, med):
    _medium.save_medium_file(med, path)


def _fmt(v):
    return repr(float(v))


def write_reference_case(d, spec, med, output_interval=None, with_born=True):
    """Write case.cfg + medium_map.bin + material files under directory d so
    that the reference solver loads exactly the materials of ``spec``.

    Lattice materials are given a born= OUTCAR (needed for multi-rank runs,
    fdtd.c:2621-2704) and their eps through a one-row dielectric= table, which
    overrides the OUTCAR tensor (fdtd.c:2433-2441); switchxz materials get
    pre-permuted files so the reference's own permutation restores them."""
    os.makedirs(d, exist_ok=True)
    g = spec["grid"]
    write_medium(os.path.join(d, "medium_map.bin"), med)
    oi = output_interval or g["tsteps"] + 1
    lines = ["nx = %d" % g["nx"], "ny = %d" % g["ny"], "nz = %d" % g["nz"], "tsteps = %d" % g["tsteps"],
             "delta_x = %s" % _fmt(g["dx"]), "delta_y = %s" % _fmt(g["dy"]), "delta_z = %s" % _fmt(g["dz"]),
             "dt = %s" % _fmt(g["dt"]), "pml_thickness = %d" % g["pml_thickness"],
             "pml_reflection = %s" % _fmt(g["pml_reflection"]), "output_interval = %d" % oi,
             "output_component = Ez", "output_slice_axis = z", "output_slice_position = %d" % (g["nz"] // 2),
             "medium_file = medium_map.bin"]
    names = {0: "gaussian", 1: "sinusoidal", 2: "ricker", 3: "dipole"}
    for s in spec["sources"]:
        lines.append("source type=%s, x=%d, y=%d, z=%d, amplitude=%s, frequency=%s, t0=%s, tau=%s, component=%s" % (
            names[s["type"]], s["x"], s["y"], s["z"], _fmt(s["amplitude"]), _fmt(s["frequency"]), _fmt(s["t0"]),
            _fmt(s["tau"]), s["component"]))
    f = g["f_source"]
    for mid, m in enumerate(spec["materials"]):
        if mid < 4:
            continue
        parts = ["id=%d" % mid, "name=%s" % m.get("name", "m%d" % mid)]
        sw = "raw" in m
        re, im = list(m["eps_re"]), list(m["eps_im"])
        if m.get("natom", 0) or sw:
            # eps through a one-row 15-column table at the source frequency
            tre, tim = (switchxz3(re), switchxz3(im)) if sw else (re, im)
            tab = "%s 0 %s %s %s %s %s %s %s %s %s %s %s %s %s\n" % (
                _fmt(f), _fmt(f / WN_TO_HZ), _fmt(tre[0]), _fmt(tre[1]), _fmt(tre[2]), _fmt(tre[4]), _fmt(tre[5]),
                _fmt(tre[8]), _fmt(tim[0]), _fmt(tim[1]), _fmt(tim[2]), _fmt(tim[4]), _fmt(tim[5]), _fmt(tim[8]))
            fn = "eps_m%d.txt" % mid
            open(os.path.join(d, fn), "w").write(tab)
            parts.append("dielectric=%s" % fn)
        else:
            keys = ["xx", "xy", "xz", "yx", "yy", "yz", "zx", "zy", "zz"]
            parts += ["eps_%s=%s" % (k, _fmt(v)) for k, v in zip(keys, re)]
            parts += ["eps_%si=%s" % (k, _fmt(v)) for k, v in zip(keys, im)]
        if m.get("sigma"):
            parts.append("sigma=%s" % _fmt(m["sigma"]))
        if m.get("use_ade"):
            parts.append("use_ADE=1")
        if m.get("anglez"):
            parts.append("anglez=%s" % _fmt(m["anglez"]))
        if m.get("anglex"):
            parts.append("anglex=%s" % _fmt(m["anglex"]))
        n = m.get("natom", 0)
        if n:
            src = m["raw"] if sw else m
            for key in ("qezv", "qezm", "dmm"):
                fn = "%s_m%d.txt" % (key, mid)
                np.savetxt(os.path.join(d, fn), np.asarray(src[key]), fmt="%.17g")
            parts += ["natom=%d" % n, "QeZV=qezv_m%d.txt" % mid, "QeZM=qezm_m%d.txt" % mid, "DMM=dmm_m%d.txt" % mid]
            if with_born:
                Z = np.asarray(src["born"] if sw else m["born"]).reshape(n, 3, 3)
                out = [" MACROSCOPIC STATIC DIELECTRIC TENSOR (including local field effects)", " " + "-" * 54]
                for r in range(3):
                    out.append("  %12.6f %12.6f %12.6f" % tuple(re[r * 3:r * 3 + 3]))
                out += [" " + "-" * 54, "", " BORN EFFECTIVE CHARGES (including local field effects)", " " + "-" * 60]
                for a in range(n):
                    out.append(" ion %4d" % (a + 1))
                    for r in range(3):
                        out.append("    %d %11.5f %11.5f %11.5f" % ((r + 1,) + tuple(Z[a, r])))
                fn = "outcar_m%d.txt" % mid
                open(os.path.join(d, fn), "w").write("\n".join(out) + "\n")
                parts.append("born=%s" % fn)
            if sw:
                parts.append("switchxz=1")
        lines.append("material " + ", ".join(parts))
    mon = spec.get("monitor")
    if mon:
        lines.append("monitor_point = x=%d,y=%d,z=%d,atom=%d,dir=%s" % (mon["x"], mon["y"], mon["z"], mon["atom"],
                                                                       "xyz"[mon["dir"]]))
        lines.append("monitor_output_file = monitor_point.dat")
    text = "\n".join(lines) + "\n"
    if spec.get("medium"):
        # the geometry that draws medium_map.bin: python3 -m php_fdtd_amd.medium case.cfg
        text += "\n" + _medium.format_section(spec["medium"])
    open(os.path.join(d, "case.cfg"), "w").write(text)
    return os.path.join(d, "case.cfg")
