"""Medium generator and medium_map.bin (generate_medium5.py restated in
php_fdtd_amd.medium).

The bodies are checked against exact rational arithmetic, not against a
second copy of the float formulas: a hexahedron by Cramer's rule over
Fractions, cylinders and cones by the square-root-free form of the
projection test (0 <= u.v <= |v|^2, |u|^2 |v|^2 - (u.v)^2 <= r(t)^2 |v|^2).
Each case first checks that no lattice point lies within 1e-6 of a boundary
it does not sit on exactly, so float rounding cannot decide any cell and the
comparison is exact.  Where cells do sit exactly on a boundary the inputs are
chosen so the float path is exact too (integer coordinates, unit spacing), and
the reference's inclusive tests (>=, <=, the 1e-9 hexahedron tolerance) are
pinned by those cells.
"""
import os
import struct
import subprocess
import sys
from fractions import Fraction as Fr

import numpy as np
import pytest

from php_fdtd_amd import medium as M
from php_fdtd_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def cfg(grid, body, spacing=None):
    lines = ["[grid]"] + ["%s = %d" % (k, n) for k, n in zip(("nx", "ny", "nz"), grid)]
    if spacing:
        lines += ["%s = %r" % (k, d) for k, d in zip(("delta_x", "delta_y", "delta_z"), spacing)]
    return "\n".join(lines) + "\n\n[medium]\n" + body + "\n\n[source]\nsource type=dipole, x=1, y=1, z=1\n"


# --------------------------------------------------------------------------
# medium_map.bin
# --------------------------------------------------------------------------
def test_file_format_offsets(tmp_path):
    """12 + 4 ((i ny + j) nz + k) is cell (i, j, k) (README.md:520-535)."""
    rng = np.random.default_rng(1)
    med = rng.integers(0, 9, size=(3, 4, 5)).astype(np.int32)
    p = tmp_path / "m.bin"
    M.save_medium_file(med, str(p))
    raw = p.read_bytes()
    assert len(raw) == 12 + 4 * 60
    assert struct.unpack("iii", raw[:12]) == (3, 4, 5)
    for i, j, k in [(0, 0, 0), (2, 3, 4), (1, 2, 3), (2, 0, 1), (0, 3, 0)]:
        off = 12 + 4 * ((i * 4 + j) * 5 + k)
        assert struct.unpack("i", raw[off:off + 4])[0] == med[i, j, k]
    back = M.load_medium_file(str(p), shape=(3, 4, 5))
    assert back.dtype == np.int32 and np.array_equal(back, med)


def test_file_format_errors(tmp_path):
    med = np.ones((2, 3, 4), np.int32)
    p = tmp_path / "m.bin"
    M.save_medium_file(med, str(p))
    with pytest.raises(ValueError, match="grid"):
        M.load_medium_file(str(p), shape=(2, 3, 5))
    p.write_bytes(p.read_bytes()[:-4])
    with pytest.raises(ValueError, match="header"):
        M.load_medium_file(str(p))
    with pytest.raises(ValueError):
        M.save_medium_file(np.ones((2, 3), np.int32), str(tmp_path / "x.bin"))


# --------------------------------------------------------------------------
# cfg parsing and cubes
# --------------------------------------------------------------------------
def test_example_cube():
    # the cube line of the reference's examples/example.cfg on its 80x80x48 grid
    text = cfg((80, 80, 48), '# MoO3 block\ncubes = [ {"xrange": (30, 49), "yrange": (30, 49), "zrange": (16, 24), '
                             '"material": 4} ]\nmedium_file = ./medium_map.bin')
    med = M.read_config(text)
    assert med.shape == (80, 80, 48) and med.dtype == np.int32
    ref = np.zeros((80, 80, 48), np.int32)
    ref[30:50, 30:50, 16:25] = 4
    assert np.array_equal(med, ref) and int((med == 4).sum()) == 20 * 20 * 9


def test_cubes_order_clipping_comments():
    body = ('cubes = [    {"xrange": (6, 9), "yrange": (0, 11), "zrange": (0, 11), "material": 4},'
            '{"xrange": (0, 5), "yrange": (0, 11), "zrange": (0, 11), "material": 1},\n'
            '   # a comment line inside the list\n'
            '   {"xrange": (8, 30), "yrange": (2, 3), "zrange": (10, 99), "material": 7},\n'
            '   {"xrange": (40, 50), "yrange": (0, 1), "zrange": (0, 1), "material": 9},\n'
            '   {"yrange": (0, 1), "xrange": (0, 1), "zrange": (0, 1), "material": 8} ]')
    med = M.read_config(cfg((12, 12, 12), body))
    ref = np.zeros((12, 12, 12), np.int32)
    ref[6:10] = 4
    ref[0:6] = 1
    ref[8:12, 2:4, 10:12] = 7      # clipped to the grid, drawn over the first cube
    # the cube wholly outside draws nothing; keys out of order are not a cube
    assert np.array_equal(med, ref)
    assert not np.any(med == 8) and not np.any(med == 9)


def test_grid_keys_and_defaults():
    assert M.grid_of("nothing here") == ((50, 50, 50), (1.0, 1.0, 1.0))
    shape, spacing = M.grid_of("nx = 7\nny=8\n nz =  9\ndelta_x = 2e-8\ndelta_y=1.5E-8\ndelta_z = 1e-08\n")
    assert shape == (7, 8, 9) and spacing == (2e-8, 1.5e-8, 1e-8)
    assert not np.any(M.read_config("nx = 4\nny = 4\nnz = 4\n"))        # no [medium] section
    assert M.read_config("[medium]\n").shape == (50, 50, 50)
    # the section ends at the next line opening with '['
    text = cfg((6, 6, 6), "") + 'cubes = [ {"xrange": (0, 5), "yrange": (0, 5), "zrange": (0, 5), "material": 3} ]\n'
    assert not np.any(M.read_config(text))


# --------------------------------------------------------------------------
# hexahedra
# --------------------------------------------------------------------------
def hexa_exact(V, shape):
    """Cells p with p = v0 + a u + b v + c w, 0 <= a, b, c <= 1 (Cramer's rule
    over Fractions), inside the vertices' bounding box; and the smallest
    distance of any nonzero coordinate from 0 or 1."""
    V = [[Fr(c) for c in v] for v in V]
    u, v, w = ([V[i][k] - V[0][k] for k in range(3)] for i in (1, 3, 4))

    def det(a, b, c):
        return (a[0] * (b[1] * c[2] - b[2] * c[1]) - b[0] * (a[1] * c[2] - a[2] * c[1]) +
                c[0] * (a[1] * b[2] - a[2] * b[1]))

    D = det(u, v, w)
    lo = [max(0, int(np.floor(float(min(x[k] for x in V))))) for k in range(3)]
    hi = [min(n - 1, int(np.ceil(float(max(x[k] for x in V))))) for k, n in enumerate(shape)]
    inside = np.zeros(shape, bool)
    margin = 1.0
    for i in range(lo[0], hi[0] + 1):
        for j in range(lo[1], hi[1] + 1):
            for k in range(lo[2], hi[2] + 1):
                q = [Fr(i) - V[0][0], Fr(j) - V[0][1], Fr(k) - V[0][2]]
                abc = [det(q, v, w) / D, det(u, q, w) / D, det(u, v, q) / D]
                for x in abc:
                    for edge in (0, 1):
                        if x != edge:
                            margin = min(margin, abs(float(x - edge)))
                inside[i, j, k] = all(0 <= x <= 1 for x in abc)
    return inside, margin


def hexa_text(V, material):
    return "hexahedra = [ {\"vertices\": [%s], \"material\": %d} ]" % (
        ", ".join("(%s)" % ", ".join(repr(c) for c in v) for v in V), material)


def test_hexahedron_box_equals_cube():
    # the README's 4x3x3 box (README.md:481-486) moved by whole cells
    V = [(6, 6, 3), (9, 6, 3), (9, 8, 3), (6, 8, 3), (6, 6, 5), (9, 6, 5), (9, 8, 5), (6, 8, 5)]
    med = M.read_config(cfg((16, 16, 16), hexa_text(V, 6)))
    ref = np.zeros((16, 16, 16), np.int32)
    ref[6:10, 6:9, 3:6] = 6           # faces included: the 1e-9 tolerance
    assert np.array_equal(med, ref) and int((med == 6).sum()) == 36


@pytest.mark.parametrize("shear", [((5, 1, 0), (1, 4, 0), (1, 1, 5)), ((4, -2, 1), (2, 5, -1), (-1, 2, 6))])
def test_hexahedron_sheared_exact(shear):
    u, v, w = shear
    v0 = (6, 7, 5)

    def add(*ps):
        return tuple(sum(c) for c in zip(*ps))

    V = [v0, add(v0, u), add(v0, u, v), add(v0, v), add(v0, w), add(v0, u, w), add(v0, u, v, w), add(v0, v, w)]
    shape = (20, 20, 18)
    med = np.zeros(shape, np.int32)
    n = M.fill_hexahedron(med, dict(vertices=V, material=3))
    exact, margin = hexa_exact(V, shape)
    assert margin > 1e-6
    assert n == int(exact.sum()) > 0
    assert np.array_equal(med == 3, exact)


def test_hexahedron_bounding_box_quirk():
    """Only v0, v1, v3, v4 span the body, but all 8 vertices bound the search
    (generate_medium5.py:163-175, 180-183): vertices v2, v5, v6, v7 drawn
    inward cut the body to their box."""
    V = [(2, 2, 2), (10, 2, 2), (4, 4, 4), (2, 10, 2), (2, 2, 10), (4, 4, 4), (4, 4, 4), (4, 4, 4)]
    med = np.zeros((14, 14, 14), np.int32)
    M.fill_hexahedron(med, dict(vertices=V, material=2))
    exact, margin = hexa_exact(V, med.shape)
    assert margin > 1e-6 and np.array_equal(med == 2, exact)
    assert int(exact.sum()) == 9 * 9 * 9   # the 9^3 cube of the bounding box, not 9^3 of a larger body


def test_hexahedron_rejects():
    flat = [(1, 1, 1), (5, 1, 1), (5, 5, 1), (1, 5, 1), (1, 1, 1), (5, 1, 1), (5, 5, 1), (1, 5, 1)]
    assert not np.any(M.read_config(cfg((8, 8, 8), hexa_text(flat, 5))))          # singular frame
    seven = hexa_text(flat[:7], 5)
    assert not np.any(M.read_config(cfg((8, 8, 8), seven)))                      # 7 vertices
    assert M.parse_hexahedra('hexahedra = [ 3, {"material": 1} ]') == []         # not dicts / no vertices
    eight = "[" + ", ".join(["(0, 0, 0)"] * 7 + ["(1, 2, 3)"]) + "]"
    bad = "[" + ", ".join(["(0, 0, 0)"] * 7 + ["(0, 0, 'a')"]) + "]"
    assert M.parse_hexahedra('hexahedra = [ {"vertices": %s, "material": 1} ' % eight) == []  # unmatched
    assert M.parse_hexahedra('HEXAHEDRA = [ {"vertices": %s, "material": 1} ]' % bad) == []   # not numbers
    assert M.parse_hexahedra('hexahedra = [ {"vertices": [(0, 0, 0)] * 8, "material": 1} ]') == []  # not a literal
    ok = M.parse_hexahedra('Hexahedra = [ {"vertices": %s, "material": 1} ]  # note' % eight)
    assert len(ok) == 1 and ok[0]["vertices"][7] == (1.0, 2.0, 3.0)


# --------------------------------------------------------------------------
# cylinders and cones
# --------------------------------------------------------------------------
def axis_exact(p1, p2, r1, r2, spacing, shape):
    """Exact membership for the frustum r(t) = r1 + (r2 - r1) t / |v| (a
    cylinder when r1 == r2) with every quantity a Fraction of the float
    inputs, and the smallest relative distance of any cell from a boundary it
    does not sit on."""
    d = [Fr(x) for x in spacing]
    a = [Fr(x) * dk for x, dk in zip(p1, d)]
    b = [Fr(x) * dk for x, dk in zip(p2, d)]
    v = [bk - ak for ak, bk in zip(a, b)]
    vv = sum(x * x for x in v)
    R1, R2 = Fr(r1), Fr(r2)
    rmax = max(float(r1), float(r2))
    lo = [max(0, int(np.floor(min(p1[k], p2[k]) - rmax / spacing[k]))) for k in range(3)]
    hi = [min(n - 1, int(np.ceil(max(p1[k], p2[k]) + rmax / spacing[k]))) for k, n in enumerate(shape)]
    inside = np.zeros(shape, bool)
    margin = 1.0
    for i in range(lo[0], hi[0] + 1):
        for j in range(lo[1], hi[1] + 1):
            for k in range(lo[2], hi[2] + 1):
                u = [Fr(i) * d[0] - a[0], Fr(j) * d[1] - a[1], Fr(k) * d[2] - a[2]]
                uv = sum(x * y for x, y in zip(u, v))
                uu = sum(x * x for x in u)
                rad = R1 + (R2 - R1) * uv / vv
                lhs = uu * vv - uv * uv          # |v|^2 x squared distance from the axis
                rhs = rad * rad * vv
                for x, edge, scale in ((uv, 0, vv), (uv, vv, vv), (lhs, rhs, vv * rmax * rmax)):
                    if x != edge:
                        margin = min(margin, abs(float(x - edge)) / float(scale))
                inside[i, j, k] = 0 <= uv <= vv and lhs <= rhs
    return inside, margin


def test_cylinder_axis_aligned_gauss_count():
    """Axis along z, unit cells, radius 5: 81 lattice points of the disc
    i^2 + j^2 <= 25 (Gauss's circle count, the 12 on the circle included) on
    each of the 26 planes."""
    text = cfg((40, 40, 40), 'cylinders = [ {"p1": (20, 20, 5), "p2": (20, 20, 30), "radius": 5.0, "material": 6} ]')
    med = M.read_config(text)
    disc = sum(1 for i in range(-5, 6) for j in range(-5, 6) if i * i + j * j <= 25)
    assert disc == 81
    assert int((med == 6).sum()) == 81 * 26
    assert np.all(med[:, :, :5] == 0) and np.all(med[:, :, 31:] == 0)
    assert med[25, 20, 5] == 6 and med[23, 24, 30] == 6 and med[24, 24, 17] == 0


def test_cylinder_anisotropic_spacing():
    """Radius in metres over 20 x 10 x 10 nm cells (generate_medium5.py:272-346):
    the disc (20 i)^2 + (10 j)^2 <= 55^2 nm^2."""
    sp = (2e-8, 1e-8, 1e-8)
    text = cfg((30, 30, 20), 'cylinders = [ {"p1": (15, 15, 2), "p2": (15, 15, 12), "radius": 5.5e-8, "material": 2} ]',
               spacing=sp)
    med = M.read_config(text)
    disc = sum(1 for i in range(-4, 5) for j in range(-7, 8) if 400 * i * i + 100 * j * j <= 3025)
    assert int((med == 2).sum()) == disc * 11
    exact, margin = axis_exact((15, 15, 2), (15, 15, 12), 5.5e-8, 5.5e-8, sp, (30, 30, 20))
    assert margin > 1e-6 and np.array_equal(med == 2, exact)


@pytest.mark.parametrize("p1,p2,r", [((2.5, 3.5, 0.5), (12.5, 9.5, 15.5), 2.75),
                                     ((13.5, 2.5, 3.5), (3.5, 12.5, 14.5), 3.25)])
def test_cylinder_oblique_exact(p1, p2, r):
    shape = (18, 18, 20)
    med = np.zeros(shape, np.int32)
    n = M.fill_cylinder(med, dict(p1=p1, p2=p2, radius=r, material=5))
    exact, margin = axis_exact(p1, p2, r, r, (1.0, 1.0, 1.0), shape)
    assert margin > 1e-6
    assert n == int(exact.sum()) > 50 and np.array_equal(med == 5, exact)


def test_cylinder_rejects_and_parse():
    got = M.parse_cylinders('cylinders = [ {"p1": (212, 500, 506), "p2": (212, 700, 494), "radius": 120e-9, '
                            '"material": 6} ]')
    assert got == [dict(p1=(212.0, 500.0, 506.0), p2=(212.0, 700.0, 494.0), material=6, radius=120e-9)]
    assert M.parse_cylinders('cylinders = [ {"p1": (1, 1, 1), "radius": 1.0, "material": 6} ]') == []
    text = cfg((8, 8, 8), 'cylinders = [ {"p1": (4, 4, 4), "p2": (4, 4, 4), "radius": 2.0, "material": 6} ]')
    assert not np.any(M.read_config(text))                                      # zero length


@pytest.mark.parametrize("p1,p2,r1,r2", [((2.5, 3.5, 0.5), (12.5, 9.5, 15.5), 1.25, 4.5),
                                         ((9.5, 9.5, 1.5), (9.5, 9.5, 16.5), 0.75, 6.25),
                                         ((14.5, 3.5, 2.5), (4.5, 13.5, 15.5), 4.0, 0.5)])
def test_cone_exact(p1, p2, r1, r2):
    shape = (18, 18, 20)
    med = np.zeros(shape, np.int32)
    n = M.fill_cone(med, dict(p1=p1, p2=p2, r1=r1, r2=r2, material=7))
    exact, margin = axis_exact(p1, p2, r1, r2, (1.0, 1.0, 1.0), shape)
    assert margin > 1e-6
    assert n == int(exact.sum()) > 50 and np.array_equal(med == 7, exact)


def test_cone_with_equal_radii_is_the_cylinder():
    rng = np.random.default_rng(7)
    sp = (1e-8, 1.5e-8, 2e-8)
    for _ in range(6):
        p1 = tuple(rng.uniform(2, 28, 3))
        p2 = tuple(rng.uniform(2, 28, 3))
        r = float(rng.uniform(1e-8, 8e-8))
        a = np.zeros((32, 32, 32), np.int32)
        b = np.zeros((32, 32, 32), np.int32)
        M.fill_cylinder(a, dict(p1=p1, p2=p2, radius=r, material=3), sp)
        M.fill_cone(b, dict(p1=p1, p2=p2, r1=r, r2=r, material=3), sp)
        assert np.array_equal(a, b) and a.any()


def test_body_order():
    """cubes, then hexahedra, then cylinders, then cones, whatever the text
    order (generate_medium5.py:77-84)."""
    body = "\n".join([
        'cones = [ {"p1": (4, 4, 0), "p2": (4, 4, 7), "r1": 1.0, "r2": 1.0, "material": 4} ]',
        'cylinders = [ {"p1": (4, 4, 0), "p2": (4, 4, 7), "radius": 2.0, "material": 3} ]',
        hexa_text([(2, 2, 0), (6, 2, 0), (6, 6, 0), (2, 6, 0), (2, 2, 7), (6, 2, 7), (6, 6, 7), (2, 6, 7)], 2),
        'cubes = [ {"xrange": (0, 7), "yrange": (0, 7), "zrange": (0, 7), "material": 1} ]'])
    med = M.read_config(cfg((8, 8, 8), body))
    plane = med[:, :, 3]
    assert plane[0, 0] == 1 and plane[2, 6] == 2 and plane[2, 4] == 3 and plane[4, 4] == 4 and plane[5, 4] == 4
    assert plane[6, 4] == 3 and plane[4, 6] == 3 and plane[6, 6] == 2


def test_format_section_round_trip():
    geo = dict(cubes=[dict(xrange=(1, 5), yrange=(0, 9), zrange=(2, 3), material=4)],
               hexahedra=[dict(vertices=[(1.0, 1.0, 1.0), (4.0, 1.0, 1.0), (4.0, 3.5, 1.0), (1.0, 3.5, 1.0),
                                         (1.5, 1.0, 4.0), (4.5, 1.0, 4.0), (4.5, 3.5, 4.0), (1.5, 3.5, 4.0)],
                               material=2)],
               cylinders=[dict(p1=(2.0, 2.0, 1.0), p2=(7.0, 8.0, 9.0), radius=2.5e-8, material=6)],
               cones=[dict(p1=(5.0, 5.0, 2.0), p2=(5.0, 5.0, 9.0), r1=1e-8, r2=3.3e-8, material=5)])
    text = "nx = 10\nny = 10\nnz = 11\ndelta_x = 1e-8\ndelta_y = 1e-8\ndelta_z = 1e-8\n" + M.format_section(geo)
    assert M.parse_geometry(M.medium_section(text)) == geo
    assert np.array_equal(M.read_config(text), M.build((10, 10, 11), (1e-8, 1e-8, 1e-8), geo))


# --------------------------------------------------------------------------
# workloads, command line
# --------------------------------------------------------------------------
def test_workload_slabs():
    for build, n, z0, z1 in ((W.lorentz_slab, (100, 100, 100), 40, 60), (W.lattice4_slab, (160, 160, 120), 50, 70)):
        spec, med = build()
        ref = np.zeros(n, np.int32)
        ref[:, :, z0:z1] = 4
        assert np.array_equal(med, ref)
        assert spec["medium"]["cubes"][0]["zrange"] == (z0, z1 - 1)


def test_tip_cone_per_plane():
    """The 1024x1024x640 tip at 1/8 scale: on plane z the tip is the disc
    (x - cx)^2 + (y - cy)^2 <= (r0 + (z - apex) tan 30deg)^2 in cells; cells
    exactly on a circle (integer radius at the apex) may go either way."""
    spec, med = W.tip_slab(8)
    cone = spec["medium"]["cones"][0]
    cx, cy, apex = (int(c) for c in cone["p1"])
    top = int(cone["p2"][2]) + 1
    r0 = cone["r1"] / spec["grid"]["dx"]
    t = np.tan(np.radians(30.0))
    x = np.arange(med.shape[0])[:, None] - cx
    y = np.arange(med.shape[1])[None, :] - cy
    r2 = x * x + y * y
    for z in range(med.shape[2]):
        got = med[:, :, z] == 5
        if not apex <= z < top:
            assert not got.any()
            continue
        rad2 = (r0 + (z - apex) * t) ** 2
        near = np.abs(r2 - rad2) <= 1e-6 * rad2
        want = r2 <= rad2
        assert np.array_equal(got[~near], want[~near]), z
    slab = spec["medium"]["cubes"][0]["zrange"]
    assert np.all(med[:, :, slab[0]:slab[1] + 1] == 4)


def test_command_line(tmp_path):
    text = cfg((10, 12, 14), 'cubes = [ {"xrange": (1, 3), "yrange": (2, 8), "zrange": (0, 13), "material": 4} ]\n'
                             'cylinders = [ {"p1": (5, 6, 1), "p2": (5, 6, 12), "radius": 3e-8, "material": 2} ]',
               spacing=(1e-8, 1e-8, 1e-8))
    c = tmp_path / "sim.cfg"
    c.write_text(text)
    out = tmp_path / "medium_map.bin"
    r = subprocess.run([sys.executable, "-m", "php_fdtd_amd.medium", str(c), str(out)], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert np.array_equal(M.load_medium_file(str(out), (10, 12, 14)), M.read_config(str(c)))
    r = subprocess.run([sys.executable, "-m", "php_fdtd_amd.medium", str(tmp_path / "missing.cfg")], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "not found" in r.stdout


def test_reference_case_redraws_its_medium(tmp_path):
    spec, med = W.lorentz_slab()
    path = W.write_reference_case(str(tmp_path / "c0"), spec, med, with_born=False)
    assert np.array_equal(M.read_config(path), M.load_medium_file(str(tmp_path / "c0" / "medium_map.bin")))
