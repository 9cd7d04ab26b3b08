"""Medium maps: the ``[medium]`` geometry of a cfg file and ``medium_map.bin``.

The solver's material input is an int32 id per cell (0 = vacuum).  The
reference draws it from the ``[medium]`` section of the cfg with
generate_medium5.py; this module builds the same array from the same text:

* ``cubes``      inclusive index boxes, clipped to the grid
                 (generate_medium5.py:49-74);
* ``hexahedra``  parallelepipeds given by 8 vertices; a cell belongs to one
                 when its affine coordinates in the frame v0; v1-v0, v3-v0,
                 v4-v0 lie in [-1e-9, 1 + 1e-9] and it sits inside the
                 vertices' bounding box (generate_medium5.py:155-209);
* ``cylinders``  end points p1, p2 in cell indices, radius in metres, cells
                 measured in metres through delta_x/y/z
                 (generate_medium5.py:272-346);
* ``cones``      (no counterpart in the reference) a frustum between p1 and
                 p2 with radius r1 at p1 and r2 at p2, in metres; with
                 r1 == r2 it is the cylinder, cell for cell.  The
                 1024x1024x640 tip workload needs it.

Bodies are drawn in that order, each overwriting what came before
(generate_medium5.py:77-84).  ``medium_map.bin`` is the int32 header
nx, ny, nz followed by the ids with x slowest and z fastest, native byte
order (generate_medium5.py:348-372; README.md:520-535).

Command line, as the reference's ``python3 generate_medium5.py cfg [out]``
(generate_medium5.py:374-389)::

    python3 -m php_fdtd_amd.medium simulation.cfg [medium_map.bin]
"""
import ast
import os
import re
import sys

import numpy as np

DEFAULT_SHAPE = (50, 50, 50)        # generate_medium5.py:10
DEFAULT_SPACING = (1.0, 1.0, 1.0)   # generate_medium5.py:11
HEX_TOL = 1e-9                      # generate_medium5.py:197
KINDS = ("cubes", "hexahedra", "cylinders", "cones")


def _note(msg):
    print("[medium] " + msg, file=sys.stderr)


# --------------------------------------------------------------------------
# cfg text
# --------------------------------------------------------------------------
def grid_of(text):
    """(nx, ny, nz), (dx, dy, dz) of a cfg: the first ``key = value`` match of
    each key anywhere in the text, defaults 50 cells and 1.0
    (generate_medium5.py:14-31)."""
    shape = list(DEFAULT_SHAPE)
    for a, key in enumerate(("nx", "ny", "nz")):
        m = re.search(key + r"\s*=\s*(\d+)", text)
        if m:
            shape[a] = int(m.group(1))
    spacing = list(DEFAULT_SPACING)
    for a, key in enumerate(("delta_x", "delta_y", "delta_z")):
        m = re.search(key + r"\s*=\s*([\d.eE+-]+)", text)
        if m:
            spacing[a] = float(m.group(1))
    return tuple(shape), tuple(spacing)


def medium_section(text):
    """Body of the ``[medium]`` section (up to the next line opening with '['),
    or None (generate_medium5.py:40-44)."""
    m = re.search(r"\[medium\](.*?)(?=\n\[|\Z)", text, re.DOTALL | re.IGNORECASE)
    return m.group(1).strip() if m else None


_PAIR = r"\(\s*(\d+)\s*,\s*(\d+)\s*\)"
_CUBE = re.compile(r"\{\s*" + r",\s*".join(r'"%s"\s*:\s*%s' % (k, _PAIR) for k in ("xrange", "yrange", "zrange")) +
                   r',\s*"material"\s*:\s*(\d+)\s*\}')


def parse_cubes(section):
    """Cubes of ``cubes = [...]``: the list runs to its first ']'; comments
    inside it are dropped; each entry must spell its keys in the order xrange,
    yrange, zrange, material with non-negative integers; anything else is
    skipped silently (generate_medium5.py:49-59)."""
    m = re.search(r"cubes\s*=\s*\[(.*?)\]", section, re.DOTALL)
    if not m:
        return []
    body = re.sub(r"\s+", " ", re.sub(r"#.*", "", m.group(1).strip()).replace("\n", " "))
    out = []
    for c in _CUBE.finditer(body):
        v = [int(g) for g in c.groups()]
        out.append(dict(xrange=(v[0], v[1]), yrange=(v[2], v[3]), zrange=(v[4], v[5]), material=v[6]))
    return out


def _bracket_list(section, keyword):
    """The Python literal list after ``keyword =`` (comments dropped, keyword
    case-insensitive, brackets balanced), or [] when absent or unreadable
    (generate_medium5.py:88-128, 211-251)."""
    text = re.sub(r"#.*", "", section)
    m = re.search(keyword + r"\s*=", text, re.IGNORECASE)
    if not m:
        return []
    rest = text[m.end():]
    start = rest.find("[")
    if start < 0:
        return []
    depth = 0
    for i in range(start, len(rest)):
        if rest[i] == "[":
            depth += 1
        elif rest[i] == "]":
            depth -= 1
            if depth == 0:
                break
    else:
        _note("unmatched brackets in the %s list" % keyword)
        return []
    try:
        return ast.literal_eval(rest[start:i + 1])
    except Exception as e:  # a malformed list draws nothing, as in the reference
        _note("cannot read the %s list: %s" % (keyword, e))
        return []


def parse_hexahedra(section):
    """Entries with 'vertices' (8 points) and 'material'; others are skipped
    with a note (generate_medium5.py:130-151)."""
    out = []
    for item in _bracket_list(section, "hexahedra"):
        if not isinstance(item, dict) or "vertices" not in item or "material" not in item:
            _note("skipping hexahedron %r" % (item,))
            continue
        if len(item["vertices"]) != 8:
            _note("skipping hexahedron with %d vertices" % len(item["vertices"]))
            continue
        try:
            verts = [tuple(float(c) for c in v) for v in item["vertices"]]
        except (TypeError, ValueError) as e:
            _note("skipping hexahedron: %s" % e)
            continue
        out.append(dict(vertices=verts, material=item["material"]))
    return out


def _axis_items(section, keyword, radius_keys):
    out = []
    for item in _bracket_list(section, keyword):
        need = ("p1", "p2") + radius_keys + ("material",)
        if not isinstance(item, dict) or any(k not in item for k in need):
            _note("skipping %s entry %r" % (keyword, item))
            continue
        d = dict(p1=tuple(float(c) for c in item["p1"]), p2=tuple(float(c) for c in item["p2"]),
                 material=int(item["material"]))
        for k in radius_keys:
            d[k] = float(item[k])
        out.append(d)
    return out


def parse_cylinders(section):
    """Entries with p1, p2, radius, material (generate_medium5.py:253-269)."""
    return _axis_items(section, "cylinders", ("radius",))


def parse_cones(section):
    """Entries with p1, p2, r1, r2, material (frusta; no reference
    counterpart)."""
    return _axis_items(section, "cones", ("r1", "r2"))


def parse_geometry(section):
    if section is None:
        return {k: [] for k in KINDS}
    return dict(cubes=parse_cubes(section), hexahedra=parse_hexahedra(section),
                cylinders=parse_cylinders(section), cones=parse_cones(section))


def format_section(geometry):
    """``[medium]`` text that parses back to ``geometry`` (one line per list)."""
    lines = ["[medium]"]
    cubes = geometry.get("cubes", [])
    if cubes:
        lines.append("cubes = [ " + ", ".join(
            '{"xrange": (%d, %d), "yrange": (%d, %d), "zrange": (%d, %d), "material": %d}' % (
                c["xrange"] + c["yrange"] + c["zrange"] + (c["material"],)) for c in cubes) + " ]")

    def pt(p):
        return "(" + ", ".join(repr(float(v)) for v in p) + ")"

    if geometry.get("hexahedra"):
        lines.append("hexahedra = [ " + ", ".join(
            '{"vertices": [%s], "material": %d}' % (", ".join(pt(v) for v in h["vertices"]), h["material"])
            for h in geometry["hexahedra"]) + " ]")
    if geometry.get("cylinders"):
        lines.append("cylinders = [ " + ", ".join(
            '{"p1": %s, "p2": %s, "radius": %r, "material": %d}' % (pt(c["p1"]), pt(c["p2"]), float(c["radius"]),
                                                                   c["material"]) for c in geometry["cylinders"]) + " ]")
    if geometry.get("cones"):
        lines.append("cones = [ " + ", ".join(
            '{"p1": %s, "p2": %s, "r1": %r, "r2": %r, "material": %d}' % (
                pt(c["p1"]), pt(c["p2"]), float(c["r1"]), float(c["r2"]), c["material"]) for c in geometry["cones"]) +
                     " ]")
    return "\n".join(lines) + "\n"


# --------------------------------------------------------------------------
# drawing
# --------------------------------------------------------------------------
def fill_cube(med, cube):
    """Inclusive index box clipped to the grid (generate_medium5.py:67-74)."""
    lo = [max(r[0], 0) for r in (cube["xrange"], cube["yrange"], cube["zrange"])]
    hi = [min(r[1] + 1, n) for r, n in zip((cube["xrange"], cube["yrange"], cube["zrange"]), med.shape)]
    if all(a < b for a, b in zip(lo, hi)):
        med[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = cube["material"]
        return int(np.prod([b - a for a, b in zip(lo, hi)]))
    return 0


def _box_points(lo, hi):
    """Index grids (X, Y, Z) of the inclusive box lo..hi, empty if inverted."""
    return np.meshgrid(*[np.arange(a, b + 1) for a, b in zip(lo, hi)], indexing="ij")


def fill_hexahedron(med, hexa):
    """Cells whose affine coordinates (a, b, c), p = v0 + a u + b v + c w with
    u = v1 - v0, v = v3 - v0, w = v4 - v0, lie in [-tol, 1 + tol], searched
    over the vertices' bounding box (floor of the minima, ceil of the maxima,
    clipped to the grid).  A singular frame draws nothing
    (generate_medium5.py:155-209)."""
    V = np.asarray(hexa["vertices"], dtype=np.float64)
    lo = [max(0, int(np.floor(V[:, a].min()))) for a in range(3)]
    hi = [min(n - 1, int(np.ceil(V[:, a].max()))) for a, n in enumerate(med.shape)]
    frame = np.column_stack((V[1] - V[0], V[3] - V[0], V[4] - V[0]))
    try:
        inv = np.linalg.inv(frame)
    except np.linalg.LinAlgError:
        _note("degenerate hexahedron (singular frame), skipped")
        return 0
    X, Y, Z = _box_points(lo, hi)
    pts = np.column_stack((X.ravel(), Y.ravel(), Z.ravel()))
    abc = inv @ (pts - V[0]).T
    inside = np.all((abc >= -HEX_TOL) & (abc <= 1 + HEX_TOL), axis=0).reshape(X.shape)
    med[X[inside], Y[inside], Z[inside]] = hexa["material"]
    return int(inside.sum())


def _fill_axis_body(med, p1, p2, r1, r2, spacing, material):
    """Cells of the body around the axis p1 -> p2 (cell indices) whose radius
    runs linearly from r1 at p1 to r2 at p2 (metres).  Cell centres, ends and
    the axis are taken to metres through the spacing; a cell belongs when its
    projection t on the unit axis satisfies 0 <= t <= |axis| and its squared
    distance from the axis is at most radius(t)^2.  The search box is the end
    points' box grown by max(r1, r2) / spacing per axis, floor / ceil, clipped
    to the grid (generate_medium5.py:272-346 for r1 == r2)."""
    d = np.asarray(spacing, dtype=np.float64)
    a = np.asarray(p1, dtype=np.float64)
    b = np.asarray(p2, dtype=np.float64)
    a_m, b_m = a * d, b * d
    axis = b_m - a_m
    length = np.linalg.norm(axis)
    if length < 1e-12:
        _note("zero-length axis, skipped")
        return 0
    unit = axis / length
    rmax = max(r1, r2)
    lo = [max(0, int(np.floor(min(a[k], b[k]) - rmax / d[k]))) for k in range(3)]
    hi = [min(n - 1, int(np.ceil(max(a[k], b[k]) + rmax / d[k]))) for k, n in enumerate(med.shape)]
    X, Y, Z = _box_points(lo, hi)
    ux = X * d[0] - a_m[0]
    uy = Y * d[1] - a_m[1]
    uz = Z * d[2] - a_m[2]
    t = ux * unit[0] + uy * unit[1] + uz * unit[2]
    ox = ux - t * unit[0]
    oy = uy - t * unit[1]
    oz = uz - t * unit[2]
    off2 = ox ** 2 + oy ** 2 + oz ** 2
    # r1 == r2 gives rad == r1 exactly (r1 + 0.0), the cylinder's test
    rad = r1 + (r2 - r1) * (t / length)
    inside = (t >= 0.0) & (t <= length) & (off2 <= rad * rad)
    med[X[inside], Y[inside], Z[inside]] = material
    return int(inside.sum())


def fill_cylinder(med, cyl, spacing=DEFAULT_SPACING):
    return _fill_axis_body(med, cyl["p1"], cyl["p2"], cyl["radius"], cyl["radius"], spacing, cyl["material"])


def fill_cone(med, cone, spacing=DEFAULT_SPACING):
    return _fill_axis_body(med, cone["p1"], cone["p2"], cone["r1"], cone["r2"], spacing, cone["material"])


def build(shape, spacing, geometry):
    """int32 medium of ``shape`` with cubes, hexahedra, cylinders, cones drawn
    in that order over vacuum."""
    med = np.zeros(tuple(int(n) for n in shape), dtype=np.int32)
    for c in geometry.get("cubes", []):
        fill_cube(med, c)
    for h in geometry.get("hexahedra", []):
        fill_hexahedron(med, h)
    for c in geometry.get("cylinders", []):
        fill_cylinder(med, c, spacing)
    for c in geometry.get("cones", []):
        fill_cone(med, c, spacing)
    return med


def read_config(config):
    """Medium of a cfg file (path) or cfg text: grid from its keys, bodies from
    its ``[medium]`` section; all vacuum without one
    (generate_medium5.py:8-86)."""
    text = config
    if "\n" not in config and os.path.exists(config):
        with open(config) as f:
            text = f.read()
    shape, spacing = grid_of(text)
    return build(shape, spacing, parse_geometry(medium_section(text)))


# --------------------------------------------------------------------------
# medium_map.bin
# --------------------------------------------------------------------------
def save_medium_file(med, path):
    """int32 nx, ny, nz then the ids, x slowest, z fastest
    (generate_medium5.py:361-366)."""
    med = np.asarray(med)
    if med.ndim != 3:
        raise ValueError("medium must be 3-D, got shape %r" % (med.shape,))
    with open(path, "wb") as f:
        f.write(np.asarray(med.shape, dtype=np.int32).tobytes())
        f.write(np.ascontiguousarray(med, dtype=np.int32).tobytes())


def load_medium_file(path, shape=None):
    """The int32 (nx, ny, nz) array of a medium_map.bin; raises ValueError when
    the size disagrees with the header or the header with ``shape`` (the
    solver's own check, fdtd.c:718-959)."""
    raw = np.fromfile(path, dtype=np.int32)
    if raw.size < 3:
        raise ValueError("%s: no medium header" % path)
    dims = tuple(int(v) for v in raw[:3])
    if min(dims) <= 0 or raw.size != 3 + dims[0] * dims[1] * dims[2]:
        raise ValueError("%s: header %r does not match %d payload ints" % (path, dims, raw.size - 3))
    if shape is not None and tuple(shape) != dims:
        raise ValueError("%s: medium is %r, the grid is %r" % (path, dims, tuple(shape)))
    return raw[3:].reshape(dims)


def summary(med):
    ids, counts = np.unique(med, return_counts=True)
    n = med.size
    return "\n".join("  material %d: %d cells (%.2f%%)" % (i, c, 100.0 * c / n) for i, c in zip(ids, counts))


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if not argv:
        print("usage: python3 -m php_fdtd_amd.medium <config_file> [output_file]")
        return 1
    cfg = argv[0]
    out = argv[1] if len(argv) > 1 else "medium_map.bin"
    if not os.path.exists(cfg):
        print("Error: Config file '%s' not found" % cfg)
        return 1
    med = read_config(cfg)
    save_medium_file(med, out)
    print("medium %dx%dx%d -> %s (%d bytes)" % (med.shape + (out, os.path.getsize(out))))
    print(summary(med))
    return 0


if __name__ == "__main__":
    sys.exit(main())
