import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def read_generators(path):
    """Parser of import_generators (src/tessellation.rs:126-148): `id x y z`, id ignored,
    unparsable or missing coordinates become 0.0."""
    pts = []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if not parts:
                continue
            vals = []
            for k in range(1, 4):
                try:
                    vals.append(float(parts[k]))
                except (IndexError, ValueError):
                    vals.append(0.0)
            pts.append(vals)
    return np.asarray(pts, dtype=np.float64)


def read_voropp_vertices(path):
    cells = {}
    cur = None
    with open(path) as fh:
        for line in fh:
            line = line.strip()
            if not line:
                continue
            if line.startswith("Cell"):
                cur = int(line.split()[1].rstrip(":"))
                cells[cur] = []
            elif cur is not None:
                vals = [float(x) for x in line.split()]
                if len(vals) == 3:
                    cells[cur].append(vals)
    return {k: np.asarray(v) for k, v in cells.items()}


def bits(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.uint64)


def assert_same_cell(gpu_cell, ora_cell, i):
    """Bit-exact equality of the four Cell3DFaces arrays."""
    assert gpu_cell is not None and ora_cell is not None, i
    assert np.array_equal(gpu_cell.face_counts(), ora_cell.face_counts), f"cell {i}: face counts"
    assert np.array_equal(gpu_cell.face_neighbors(), ora_cell.face_neighbors), f"cell {i}: neighbours"
    assert np.array_equal(gpu_cell.face_indices(), ora_cell.face_indices), f"cell {i}: face indices"
    assert np.array_equal(bits(gpu_cell.vertices()), bits(ora_cell.vertices)), f"cell {i}: vertices"


def oracle_wall(o, wall):
    """Adds a vorothree_amd Wall to an OracleTess with the same geometry parameters."""
    g = wall.geometry
    name = type(g).__name__
    wid = wall.id()
    if name == "PlaneGeometry":
        o.add_wall_plane(wid, g.point, g.normal_raw)
    elif name == "SphereGeometry":
        o.add_wall_sphere(wid, g.center, g.radius)
    elif name == "CylinderGeometry":
        o.add_wall(wid, 2, g.center, g.axis_raw, g.radius)
    elif name == "ConeGeometry":
        o.add_wall(wid, 3, g.tip, g.axis_raw, g.angle)
    elif name == "TorusGeometry":
        o.add_wall(wid, 4, g.center, g.axis_raw, g.major_radius, g.minor_radius)
    elif name == "ConvexPolyhedronGeometry":
        o.add_wall(wid, 5, pts=g.planes)
    else:
        o.add_wall(wid, 6, r=g.tube_radius, pts=g.samples, closed=g.closed)
