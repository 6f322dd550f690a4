"""Every wall kind of src/wall/wall_3d.rs on the device against the oracle, bit for bit: the wall
filter (kept generators), every cell's vertices, faces and neighbour ids, volumes and centroids.
Each case adds generators exactly on the wall's singular set (axis, tip, curve sample) so the
"no plane emitted" branches run too."""
import math

import numpy as np
import pytest

from conftest import assert_same_cell, bits, oracle_wall
from oracle import OracleTess
from vorothree_amd import (Algorithm3DGrid, Batch, BoundingBox, CatmullRomGeometry, ConeGeometry,
                           ConvexPolyhedronGeometry, CubicBezierGeometry, CylinderGeometry, SphereGeometry,
                           Tessellation, TorusGeometry, TrefoilKnotGeometry, Wall)
from vorothree_amd import workloads as W

pytestmark = pytest.mark.gpu

UNIT = BoundingBox((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))


def walls_for(name):
    if name == "cylinder":
        return [CylinderGeometry((0.5, 0.5, 0.0), (0.0, 0.0, 1.0), 0.3)], [[0.5, 0.5, 0.4]]
    if name == "cylinder-tilted":
        return [CylinderGeometry((0.5, 0.5, 0.5), (1.0, 1.0, 0.3), 0.25)], [[0.5, 0.5, 0.5]]
    if name == "cone":
        return [ConeGeometry((0.5, 0.5, 0.05), (0.0, 0.1, 1.0), 0.4)], [[0.5, 0.5, 0.05]]
    if name == "torus":
        return [TorusGeometry((0.5, 0.5, 0.5), (0.0, 0.0, 1.0), 0.28, 0.13)], [[0.5, 0.5, 0.5], [0.78, 0.5, 0.5]]
    if name == "hexahedron":
        return [ConvexPolyhedronGeometry.new_hexahedron((0.5, 0.5, 0.5), 0.6)], []
    if name == "tetrahedron":
        return [ConvexPolyhedronGeometry.new_tetrahedron((0.5, 0.5, 0.5), 0.48)], []
    if name == "trefoil":
        k = TrefoilKnotGeometry((0.5, 0.5, 0.5), 0.12, 0.07, 120)
        return [k], [list(k.samples[7])]
    if name == "bezier-open":
        b = CubicBezierGeometry((0.1, 0.2, 0.3), (0.4, 0.95, 0.2), (0.8, 0.05, 0.8), (0.9, 0.8, 0.3), 0.09, 64, False)
        return [b], [list(b.samples[20])]
    if name == "catmull-closed":
        c = CatmullRomGeometry([[0.2, 0.2, 0.5], [0.8, 0.25, 0.4], [0.75, 0.8, 0.6], [0.25, 0.75, 0.5]], 0.08, 80, True)
        return [c], [list(c.samples[33])]
    if name == "mixed":
        return [CylinderGeometry((0.5, 0.5, 0.5), (0.0, 1.0, 0.0), 0.45),
                ConvexPolyhedronGeometry.new_hexahedron((0.5, 0.5, 0.5), 0.75),
                SphereGeometry((0.5, 0.5, 0.5), 0.5)], []
    raise KeyError(name)


CASES = ["cylinder", "cylinder-tilted", "cone", "torus", "hexahedron", "tetrahedron", "trefoil", "bezier-open",
         "catmull-closed", "mixed"]


@pytest.mark.parametrize("name", CASES)
def test_wall_kind_bit_exact(name):
    geos, extra = walls_for(name)
    pts = W.PCG32(930 + CASES.index(name)).uniform(6000, (0, 0, 0), (1, 1, 1))
    if extra:
        pts = np.concatenate([np.asarray(extra, dtype=np.float64), pts])
    walls = [Wall(-1000 - k, g) for k, g in enumerate(geos)]
    t = Tessellation(UNIT, Algorithm3DGrid(16, 16, 16, UNIT))
    for w in walls:
        t.add_wall(w)
    t.set_generators(pts)
    t.calculate()
    o = OracleTess((0, 0, 0), (1, 1, 1), 16, 16, 16)
    for w in walls:
        oracle_wall(o, w)
    o.set_generators(pts)
    o.calculate(threads=8)
    n = o.count_generators()
    assert t.count_cells() == n and n > 100
    assert np.array_equal(bits(t.generators()), bits(o.generators()))
    vol, cen, nf, _ = o.reduced()
    assert np.array_equal(bits(t.volumes()), bits(vol))
    assert np.array_equal(bits(t.centroids()), bits(cen))
    assert np.array_equal(t.face_counts(), nf)
    off, ids = t.neighbors()
    for i in range(n):
        assert np.array_equal(ids[off[i]:off[i + 1]], o.get_cell(i).face_neighbors), i
    for i in range(0, n, 53):
        assert_same_cell(t.get_cell(i), o.get_cell(i), i)
    assert set(ids[ids <= -1000]) <= {w.id() for w in walls}


def test_batch_members_with_every_wall_kind():
    names = ["cylinder", "cone", "torus", "tetrahedron", "trefoil", "catmull-closed"]
    b = Batch(keep_cells=True)
    cases = []
    for k, name in enumerate(names):
        geos, extra = walls_for(name)
        pts = W.PCG32(960 + k).uniform(2500 + 500 * k, (0, 0, 0), (1, 1, 1))
        walls = [Wall(-1000 - j, g) for j, g in enumerate(geos)]
        b.add(UNIT, (12, 12, 12), walls, pts)
        cases.append((walls, pts))
    b.calculate()
    for t, (walls, pts) in enumerate(cases):
        o = OracleTess((0, 0, 0), (1, 1, 1), 12, 12, 12)
        for w in walls:
            oracle_wall(o, w)
        o.set_generators(pts)
        o.calculate(threads=8)
        vol, cen, _, _ = o.reduced()
        assert np.array_equal(bits(b.volumes(t)), bits(vol)), names[t]
        assert np.array_equal(bits(b.centroids(t)), bits(cen)), names[t]
        off, ids = b.neighbors(t)
        for i in range(0, len(vol), 9):
            assert np.array_equal(ids[off[i]:off[i + 1]], o.get_cell(i).face_neighbors), (names[t], i)


def test_tube_wall_under_graph_replay():
    geos, _ = walls_for("trefoil")
    w = Wall(-1000, geos[0])
    g = Tessellation(UNIT, Algorithm3DGrid(14, 14, 14, UNIT), keep_cells=False, graphs=True)
    e = Tessellation(UNIT, Algorithm3DGrid(14, 14, 14, UNIT), keep_cells=False, graphs=False)
    for t in (g, e):
        t.add_wall(w)
    pts = W.PCG32(975).uniform(20000, (0, 0, 0), (1, 1, 1))
    for t in (g, e):
        t.set_generators(pts)
        t.calculate()
    for _ in range(3):
        for t in (g, e):
            t.calculate()
        assert np.array_equal(bits(g.volumes()), bits(e.volumes()))
        assert np.array_equal(bits(g.centroids()), bits(e.centroids()))
    assert g.stats()[8] >= 1


# --- geometries of the caller's, through vd_tess_add_wall_planes_per_generator ------------------

def _compare(g, o):
    n = o.count_cells()
    assert g.count_cells() == n
    assert np.array_equal(bits(g.generators()), bits(o.generators()))
    for i in range(n):
        assert_same_cell(g.get_cell(i), o.get_cell(i), i)
    vol, cen, nf, _ = o.reduced()
    assert np.array_equal(bits(g.volumes()), bits(vol))
    assert np.array_equal(bits(g.centroids()), bits(cen))


def test_host_sphere_equals_builtin_and_oracle():
    from host_geometries import PySphere
    pts = W.PCG32(1301).uniform(6000, (0, 0, 0), (1, 1, 1))
    host = Tessellation(UNIT, Algorithm3DGrid(16, 16, 16, UNIT))
    builtin = Tessellation(UNIT, Algorithm3DGrid(16, 16, 16, UNIT))
    o = OracleTess((0, 0, 0), (1, 1, 1), 16, 16, 16)
    host.add_wall(Wall(-1000, PySphere((0.5, 0.5, 0.5), 0.45)))
    builtin.add_wall(Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), 0.45)))
    o.add_wall_host(-1000, PySphere((0.5, 0.5, 0.5), 0.45))
    for t in (host, builtin, o):
        t.set_generators(pts)
    for step in range(2):
        host.calculate_sealed()
        builtin.calculate_sealed()
        o.calculate_sealed(threads=4)
        _compare(host, o)
        _compare(builtin, o)
        assert np.array_equal(host.seal_log(), o.seal_log()) and len(o.seal_log()) > 0
        assert np.array_equal(builtin.seal_log(), o.seal_log())
        host.relax()
        builtin.relax()
        o.relax()


@pytest.mark.parametrize("walls_first", [True, False])
def test_host_wall_with_varying_plane_counts_bit_exact(walls_first):
    # 1 to 8 planes per generator, beside a built-in curved wall; sealing clips by every plane of
    # the neighbour's cut
    from host_geometries import PyNearFacets
    pts = W.PCG32(1302 + int(walls_first)).uniform(8000, (0, 0, 0), (1, 1, 1))
    g = Tessellation(UNIT, Algorithm3DGrid(18, 18, 18, UNIT))
    o = OracleTess((0, 0, 0), (1, 1, 1), 18, 18, 18)

    def walls():
        g.add_wall(Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), 0.49)))
        o.add_wall_sphere(-1000, (0.5, 0.5, 0.5), 0.49)
        g.add_wall(Wall(-1001, PyNearFacets((0.5, 0.5, 0.5), 0.6)))
        o.add_wall_host(-1001, PyNearFacets((0.5, 0.5, 0.5), 0.6))

    if walls_first:
        walls()
    g.set_generators(pts)
    o.set_generators(pts)
    if not walls_first:
        walls()
    g.calculate_sealed()
    o.calculate_sealed(threads=4)
    _compare(g, o)
    assert np.array_equal(g.seal_log(), o.seal_log())
    g.relax()
    o.relax()
    g.calculate()
    o.calculate(threads=4)
    _compare(g, o)
