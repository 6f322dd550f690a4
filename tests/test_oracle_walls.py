"""The oracle's wall geometries (src/wall/wall_3d.rs) pinned by closed forms and independent
point-in-region tests:
* planar walls (convex polyhedra) clip exactly: the cell volumes sum to the polyhedron's volume
  (up to the rounding of the reference's origin-relative volume terms, ~1e-11 relative);
* curved walls clip every cell by the tangent plane at the generator's closest surface point, so
  the volumes sum to slightly more than the solid's closed-form volume (never less);
* the wall filter keeps exactly the points that an independent numpy membership test keeps;
* the sampled-curve constructors reproduce their control points where the curve interpolates.
"""
import math

import numpy as np
import pytest

from conftest import oracle_wall
from oracle import OracleTess
from vorothree_amd import (CatmullRomGeometry, ConeGeometry, ConvexPolyhedronGeometry, CubicBezierGeometry,
                           CylinderGeometry, TorusGeometry, TrefoilKnotGeometry, Wall)
from vorothree_amd import workloads as W

N = 30_000


def run(walls, n=N, seed=901, grid=24):
    pts = W.PCG32(seed).uniform(n, (0, 0, 0), (1, 1, 1))
    o = OracleTess((0, 0, 0), (1, 1, 1), grid, grid, grid)
    for w in walls:
        oracle_wall(o, w)
    o.set_generators(pts)
    o.calculate(threads=8)
    vol, _, _, _ = o.reduced()
    return pts, o, vol


def test_hexahedron_and_tetrahedron_volumes_exact():
    cube = ConvexPolyhedronGeometry.new_hexahedron((0.5, 0.5, 0.5), 0.3 * math.sqrt(3.0))
    pts, o, vol = run([Wall(-1000, cube)])
    inside = np.all((pts >= 0.2) & (pts <= 0.8), axis=1)
    assert o.count_generators() == int(inside.sum())
    assert abs(vol.sum() - 0.216) < 1e-10 * 0.216      # origin-relative volume terms round at 1e-11
    R = 0.45
    tet = ConvexPolyhedronGeometry.new_tetrahedron((0.5, 0.5, 0.5), R)
    _, _, vol = run([Wall(-1001, tet)], seed=902)
    a = R * math.sqrt(8.0 / 3.0)                     # edge of the regular tetrahedron
    exact = a ** 3 / (6.0 * math.sqrt(2.0))
    assert abs(vol.sum() - exact) < 1e-10 * exact


def test_cylinder_volume_and_filter():
    g = CylinderGeometry((0.5, 0.5, 0.0), (0.0, 0.0, 2.0), 0.3)      # axis normalised by the wall
    pts, o, vol = run([Wall(-1000, g)])
    keep = np.hypot(pts[:, 0] - 0.5, pts[:, 1] - 0.5) <= 0.3
    assert o.count_generators() == int(keep.sum())
    exact = math.pi * 0.09
    assert exact <= vol.sum() < exact * 1.01


def test_cone_volume_and_filter():
    ang, h = 0.35, 0.98
    g = ConeGeometry((0.5, 0.5, 0.02), (0.0, 0.0, 1.0), ang)
    pts, o, vol = run([Wall(-1000, g)])
    z = pts[:, 2] - 0.02
    keep = (z >= 0) & (np.hypot(pts[:, 0] - 0.5, pts[:, 1] - 0.5) <= z * math.tan(ang))
    assert o.count_generators() == int(keep.sum())
    exact = math.pi * (h * math.tan(ang)) ** 2 * h / 3.0
    assert exact <= vol.sum() < exact * 1.03


def test_torus_volume_and_filter():
    R, r = 0.28, 0.12
    g = TorusGeometry((0.5, 0.5, 0.5), (0.0, 0.0, 1.0), R, r)
    pts, o, vol = run([Wall(-1000, g)])
    rho = np.hypot(pts[:, 0] - 0.5, pts[:, 1] - 0.5)
    keep = np.hypot(rho - R, pts[:, 2] - 0.5) <= r
    assert o.count_generators() == int(keep.sum())
    exact = 2.0 * math.pi ** 2 * R * r * r
    assert exact <= vol.sum() < exact * 1.03


def polyline_dist(pts, samples, closed):
    n = len(samples)
    segs = n if closed else n - 1
    best = np.full(len(pts), np.inf)
    for i in range(segs):
        p0, p1 = samples[i], samples[(i + 1) % n]
        v = p1 - p0
        t = np.clip(((pts - p0) @ v) / (v @ v), 0.0, 1.0)
        d = np.linalg.norm(pts - (p0 + t[:, None] * v), axis=1)
        best = np.minimum(best, d)
    return best


def ring(n, radius):
    a = 2.0 * np.pi * np.arange(n) / n
    return np.stack([0.5 + radius * np.cos(a), 0.5 + radius * np.sin(a), np.full(n, 0.5)], axis=1)


class _Ring:
    """A tube geometry with explicit samples (the same wall kind as the curve geometries)."""

    def __init__(self, samples, tube_radius, closed):
        from vorothree_amd.tessellation import _Tube
        self.__class__ = type("RingTube", (_Tube,), {})
        self.samples = np.ascontiguousarray(samples)
        self.tube_radius = tube_radius
        self.closed = closed


@pytest.mark.parametrize("closed", [True, False])
def test_tube_volume_and_filter(closed):
    samples = ring(48, 0.25)
    g = _Ring(samples, 0.08, closed)
    pts, o, vol = run([Wall(-1000, g)])
    keep = polyline_dist(pts, samples, closed) <= 0.08
    assert abs(o.count_generators() - int(keep.sum())) <= 2     # boundary rounding only
    # polygonal ring of 48 segments, tube radius r: about its length times pi r^2
    length = (48 if closed else 47) * 2 * 0.25 * math.sin(math.pi / 48)
    approx = length * math.pi * 0.08 ** 2
    assert 0.9 * approx < vol.sum() < 1.15 * approx


def test_curve_constructors_hit_their_control_points():
    b = CubicBezierGeometry((0.1, 0.2, 0.3), (0.4, 0.9, 0.1), (0.8, 0.1, 0.7), (0.9, 0.8, 0.2), 0.05, 32, False)
    assert b.samples.shape == (33, 3)
    assert np.array_equal(b.samples[0], [0.1, 0.2, 0.3]) and np.array_equal(b.samples[-1], [0.9, 0.8, 0.2])
    k = TrefoilKnotGeometry((0.5, 0.5, 0.5), 0.1, 0.03, 90)
    assert k.closed and k.samples.shape == (90, 3)
    assert np.allclose(k.samples[0], [0.5, 0.4, 0.5], atol=1e-15)   # t = 0: (0, 1 - 2, 0) * scale
    cp = [[0.2, 0.2, 0.5], [0.8, 0.2, 0.5], [0.8, 0.8, 0.5], [0.2, 0.8, 0.5]]
    c = CatmullRomGeometry(cp, 0.05, 40, True)
    assert c.samples.shape == (41, 3)
    for i in range(4):                                              # interpolating spline
        assert np.array_equal(c.samples[10 * i], cp[i])
    o = CatmullRomGeometry(cp, 0.05, 30, False)
    for i in range(3):
        assert np.array_equal(o.samples[10 * i], cp[i])
    # open curve at t = 1: get_point (wall_3d.rs:853-855) moves to the last segment but keeps
    # weight 0, so the last sample repeats the second-to-last control point
    assert np.array_equal(o.samples[30], cp[2])


def sealed_pair(walls, n=12_000, seed=990, grid=18):
    pts = W.PCG32(seed).uniform(n, (0, 0, 0), (1, 1, 1))
    o = OracleTess((0, 0, 0), (1, 1, 1), grid, grid, grid)
    for w in walls:
        oracle_wall(o, w)
    o.set_generators(pts)
    o.calculate(threads=8)
    v0, _, _, _ = o.reduced()
    nbs = [o.get_cell(i).face_neighbors for i in range(len(v0))]
    o.calculate_sealed(threads=8)
    v1, _, _, _ = o.reduced()
    return o, v0, v1, nbs


def test_sealing_only_shrinks_cells_and_keeps_the_ball():
    from vorothree_amd import SphereGeometry
    R = 0.45
    o, v0, v1, nbs = sealed_pair([Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), R))])
    ball = 4.0 / 3.0 * math.pi * R ** 3
    # tangent planes of a convex wall never cut into it: the sealed union still covers the ball
    assert ball <= v1.sum() < v0.sum()
    assert np.all(v1 <= v0 + 1e-15)
    log = o.seal_log()
    assert len(log) > 0
    assert set(log[:, 0]) == set(np.nonzero(v1 < v0)[0])
    for i, j, wid in log[::7]:
        assert j in nbs[i] and wid == -1000 and -1000 in nbs[i] and -1000 in nbs[j]


def test_sealing_ignores_planar_walls():
    cube = ConvexPolyhedronGeometry.new_hexahedron((0.5, 0.5, 0.5), 0.7)
    from vorothree_amd import PlaneGeometry
    walls = [Wall(-1000, cube), Wall(-1001, PlaneGeometry((0.5, 0.5, 0.5), (1.0, 0.2, 0.0)))]
    o, v0, v1, _ = sealed_pair(walls, seed=991)
    assert len(o.seal_log()) == 0
    assert np.array_equal(v0.view(np.uint64), v1.view(np.uint64))


def test_prune_vertices_lie_on_both_faces():
    R = 0.42
    from vorothree_amd import SphereGeometry
    o, _, _, _ = sealed_pair([Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), R))], seed=992)
    o.prune_boundaries()
    ints, pos = o.prune_log()
    rows = ints.reshape(-1, 4)
    assert len(rows) > 0 and rows[:, 3].sum() == len(pos)
    gen = o.generators().reshape(-1, 3)
    k = 0
    for i, j, wid, c in rows:
        assert wid == -1000 and c >= 1
        v = pos[k:k + c]
        k += c
        # on the bisector of generators i and j, and outside or on the sphere (tangent planes)
        di = np.sum((v - gen[i]) ** 2, axis=1)
        dj = np.sum((v - gen[j]) ** 2, axis=1)
        assert np.all(np.abs(di - dj) < 1e-9)
        assert np.all(np.linalg.norm(v - 0.5, axis=1) >= R - 1e-9)


# --- geometries of the caller's (WallGeometry implemented outside the crate) -------------------

def _cells(t):
    vol, cen, nf, _ = t.reduced()
    return vol, cen, nf, [t.get_cell(i) for i in range(t.count_cells())]


def _same(a, b):
    va, ca, na, xa = a
    vb, cb, nb, xb = b
    assert np.array_equal(va.view(np.uint64), vb.view(np.uint64))
    assert np.array_equal(ca.view(np.uint64), cb.view(np.uint64))
    assert np.array_equal(na, nb)
    for p, q in zip(xa, xb):
        assert np.array_equal(p.vertices.view(np.uint64), q.vertices.view(np.uint64))
        assert np.array_equal(p.face_indices, q.face_indices)
        assert np.array_equal(p.face_neighbors, q.face_neighbors)


def test_host_sphere_equals_builtin_sphere():
    """The callback path (contains on every point, cut per generator, curved for sealing) gives the
    built-in SphereGeometry's cells and seal log bit for bit."""
    from host_geometries import PySphere
    pts = W.PCG32(1201).uniform(4000, (0, 0, 0), (1, 1, 1))
    a = OracleTess((0, 0, 0), (1, 1, 1), 14, 14, 14)
    b = OracleTess((0, 0, 0), (1, 1, 1), 14, 14, 14)
    a.add_wall_sphere(-1000, (0.5, 0.5, 0.5), 0.45)
    b.add_wall_host(-1000, PySphere((0.5, 0.5, 0.5), 0.45))
    for t in (a, b):
        t.set_generators(pts)
        t.calculate_sealed(threads=4)
    assert a.count_generators() == b.count_generators() > 0
    _same(_cells(a), _cells(b))
    assert np.array_equal(a.seal_log(), b.seal_log()) and len(a.seal_log()) > 0


def test_host_octahedron_equals_builtin_polyhedron_and_fills_it():
    from host_geometries import PyOctahedron, octahedron_planes
    pts = W.PCG32(1202).uniform(3000, (0, 0, 0), (1, 1, 1))
    planes = octahedron_planes((0.5, 0.5, 0.5), 0.45)
    a = OracleTess((0, 0, 0), (1, 1, 1), 12, 12, 12)
    b = OracleTess((0, 0, 0), (1, 1, 1), 12, 12, 12)
    a.add_wall(-1000, 5, pts=[list(q) + list(n) for q, n in planes])
    b.add_wall_host(-1000, PyOctahedron((0.5, 0.5, 0.5), 0.45))
    for t in (a, b):
        t.set_generators(pts)
        t.calculate(threads=4)
    _same(_cells(a), _cells(b))
    vol, _, _, _ = b.reduced()
    assert abs(vol.sum() - 4.0 / 3.0 * 0.45 ** 3) < 1e-12
