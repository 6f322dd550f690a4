"""GPU parity: libvoronoid_hip.so against the CPU oracle on the same seeded inputs.

Bar: bit-exact cells (vertex coordinates, face counts, face vertex indices, face neighbour ids)
and bit-exact reductions (volume, centroid, face count) -- which is stronger than the 1e-12
relative volume tolerance and the exact neighbour sets the task asks for. At full size the
checks are a bit-exact sample of cells plus size-independent properties (volume sums, neighbour
reciprocity).
"""
import numpy as np
import pytest

from conftest import GOLDEN, assert_same_cell, bits, read_generators, read_voropp_vertices
from oracle import OracleTess
from vorothree_amd import (Algorithm3DGrid, BoundingBox, PlaneGeometry, SphereGeometry,
                           Tessellation, Wall, WALL_ID_MAX, shell_prefix)
from vorothree_amd import workloads as W

pytestmark = pytest.mark.gpu


def build_pair(lo, hi, grid, points, walls=(), keep_cells=True, walls_first=False):
    b = BoundingBox(lo, hi)
    g = Tessellation(b, Algorithm3DGrid(*grid, b), keep_cells=keep_cells)
    o = OracleTess(lo, hi, *grid)

    def add_walls():
        for w in walls:
            if w[0] == "sphere":
                g.add_wall(Wall(w[1], SphereGeometry(w[2], w[3])))
                o.add_wall_sphere(w[1], w[2], w[3])
            else:
                g.add_wall(Wall(w[1], PlaneGeometry(w[2], w[3])))
                o.add_wall_plane(w[1], w[2], w[3])

    if walls_first:
        add_walls()
    g.set_generators(points)
    o.set_generators(points)
    if not walls_first:
        add_walls()
    return g, o


def check_all_cells(g, o, threads=8):
    g.calculate()
    o.calculate(threads=threads)
    n = o.count_cells()
    assert g.count_cells() == n
    assert np.array_equal(bits(g.generators()), bits(o.generators()))
    for i in range(n):
        assert_same_cell(g.get_cell(i), o.get_cell(i), i)
    vol, cen, nf, _ = o.reduced()
    assert np.array_equal(bits(g.volumes()), bits(vol))
    assert np.array_equal(bits(g.centroids()), bits(cen))
    assert np.array_equal(g.face_counts(), nf)
    return vol


def test_config1_plain_1k_bit_exact():
    w = W.config1()
    g, o = build_pair(w.lo, w.hi, w.grid, w.points)
    vol = check_all_cells(g, o)
    assert abs(vol.sum() - 1000.0) <= 1e-9 * 1000.0
    off, ids = g.neighbors()
    for i in range(g.count_cells()):
        assert np.array_equal(ids[off[i]:off[i + 1]], o.get_cell(i).face_neighbors)


def test_voropp_vertex_positions_sphere_wall():
    pts = read_generators(f"{GOLDEN}/voropp/vertex_positions/generators.txt")
    g, o = build_pair((-10.0,) * 3, (10.0,) * 3, (6, 6, 6), pts,
                      [("sphere", WALL_ID_MAX, (0.0, 0.0, 0.0), 8.0)], walls_first=True)
    check_all_cells(g, o)
    expected = read_voropp_vertices(f"{GOLDEN}/voropp/vertex_positions/output.txt")
    for cid, ev in expected.items():
        v = g.get_cell(cid).vertices().reshape(-1, 3)
        assert len(v) == len(ev), cid
        for e in ev:
            assert np.any(np.all(np.abs(v - e) < 1e-4, axis=1)), (cid, e)


def test_voropp_face_counts_box():
    pts = read_generators(f"{GOLDEN}/voropp/face_counts/generators.txt")
    g, o = build_pair((-10.0,) * 3, (10.0,) * 3, (6, 6, 6), pts)
    check_all_cells(g, o)
    exp = np.loadtxt(f"{GOLDEN}/voropp/face_counts/output.txt", dtype=np.int64)
    nf = g.face_counts()
    for i, f in exp:
        assert nf[i] == f, (i, nf[i], f)


def test_reference_integration_cases():
    # tests/integration.rs:22-46, 48-79; tests/mapping.rs:3-36
    b = BoundingBox((0, 0, 0), (100, 100, 100))
    t = Tessellation(b, Algorithm3DGrid(10, 10, 10, b))
    t.set_generators([10.0, 10.0, 10.0, 90.0, 90.0, 90.0])
    t.calculate()
    assert t.count_generators() == 2 and t.count_cells() == 2
    assert t.get_cell(0).id() == 0 and t.get_cell(1).id() == 1
    assert abs(t.get_cell(0).volume() + t.get_cell(1).volume() - 1e6) < 1e-3
    assert t.get_cell(2) is None
    pts = [[x, y, z] for x in (25.0, 75.0) for y in (25.0, 75.0) for z in (25.0, 75.0)]
    t.set_generators(pts)
    t.calculate()
    for i in range(8):
        assert abs(t.get_cell(i).volume() - 125000.0) < 1e-3
    b1 = BoundingBox((0, 0, 0), (1, 1, 1))
    t1 = Tessellation(b1, Algorithm3DGrid(4, 4, 4, b1))
    t1.set_generators([[x, y, z] for x in (0.25, 0.75) for y in (0.25, 0.75) for z in (0.25, 0.75)])
    counts = t1.map(lambda c: len(c.face_counts()))
    assert counts == [6] * 8


def test_walls_plane_and_sphere_small():
    g0 = W.PCG32(11)
    pts = g0.uniform(3000, (0, 0, 0), (2.0, 1.0, 1.5))
    walls = [("sphere", -1000, (1.0, 0.5, 0.75), 0.6), ("plane", -1001, (1.0, 0.5, 0.75), (0.3, -1.0, 0.2))]
    g, o = build_pair((0.0, 0.0, 0.0), (2.0, 1.0, 1.5), (12, 6, 9), pts, walls)
    check_all_cells(g, o)
    g2, o2 = build_pair((0.0, 0.0, 0.0), (2.0, 1.0, 1.5), (12, 6, 9), pts, walls[1:], walls_first=True)
    vol = check_all_cells(g2, o2)
    assert vol.sum() > 0


def test_degenerate_inputs():
    # duplicates (zero normal), points outside the box, a single point, an empty set
    pts = np.array([[0.5, 0.5, 0.5], [0.5, 0.5, 0.5], [0.2, 0.7, 0.1], [1.5, -0.2, 0.5],
                    [0.9, 0.9, 0.9], [0.2, 0.7, 0.1]])
    g, o = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (3, 3, 3), pts)
    check_all_cells(g, o)
    g1, o1 = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (2, 2, 2), pts[:1])
    check_all_cells(g1, o1)
    assert abs(g1.get_cell(0).volume() - 1.0) < 1e-15
    g0, o0 = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (2, 2, 2), np.zeros((0, 3)))
    g0.calculate()
    assert g0.count_cells() == 0 and g0.get_cell(0) is None


@pytest.mark.parametrize("jitter", [0.0, 1e-10, 1e-7])
@pytest.mark.parametrize("wall", [False, True])
def test_lattice_bit_exact(jitter, wall):
    # a 10^3 lattice, exact or jittered around the clip tolerance: candidate planes through existing
    # vertices (distances at or near +-1e-9), four-fold vertices and crossings that share edges
    n = 10
    ax = (np.arange(n) + 0.5) / n
    pts = np.stack(np.meshgrid(ax, ax, ax, indexing="ij"), axis=-1).reshape(-1, 3)
    if jitter:
        pts = pts + W.PCG32(11).uniform(len(pts), (-jitter,) * 3, (jitter,) * 3)
    walls = [("sphere", -1000, (0.5, 0.5, 0.5), 0.45)] if wall else []
    for grid in ((n, n, n), (7, 7, 7)):
        g, o = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), grid, pts, walls)
        vol = check_all_cells(g, o)
        if not wall:
            # vertices within the clip tolerance (1e-9) of a plane are kept on both sides, so a
            # jittered lattice overlaps by a tolerance-sized volume, as the oracle does
            assert abs(vol.sum() - 1.0) < (1e-12 if jitter == 0.0 else 1e-6)


def test_uniform_sphere_50k_bit_exact():
    w = W.config2(n=50_000, grid=37)
    g, o = build_pair(w.lo, w.hi, w.grid, w.points, w.walls)
    vol = check_all_cells(g, o)
    ball = 4.0 / 3.0 * np.pi * 0.125
    assert vol.sum() >= ball and (vol.sum() - ball) / ball < 0.01


def test_clustered_60k_bit_exact():
    g0 = W.PCG32(3)
    pts = W.clustered_points(g0, 60_000, 8, 0.03, 0.8)
    g, o = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (31, 31, 31), pts)
    vol = check_all_cells(g, o)
    assert abs(vol.sum() - 1.0) < 1e-9
    st = g.stats()
    assert st[0] == 60_000


def test_relax_matches_oracle():
    w = W.config1()
    g, o = build_pair(w.lo, w.hi, w.grid, w.points)
    for _ in range(3):
        g.calculate()
        o.calculate(threads=8)
        g.relax()
        o.relax()
        assert np.array_equal(bits(g.generators()), bits(o.generators()))
    check_all_cells(g, o)


def fibonacci_sphere(n, r, c):
    k = np.arange(n) + 0.5
    phi = np.arccos(1.0 - 2.0 * k / n)
    th = np.pi * (1.0 + 5 ** 0.5) * k
    return np.stack([c[0] + r * np.cos(th) * np.sin(phi), c[1] + r * np.sin(th) * np.sin(phi),
                     c[2] + r * np.cos(phi)], axis=1)


@pytest.mark.parametrize("nshell", [70, 400, 1500])
def test_large_cells_move_to_larger_tiers(nshell):
    # a generator surrounded by nshell points on a sphere: its cell has about nshell faces
    # (70 -> past the 64-vertex fast tier; 400 -> past the 256-vertex tier; 1500 -> about 3,000
    # vertices, past the 1,024-vertex LDS tier into the global-memory tier)
    bg = W.PCG32(51).uniform(300, (0, 0, 0), (1, 1, 1))
    pts = np.concatenate([[[0.5, 0.5, 0.5]], fibonacci_sphere(nshell, 0.2, (0.5, 0.5, 0.5)), bg])
    g, o = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (5, 5, 5), pts)
    check_all_cells(g, o)
    assert g.stats()[1] > 0, "no cell reached a larger tier"
    # the background points inside the shell cut some of its faces away
    assert g.face_counts()[0] >= (nshell // 2 if nshell < 1000 else 600)
    if nshell >= 1500:
        assert o.reduced([0])[3][0, 6] > 1024, "the cell never outgrew the LDS tiers"


def test_sparse_grid_extends_shell_prefix():
    # 40 generators in a 48^3 grid: search radii span many bins, past the first shell prefix
    pts = W.PCG32(52).uniform(40, (0, 0, 0), (1, 1, 1))
    g, o = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (48, 48, 48), pts)
    vol = check_all_cells(g, o)
    assert g.stats()[2] > 0, "the shell prefix was never extended"
    assert abs(vol.sum() - 1.0) < 1e-12


def test_arena_growth_keeps_cells():
    # an arena of one vertex per cell fills at once: it doubles until the cells fit (the waves'
    # arena chunks shrink with it), and every cell comes out as with the default arena
    w = W.config2(n=40_000, grid=34)
    b = BoundingBox(w.lo, w.hi)
    ref = Tessellation(b, Algorithm3DGrid(*w.grid, b))
    small = Tessellation(b, Algorithm3DGrid(*w.grid, b))
    small.set_arena(1)
    for t in (ref, small):
        t.add_wall(Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), 0.5)))
        t.set_generators(w.points)
        t.calculate()
    n = ref.count_cells()
    assert small.count_cells() == n
    assert np.array_equal(bits(small.volumes()), bits(ref.volumes()))
    assert small.stats()[6] == ref.stats()[6] and small.stats()[7] == ref.stats()[7]
    for i in range(0, n, 37):
        c, r = small.get_cell(i), ref.get_cell(i)
        assert np.array_equal(bits(c.vertices()), bits(r.vertices()))
        assert np.array_equal(c.face_counts(), r.face_counts())
        assert np.array_equal(c.face_indices(), r.face_indices())
        assert np.array_equal(c.face_neighbors(), r.face_neighbors())
    off_a, ids_a = small.neighbors()
    off_b, ids_b = ref.neighbors()
    assert np.array_equal(off_a, off_b) and np.array_equal(ids_a, ids_b)


def test_arena_gives_up_with_an_error_and_no_results():
    # one launch with an arena of one vertex per cell cannot hold the cells: calculate() reports
    # VD_ECAPACITY and leaves no cells to read (no half-written arena offsets); with the default
    # attempts the same handle then computes every cell as the oracle does
    from vorothree_amd._lib import VD_ECAPACITY, VD_ESTATE, VdError
    w = W.config2(n=20_000, grid=27)
    g, o = build_pair(w.lo, w.hi, w.grid, w.points, w.walls)
    g.set_arena(1)
    g.set_attempts(1)
    with pytest.raises(VdError) as ei:
        g.calculate()
    assert ei.value.code == VD_ECAPACITY
    assert g.count_cells() == 0 and g.get_cell(0) is None
    with pytest.raises(VdError) as es:
        g.volumes()
    assert es.value.code == VD_ESTATE
    g.set_attempts(8)
    check_all_cells(g, o)


def test_voids_extend_once_and_probe_on_every_later_calculate():
    # 8 clusters in empty space: cells next to the voids walk past the first shell prefix; the
    # prefix grows once to the length their radius needs and they rerun in the skip-ahead
    # instantiation; later calculates on the extended prefix use it from the first launch
    n = 30_000
    pts = W.clustered_points(W.PCG32(57), n, 8, 0.05, 1.0)
    grid = W.grid_for(n, (0, 0, 0), (1, 1, 1), 0.35)
    g, o = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), grid, pts, keep_cells=False)
    g.calculate()
    o.calculate(threads=8)
    st = g.stats()
    assert st[2] >= 1, "the shell prefix was never extended"
    assert st[1] > 0, "cells next to the voids run in the larger tier too"
    vol, cen, nf, _ = o.reduced()
    assert np.array_equal(bits(g.volumes()), bits(vol))
    assert np.array_equal(bits(g.centroids()), bits(cen))
    assert np.array_equal(g.face_counts(), nf)
    g.calculate()                                   # extended prefix: the first launch probes
    assert g.stats()[2] == 0
    assert np.array_equal(bits(g.volumes()), bits(vol))
    assert np.array_equal(bits(g.centroids()), bits(cen))
    g.relax()
    o.relax()
    g.calculate()
    o.calculate(threads=8)
    vol, cen, nf, _ = o.reduced()
    assert np.array_equal(bits(g.generators()), bits(o.generators()))
    assert np.array_equal(bits(g.volumes()), bits(vol))
    assert np.array_equal(bits(g.centroids()), bits(cen))
    off, ids = g.neighbors()
    for i in range(0, n, 97):
        assert np.array_equal(ids[off[i]:off[i + 1]], o.get_cell(i).face_neighbors), i


@pytest.mark.parametrize("n,keep", [(30_000, False), (12_000, True)])
def test_far_cell_lists_bit_exact(n, keep):
    # cells next to the voids rerun over per-cell lists of the shell entries whose bins hold
    # generators: forced for every such cell (1), the default, and never (0) all give the oracle's
    # bits, full cells included
    pts = W.clustered_points(W.PCG32(71), n, 8, 0.05, 1.0)
    grid = W.grid_for(n, (0, 0, 0), (1, 1, 1), 0.35)
    o = OracleTess((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), *grid)
    o.set_generators(pts)
    o.calculate(threads=8)
    vol, cen, nf, _ = o.reduced()
    for min_entries in (1, None, 0):
        g, _ = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), grid, pts, keep_cells=keep)
        if min_entries is not None:
            g.set_far_list(min_entries)
        g.calculate()
        assert g.stats()[2] >= 1, "no cell reached across empty space"
        assert np.array_equal(bits(g.volumes()), bits(vol)), min_entries
        assert np.array_equal(bits(g.centroids()), bits(cen)), min_entries
        assert np.array_equal(g.face_counts(), nf), min_entries
        off, ids = g.neighbors()
        for i in range(0, n, 53):
            assert np.array_equal(ids[off[i]:off[i + 1]], o.get_cell(i).face_neighbors), (min_entries, i)
        if keep:
            for i in range(0, n, 97):
                assert_same_cell(g.get_cell(i), o.get_cell(i), i)
        g.calculate()                       # the extended prefix from the first launch
        assert np.array_equal(bits(g.volumes()), bits(vol)), min_entries
        g.close()


def test_many_handles_share_side_streams():
    # 48 handles alive at once, most with cells next to voids whose reruns go to a side stream;
    # the handles share the library's few side streams per device, also when two threads
    # calculate at the same time; every handle keeps the oracle's bits
    import threading
    n = 30_000
    seeds = (57, 58, 59)
    lo, hi = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    grid = W.grid_for(n, lo, hi, 0.35)
    pts, ref = {}, {}
    for s in seeds:
        pts[s] = W.clustered_points(W.PCG32(s), n, 8, 0.05, 1.0)
        o = OracleTess(lo, hi, *grid)
        o.set_generators(pts[s])
        o.calculate(threads=8)
        ref[s] = o.reduced()[0]
    b = BoundingBox(lo, hi)
    hs = []
    for k in range(48):
        t = Tessellation(b, Algorithm3DGrid(*grid, b), keep_cells=False)
        t.set_generators(pts[seeds[k % 3]])
        hs.append(t)
    try:
        extended = 0
        for t in hs:
            t.calculate()
            extended += int(t.stats()[2] >= 1)
        assert extended >= 24, f"only {extended} of 48 handles extended their shell prefix"
        for k, t in enumerate(hs):
            assert np.array_equal(bits(t.volumes()), bits(ref[seeds[k % 3]])), k
        errs = []

        def run(part):
            try:
                for t in part:
                    t.calculate()
            except Exception as e:      # reported by the main thread
                errs.append(e)

        th = [threading.Thread(target=run, args=(hs[j::2],)) for j in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        assert not errs, errs
        for k, t in enumerate(hs):
            assert np.array_equal(bits(t.volumes()), bits(ref[seeds[k % 3]])), k
    finally:
        for t in hs:
            t.close()


@pytest.mark.parametrize("lo,hi,grid,k", [
    ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (16, 16, 16), 7),
    ((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (20, 20, 20), 64),       # the whole table
    ((0.0, 0.0, 0.0), (2.0, 0.7, 1.3), (13, 5, 9), 9),          # anisotropic bins
    ((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), (40, 40, 40), 28),
])
def test_device_shell_prefix_equals_host(lo, hi, grid, k):
    # calculate() builds its prefixes on the device (keys + stable radix sort); the host build is
    # checked against the oracle's stable order in the CPU suite
    b = BoundingBox(lo, hi)
    d_host, k_host, n_host = shell_prefix(b, *grid, k)
    d_dev, k_dev, n_dev = shell_prefix(b, *grid, k, device=True)
    assert np.array_equal(d_host, d_dev)
    assert np.array_equal(bits(k_host), bits(k_dev))
    assert bits(np.array([n_host])) == bits(np.array([n_dev]))


@pytest.mark.parametrize("edge", [4.64e-4, 7.94e-4, 1e-3, 1e-4])
def test_centroid_guard_matches_oracle(edge):
    # |sum det| = 6V against 1e-9 (cell_3d_faces.rs:195-197): boxes on both sides of the guard
    lo = (0.25, 0.5, 0.75)
    hi = tuple(x + edge for x in lo)
    pts = np.array([[lo[0] + 0.3 * edge, lo[1] + 0.6 * edge, lo[2] + 0.5 * edge]])
    g, o = build_pair(lo, hi, (1, 1, 1), pts)
    check_all_cells(g, o)
    zero = 6.0 * edge ** 3 < 1e-9
    assert np.array_equal(g.centroids()[0], [0.0, 0.0, 0.0]) == zero


def test_empty_cell_and_relax_match_oracle():
    # a generator outside the box (no bounding-box check, tessellation.rs:43-67) ends empty; relax
    # keeps its position (tessellation.rs:486-490)
    pts = np.array([[5.0, 0.5, 0.5], [0.5, 0.5, 0.5], [0.2, 0.3, 0.4], [0.8, 0.7, 0.1]])
    g, o = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (2, 2, 2), pts)
    check_all_cells(g, o)
    assert g.get_cell(0).is_empty() and g.volumes()[0] == 0.0
    g.relax()
    o.relax()
    assert np.array_equal(bits(g.generators()), bits(o.generators()))
    assert np.array_equal(g.generators().reshape(-1, 3)[0], pts[0])
    check_all_cells(g, o)


def test_face_of_256_vertices_is_an_error():
    # 300 generators on a circle around one in the xy plane: the centre cell's box faces at z = 0
    # and z = 1 get 300 vertices. The reference stores face counts as u8 (cell_3d_faces.rs:455,
    # 491), which wraps at 256 and corrupts the cell; the library reports VD_ECAPACITY instead.
    from vorothree_amd._lib import VD_ECAPACITY, VdError
    a = 2.0 * np.pi * np.arange(300) / 300
    ring = np.stack([0.5 + 0.2 * np.cos(a), 0.5 + 0.2 * np.sin(a), np.full(300, 0.5)], axis=1)
    pts = np.concatenate([[[0.5, 0.5, 0.5]], ring])
    b = BoundingBox((0, 0, 0), (1, 1, 1))
    t = Tessellation(b, Algorithm3DGrid(4, 4, 4, b))
    t.set_generators(pts)
    with pytest.raises(VdError) as e:
        t.calculate()
    assert e.value.code == VD_ECAPACITY and "256" in str(e.value)
    assert t.count_cells() == 0


def test_bin_boundaries_and_box_faces():
    # coordinates exactly on bin boundaries and on the box faces (x = max maps to the last bin
    # through grid_limit = n - 1e-5, algo_3d_grid.rs:85-87,98-107), with random points around them
    n = 6
    k = np.arange(n + 1) / n
    edge = np.stack(np.meshgrid(k, k[::2], k[1::2], indexing="ij"), axis=-1).reshape(-1, 3)
    pts = np.concatenate([edge, W.PCG32(1501).uniform(400, (0, 0, 0), (1, 1, 1))])
    g, o = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (n, n, n), pts)
    vol = check_all_cells(g, o)
    assert abs(vol.sum() - 1.0) < 1e-12


def test_box_far_from_the_origin():
    # the volume and centroid fans are origin-relative (cell_3d_faces.rs:106-205): a unit box at
    # 1e6 cancels most of each determinant's digits, the same way on both sides
    lo = (1e6, -2e6, 5e5)
    hi = (1e6 + 1.0, -2e6 + 1.0, 5e5 + 1.0)
    pts = W.PCG32(1502).uniform(2000, lo, hi)
    g, o = build_pair(lo, hi, (12, 12, 12), pts, [("sphere", -1000, (1e6 + 0.5, -2e6 + 0.5, 5e5 + 0.5), 0.5)])
    check_all_cells(g, o)


@pytest.mark.parametrize("shape", ["plane", "line"])
def test_coplanar_and_collinear_generators(shape):
    # every cell spans the box across the degenerate directions: two large faces per cell on the
    # box sides, and ties in every bisector distance across the plane or along the line
    rng = W.PCG32(1503 if shape == "plane" else 1504)
    if shape == "plane":
        xy = rng.uniform(500, (0, 0, 0), (1, 1, 1))
        pts = np.stack([xy[:, 0], xy[:, 1], np.full(500, 0.5)], axis=1)
    else:
        t = rng.uniform(200, (0, 0, 0), (1, 1, 1))[:, 0]
        pts = np.stack([t, 0.3 + 0.4 * t, np.full(200, 0.25)], axis=1)
    g, o = build_pair((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (8, 8, 8), pts)
    vol = check_all_cells(g, o)
    assert abs(vol.sum() - 1.0) < 1e-12
