"""GPU parity at the full sizes of the 1M-uniform-sphere and 4M-clustered workloads, and at 300k
points of the clustered distribution.

Every cell's reductions (volume, centroid, face count) and neighbour ids (face order) are compared
bit for bit with the oracle run over all cells; full cells (vertices, face arrays) are compared on
a 20,000-cell sample. The clustered workload is 80% of the points from 64 Gaussian clusters
(sigma 0.02) and 20% uniform, seed 3, on the grid the density rule picks for it
(vd_grid_for_points at 0.35 generators per bin: 192^3 at 300k, 474^3 at 4M). The oracle grows
its shell order lazily, as the library does, so grids of this size fit in memory.
"""
import os

import numpy as np
import pytest

from conftest import assert_same_cell, bits
from oracle import OracleTess
from vorothree_amd import (Algorithm3DGrid, BoundingBox, SphereGeometry, Tessellation, Wall,
                           grid_for_points)
from vorothree_amd import workloads as W

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, os.cpu_count() or 1))


def run_both(w, grid, walls=()):
    b = BoundingBox(w.lo, w.hi)
    g = Tessellation(b, Algorithm3DGrid(*grid, b), keep_cells=True)
    o = OracleTess(w.lo, w.hi, *grid)
    for wl in walls:
        g.add_wall(Wall(wl[1], SphereGeometry(wl[2], wl[3])))
        o.add_wall_sphere(wl[1], wl[2], wl[3])
    g.set_generators(w.points)
    o.set_generators(w.points)
    g.calculate()
    o.calculate(threads=THREADS)
    return g, o


def compare_all(g, o, sample=None):
    """Bit-exact reductions and neighbour ids of every cell; full cells of a sample (or all)."""
    n = o.count_cells()
    assert g.count_cells() == n
    assert np.array_equal(bits(g.generators()), bits(o.generators()))
    vol, cen, nf, _ = o.reduced()
    gv = g.volumes()
    assert np.array_equal(bits(gv), bits(vol)), f"{np.count_nonzero(bits(gv) != bits(vol))} volumes differ"
    assert np.array_equal(bits(g.centroids()), bits(cen))
    assert np.array_equal(g.face_counts(), nf)
    goff, gids = g.neighbors()
    ooff, oids = o.neighbors()
    assert np.array_equal(goff.astype(np.uint64), ooff) and np.array_equal(gids, oids)
    idx = range(n) if sample is None else np.sort(np.random.default_rng(7).choice(n, sample, replace=False))
    for i in idx:
        assert_same_cell(g.get_cell(int(i)), o.get_cell(int(i)), int(i))
    return gv, vol


def test_full_config2_all_cells():
    # the timed path of the headline workload: full cells kept, 524,199 cells
    w = W.config2()
    g, o = run_both(w, w.grid, w.walls)
    gv, vol = compare_all(g, o, sample=20_000)
    ball = 4.0 / 3.0 * np.pi * 0.125
    assert gv.sum() == vol.sum()
    assert gv.sum() >= ball and (gv.sum() - ball) / ball < 0.01


def test_config3_distribution_300k_bit_exact():
    w = W.config3(n=300_000)
    grid = grid_for_points(BoundingBox(w.lo, w.hi), w.points, 0.35)
    assert tuple(grid) == (192, 192, 192)
    g, o = run_both(w, grid)
    gv, vol = compare_all(g, o)
    assert gv.sum() == vol.sum()
    assert abs(gv.sum() - 1.0) < 1e-8


def test_full_config3_all_cells():
    w = W.config3()
    grid = grid_for_points(BoundingBox(w.lo, w.hi), w.points, 0.35)
    assert tuple(grid) == (474, 474, 474)
    g, o = run_both(w, grid)
    gv, vol = compare_all(g, o, sample=20_000)
    # the fan triangulation is origin-relative (cell_3d_faces.rs:106-143): cluster cells of
    # 1e-9 volume carry 1e-16 absolute rounding each, so the sum is 1 only to about 1e-7
    assert gv.sum() == vol.sum()
    assert abs(gv.sum() - 1.0) < 1e-7


def test_full_config4_batch_members_bit_exact():
    # the batched workload at its full size in one batch (4,096 members, 64 to 65,536 generators,
    # grids of about 2 generators per bin, sphere and plane walls, 24.3M cells): every fifth
    # member (820 members of every size and wall kind, about 4.7M cells) bit for bit against the
    # oracle run on that member alone: volumes, centroids, face counts, generators and neighbour
    # ids (a fifth keeps the oracle's share of the test near half a minute)
    from test_gpu_batch import oracle_of, walls_of
    from vorothree_amd import Batch

    ws = W.config4()
    b = Batch()
    for w in ws:
        b.add(BoundingBox(w.lo, w.hi), w.grid, walls_of(w), w.points)
    b.calculate()
    assert sum(b.info(t)["cells"] for t in range(len(ws))) == 24_349_075
    bad = []
    for t in range(0, len(ws), 5):
        o = oracle_of(ws[t], ws[t].grid)
        o.calculate(threads=THREADS)
        vol, cen, nf, _ = o.reduced()
        goff, gids = b.neighbors(t)
        ooff, oids = o.neighbors()
        if not (np.array_equal(bits(b.volumes(t)), bits(vol)) and np.array_equal(bits(b.centroids(t)), bits(cen))
                and np.array_equal(b.face_counts(t), nf)
                and np.array_equal(bits(b.generators(t)), bits(o.generators()))
                and np.array_equal(goff.astype(np.uint64), ooff) and np.array_equal(gids, oids)):
            bad.append(t)
    assert not bad, f"members differing from the oracle: {bad[:20]}"


def test_scale_100m_uniform_one_tessellation():
    # one tessellation of 100M uniform generators with full cells kept (about 150 GB of device
    # memory: the vertex arena alone holds 4.0e9 entries, 96 GB): entry offsets past 2^31 and
    # byte offsets past 2^32 in every arena. The generators come from random_generators on the
    # device; the oracle computes a sample of cells spread over the index range (the last ones
    # included) from the same points.
    from vorothree_amd import grid_for
    n = 100_000_000
    b = BoundingBox((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    grid = grid_for(b, n, 1.0)
    g = Tessellation(b, Algorithm3DGrid(*grid, b), keep_cells=True)
    try:
        scale_checks(g, n, grid)
    finally:
        g.close()   # 150 GB of device memory: released even when a check fails


def scale_checks(g, n, grid):
    g.random_generators(n, seed=2026)
    assert g.count_generators() == n
    g.calculate()
    assert g.count_cells() == n
    gv = g.volumes()
    # the fan triangulation is origin-relative (cell_3d_faces.rs:106-143): a cell of volume 1e-8
    # near (1, 1, 1) sums determinants of order 1 and keeps about 1e-15 of rounding, so over 1e8
    # cells the sum is 1 to about 1e-6 (it is 1 + 4.3e-7 here; the sample below is bit-exact)
    assert abs(gv.sum() - 1.0) < 2e-6 and gv.min() > 0.0
    assert g.face_counts().min() >= 4
    pts = g.generators()
    o = OracleTess((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), *grid)
    o.set_generators(pts)
    del pts
    rng = np.random.default_rng(11)
    idx = np.unique(np.concatenate([rng.choice(n, 3000, replace=False), np.arange(n - 50, n),
                                    np.arange(0, 50), np.arange(n // 2 - 25, n // 2 + 25)]))
    o.calculate(subset=idx, threads=THREADS)
    vol, cen, _, _ = o.reduced(idx)
    assert np.array_equal(bits(gv[idx]), bits(vol))
    assert np.array_equal(bits(g.centroids()[idx]), bits(cen))
    for i in idx:
        assert_same_cell(g.get_cell(int(i)), o.get_cell(int(i)), int(i))
    # reciprocity across the sample. The reference's clip keeps vertices within 1e-9 of a plane
    # (cell_3d_faces.rs:303-326), so near-cospherical generators can leave a sliver face on one
    # side that the other side's clip does not make: such one-sided pairs must be rare, their face
    # must be a sliver, and the neighbour's cell must still be the oracle's bit for bit
    pairs, one_sided = 0, []
    for i in idx[::5]:
        ci = g.get_cell(int(i))
        for k, j in enumerate(ci.face_neighbors()):
            if j >= 0:
                pairs += 1
                if int(i) not in set(g.get_cell(int(j)).face_neighbors()):
                    one_sided.append((int(i), int(j), ci.face_area(k), ci.volume()))
    assert len(one_sided) <= max(1, pairs // 1000), one_sided[:10]
    if one_sided:
        js = np.array(sorted({j for _, j, _, _ in one_sided}), dtype=np.int64)
        o.calculate(subset=js, threads=THREADS)
        for j in js:
            assert_same_cell(g.get_cell(int(j)), o.get_cell(int(j)), int(j))
        for i, j, area, v in one_sided:
            assert area < 1e-6 * v ** (2.0 / 3.0), (i, j, area, v)
