"""Graph replay (VD_OPT_GRAPHS / VD_BOPT_GRAPHS): a calculate() whose launch sequence repeats the
previous one is captured once into a hipGraph and replayed. Every replay must give the eager
result bit for bit, with eager work and synchronisations of the same process between replays."""
import numpy as np
import pytest

from conftest import assert_same_cell, bits
from oracle import OracleTess
from vorothree_amd import (Algorithm3DGrid, Batch, BoundingBox, SphereGeometry, Tessellation, Wall,
                           _hip)
from vorothree_amd import workloads as W

pytestmark = pytest.mark.gpu

UNIT = BoundingBox((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))


def tess(grid, graphs, keep=True, walls=()):
    t = Tessellation(UNIT, Algorithm3DGrid(*grid, UNIT), keep_cells=keep, graphs=graphs)
    for w in walls:
        t.add_wall(w)
    return t


def same(a, b):
    assert np.array_equal(bits(a.volumes()), bits(b.volumes()))
    assert np.array_equal(bits(a.centroids()), bits(b.centroids()))
    assert np.array_equal(a.face_counts(), b.face_counts())
    oa, ia = a.neighbors()
    ob, ib = b.neighbors()
    assert np.array_equal(oa, ob) and np.array_equal(ia, ib)


def test_replays_match_eager_runs():
    n = 20_000
    g = tess((22, 22, 22), graphs=True)
    e = tess((22, 22, 22), graphs=False)
    other = tess((9, 9, 9), graphs=False, keep=False)     # unrelated eager work between replays
    for seed in range(5):
        pts = W.PCG32(700 + seed).uniform(n, (0, 0, 0), (1, 1, 1))
        g.set_generators(pts)
        g.calculate()
        e.set_generators(pts)
        e.calculate()
        same(g, e)
        other.set_generators(W.PCG32(800 + seed).uniform(3000, (0, 0, 0), (1, 1, 1)))
        other.calculate()
        _hip.synchronize()
        if seed == 2:
            o = OracleTess((0, 0, 0), (1, 1, 1), 22, 22, 22)
            o.set_generators(pts)
            o.calculate(threads=8)
            vol, _, _, _ = o.reduced()
            assert np.array_equal(bits(g.volumes()), bits(vol))
            for i in range(0, n, 997):
                assert_same_cell(g.get_cell(i), o.get_cell(i), i)
    st = g.stats()
    assert st[9] == 1, "one capture for one repeated shape"
    assert st[8] == 4, "calls 2-5 replay the graph"
    assert e.stats()[8] == 0


@pytest.mark.parametrize("kind", ["plane", "sphere"])
def test_lloyd_loop_with_walls_replays(kind):
    from vorothree_amd import PlaneGeometry
    pts = W.PCG32(711).uniform(8000, (0, 0, 0), (1, 1, 1))
    geo = PlaneGeometry((0.3, 0.0, 0.0), (1.0, 0.0, 0.0)) if kind == "plane" else \
        SphereGeometry((0.5, 0.5, 0.5), 0.45)
    walls = [Wall(-1000, geo)]
    g = tess((15, 15, 15), graphs=True, keep=False, walls=walls)
    e = tess((15, 15, 15), graphs=False, keep=False, walls=walls)
    for t in (g, e):
        t.set_generators(pts)
        t.calculate()
    for _ in range(6):
        for t in (g, e):
            t.relax()
            t.calculate()
        same(g, e)
        assert np.array_equal(bits(g.generators()), bits(e.generators()))
    if kind == "plane":
        # relaxed generators stay in the half-space, so every iteration repeats the shape
        assert g.stats()[8] >= 5


def test_sparse_grid_extension_under_graphs():
    # the first round of a replay can report shell-prefix failures; the extension changes the
    # descriptors, so the next call captures again, and the results stay exact
    g = tess((48, 48, 48), graphs=True)
    e = tess((48, 48, 48), graphs=False)
    for seed in range(3):
        pts = W.PCG32(720 + seed).uniform(40, (0, 0, 0), (1, 1, 1))
        for t in (g, e):
            t.set_generators(pts)
            t.calculate()
        same(g, e)
        assert abs(g.volumes().sum() - 1.0) < 1e-12


def test_batch_replays_match_eager():
    ws = W.config4(count=12, seed=44, nmax=5000)

    def walls_of(w):
        from vorothree_amd import PlaneGeometry
        out = []
        for wl in w.walls:
            geo = SphereGeometry(wl[2], wl[3]) if wl[0] == "sphere" else PlaneGeometry(wl[2], wl[3])
            out.append(Wall(wl[1], geo))
        return out

    bg = Batch(graphs=True)
    be = Batch(graphs=False)
    for b in (bg, be):
        for w in ws:
            b.add(BoundingBox(w.lo, w.hi), w.grid, walls_of(w), w.points)
        b.calculate()
    for _ in range(4):
        for b in (bg, be):
            b.relax()
            b.calculate()
        for t in range(len(ws)):
            assert np.array_equal(bits(bg.volumes(t)), bits(be.volumes(t))), t
            assert np.array_equal(bits(bg.centroids(t)), bits(be.centroids(t))), t
            assert np.array_equal(bits(bg.generators(t)), bits(be.generators(t))), t
        _hip.synchronize()
    # repeated calculate() on unchanged generators replays as well
    bg.calculate()
    bg.calculate()
    be.calculate()
    for t in range(len(ws)):
        assert np.array_equal(bits(bg.volumes(t)), bits(be.volumes(t))), t
    assert bg.stats()[8] >= 1
