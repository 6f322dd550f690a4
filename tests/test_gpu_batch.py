"""Batched tessellations (vd_batch_*): every member of one persistent launch against the oracle run
on that member alone, bit-exact, and against the same member computed alone on the GPU."""
import numpy as np
import pytest

from conftest import assert_same_cell, bits
from oracle import OracleTess
from vorothree_amd import (Algorithm3DGrid, Batch, BoundingBox, PlaneGeometry, SphereGeometry,
                           Tessellation, Wall, grid_for)
from vorothree_amd import workloads as W

pytestmark = pytest.mark.gpu


def small_batch(count=24, nmax=4000, seed=4):
    ws = W.config4(count=count, seed=seed, nmax=nmax)
    return ws


def walls_of(w):
    out = []
    for wl in w.walls:
        if wl[0] == "sphere":
            out.append(Wall(wl[1], SphereGeometry(wl[2], wl[3])))
        else:
            out.append(Wall(wl[1], PlaneGeometry(wl[2], wl[3])))
    return out


def oracle_of(w, grid):
    o = OracleTess(w.lo, w.hi, *grid)
    for wl in w.walls:
        if wl[0] == "sphere":
            o.add_wall_sphere(wl[1], wl[2], wl[3])
        else:
            o.add_wall_plane(wl[1], wl[2], wl[3])
    o.set_generators(w.points)
    return o


def test_batch_members_match_oracle():
    ws = small_batch()
    b = Batch(keep_cells=True)
    for w in ws:
        b.add(BoundingBox(w.lo, w.hi), w.grid, walls_of(w), w.points)
    assert len(b) == len(ws)
    b.calculate()
    for t, w in enumerate(ws):
        info = b.info(t)
        assert info["grid"] == tuple(w.grid)
        o = oracle_of(w, w.grid)
        o.calculate(threads=8)
        assert info["cells"] == o.count_cells()
        vol, cen, nf, _ = o.reduced()
        assert np.array_equal(bits(b.volumes(t)), bits(vol)), t
        assert np.array_equal(bits(b.centroids(t)), bits(cen)), t
        assert np.array_equal(b.face_counts(t), nf), t
        assert np.array_equal(bits(b.generators(t)), bits(o.generators())), t
        off, ids = b.neighbors(t)
        for i in range(0, info["cells"], 7):
            assert np.array_equal(ids[off[i]:off[i + 1]], o.get_cell(i).face_neighbors), (t, i)
        for i in range(0, info["cells"], 97):
            assert_same_cell(b.get_cell(t, i), o.get_cell(i), (t, i))


def test_batch_member_equals_single_tessellation():
    ws = small_batch(count=6, seed=14)
    b = Batch()
    for w in ws:
        b.add(BoundingBox(w.lo, w.hi), w.grid, walls_of(w), w.points)
    b.calculate()
    for t, w in enumerate(ws):
        bb = BoundingBox(w.lo, w.hi)
        s = Tessellation(bb, Algorithm3DGrid(*w.grid, bb), keep_cells=False)
        for wl in walls_of(w):
            s.add_wall(wl)
        s.set_generators(w.points)
        s.calculate()
        assert np.array_equal(bits(s.volumes()), bits(b.volumes(t)))
        assert np.array_equal(bits(s.centroids()), bits(b.centroids(t)))
        o1, i1 = s.neighbors()
        o2, i2 = b.neighbors(t)
        assert np.array_equal(o1, o2) and np.array_equal(i1, i2)


def test_batch_auto_grid_and_relax():
    ws = small_batch(count=8, seed=24, nmax=3000)
    b = Batch(per_bin=2.0)
    for w in ws:
        b.add(BoundingBox(w.lo, w.hi), None, walls_of(w), w.points)
    b.calculate()
    for t, w in enumerate(ws):
        assert b.info(t)["grid"] == grid_for(BoundingBox(w.lo, w.hi), len(w.points), 2.0)
    oracles = []
    for t, w in enumerate(ws):
        o = oracle_of(w, b.info(t)["grid"])
        o.calculate(threads=8)
        vol, _, _, _ = o.reduced()
        assert np.array_equal(bits(b.volumes(t)), bits(vol)), t
        oracles.append(o)
    # two Lloyd steps in the batch and in the oracle (tessellation.rs:482-495)
    for _ in range(2):
        b.relax()
        b.calculate()
        for t, o in enumerate(oracles):
            o.relax()
            o.calculate(threads=8)
            assert np.array_equal(bits(b.generators(t)), bits(o.generators())), t
            vol, _, _, _ = o.reduced()
            assert np.array_equal(bits(b.volumes(t)), bits(vol)), t


def test_autotune_picks_a_candidate():
    ws = small_batch(count=16, seed=34, nmax=8000)
    b = Batch()
    for w in ws:
        b.add(BoundingBox(w.lo, w.hi), None, walls_of(w), w.points)
    best = b.autotune([1.0, 2.0, 4.0])
    assert best in (1.0, 2.0, 4.0)
    b.calculate()
    t = 3
    w = ws[t]
    assert b.info(t)["grid"] == grid_for(BoundingBox(w.lo, w.hi), len(w.points), best)
    o = oracle_of(w, b.info(t)["grid"])
    o.calculate(threads=8)
    vol, _, _, _ = o.reduced()
    assert np.array_equal(bits(b.volumes(t)), bits(vol))


def test_sparse_member_extends_only_its_prefix():
    # one member needs a long shell walk (40 generators in a 48^3 grid), the others do not: the
    # longer prefix is built for that geometry only and every member still equals the oracle
    sparse = W.Workload("sparse", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (48, 48, 48),
                        W.PCG32(61).uniform(40, (0, 0, 0), (1, 1, 1)))
    dense = [W.Workload(f"dense-{k}", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (12, 12, 12),
                        W.PCG32(62 + k).uniform(3000, (0, 0, 0), (1, 1, 1))) for k in range(2)]
    ws = [dense[0], sparse, dense[1]]
    b = Batch()
    for w in ws:
        b.add(BoundingBox(w.lo, w.hi), w.grid, walls_of(w), w.points)
    b.calculate()
    assert b.stats()[2] > 0, "the shell prefix was never extended"
    for t, w in enumerate(ws):
        o = oracle_of(w, w.grid)
        o.calculate(threads=8)
        vol, cen, nf, _ = o.reduced()
        assert np.array_equal(bits(b.volumes(t)), bits(vol)), t
        assert np.array_equal(bits(b.centroids(t)), bits(cen)), t
        off, ids = b.neighbors(t)
        for i in range(0, len(vol), 11):
            assert np.array_equal(ids[off[i]:off[i + 1]], o.get_cell(i).face_neighbors), (t, i)


@pytest.mark.parametrize("far_list", [1, 0])
def test_far_lists_in_batch_members(far_list):
    # members with cells next to empty space, behind a uniform member (their bins, generators and
    # outputs start at non-zero offsets): their far cells rerun over per-cell lists (forced) or
    # over their prefixes (off), and every member equals the oracle bit for bit
    uni = W.Workload("uniform", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (14, 14, 14),
                     W.PCG32(81).uniform(2500, (0, 0, 0), (1, 1, 1)))
    voids = [W.Workload(f"void-{k}", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (30, 30, 30),
                        W.clustered_points(W.PCG32(82 + k), 9000, 8, 0.04, 1.0)) for k in range(2)]
    sparse = W.Workload("sparse", (0.0, 0.0, 0.0), (2.0, 1.0, 1.5), (40, 20, 30),
                        W.PCG32(85).uniform(60, (0, 0, 0), (2.0, 1.0, 1.5)))
    ws = [uni, voids[0], sparse, voids[1]]
    b = Batch(keep_cells=True, far_list=far_list)
    for w in ws:
        b.add(BoundingBox(w.lo, w.hi), w.grid, walls_of(w), w.points)
    b.calculate()
    assert b.stats()[2] > 0, "no cell reached across empty space"
    for t, w in enumerate(ws):
        o = oracle_of(w, w.grid)
        o.calculate(threads=8)
        vol, cen, nf, _ = o.reduced()
        assert np.array_equal(bits(b.volumes(t)), bits(vol)), t
        assert np.array_equal(bits(b.centroids(t)), bits(cen)), t
        assert np.array_equal(b.face_counts(t), nf), t
        off, ids = b.neighbors(t)
        for i in range(0, len(vol), 13):
            assert np.array_equal(ids[off[i]:off[i + 1]], o.get_cell(i).face_neighbors), (t, i)
        for i in range(0, len(vol), 211):
            assert_same_cell(b.get_cell(t, i), o.get_cell(i), (t, i))
    b.close()
