"""Seeded random small cases against the oracle, bit for bit: box shapes and offsets, grids from
1 to 13 bins per axis (anisotropic), 1 to 600 generators drawn uniform, clustered, on a jittered
lattice or with duplicates, and up to three walls of random kinds in random order. Each case runs
on both sides through set_generators, calculate and one relax step."""
import numpy as np
import pytest

from conftest import assert_same_cell, bits, oracle_wall
from oracle import OracleTess
from vorothree_amd import (Algorithm3DGrid, BoundingBox, ConeGeometry, CylinderGeometry, PlaneGeometry,
                           SphereGeometry, Tessellation, TorusGeometry, Wall)
from vorothree_amd import workloads as W

pytestmark = pytest.mark.gpu


def random_case(seed):
    rng = np.random.default_rng(seed)
    lo = rng.uniform(-5.0, 5.0, 3)
    hi = lo + rng.uniform(0.2, 4.0, 3)
    grid = tuple(int(x) for x in rng.integers(1, 14, 3))
    n = int(rng.integers(1, 601))
    kind = rng.integers(0, 4)
    pcg = W.PCG32(1000 + seed)
    if kind == 0:
        pts = pcg.uniform(n, lo, hi)
    elif kind == 1:
        c = pcg.uniform(4, lo, hi)
        pts = c[np.arange(n) % 4] + 0.05 * (hi - lo) * (pcg.uniform(n, (-1,) * 3, (1,) * 3))
    elif kind == 2:
        k = max(1, int(round(n ** (1 / 3))))
        ax = [(np.arange(k) + 0.5) / k * (hi[a] - lo[a]) + lo[a] for a in range(3)]
        pts = np.stack(np.meshgrid(*ax, indexing="ij"), axis=-1).reshape(-1, 3)
        pts = pts + pcg.uniform(len(pts), (-1e-9,) * 3, (1e-9,) * 3) * rng.integers(0, 2)
    else:
        base = pcg.uniform(max(1, n // 2), lo, hi)
        pts = np.concatenate([base, base[: n - len(base)]])   # duplicates
    mid = (lo + hi) / 2
    ext = float(np.min(hi - lo))
    walls = []
    for w in range(int(rng.integers(0, 4))):
        wk = int(rng.integers(0, 5))
        wid = -1000 - w
        axis = tuple(rng.normal(size=3))
        if wk == 0:
            g = PlaneGeometry(tuple(mid + 0.2 * ext * rng.normal(size=3)), axis)
        elif wk == 1:
            g = SphereGeometry(tuple(mid), 0.3 * ext + 0.3 * ext * rng.random())
        elif wk == 2:
            g = CylinderGeometry(tuple(mid), axis, 0.25 * ext + 0.2 * ext * rng.random())
        elif wk == 3:
            g = ConeGeometry(tuple(mid - 0.4 * ext * np.array(axis) / np.linalg.norm(axis)), axis,
                             0.3 + 0.5 * rng.random())
        else:
            g = TorusGeometry(tuple(mid), axis, 0.3 * ext, 0.1 * ext + 0.1 * ext * rng.random())
        walls.append(Wall(wid, g))
    walls_first = bool(rng.integers(0, 2))
    return tuple(lo), tuple(hi), grid, pts, walls, walls_first


def compare(g, o):
    g.calculate()
    o.calculate(threads=8)
    n = o.count_cells()
    assert g.count_cells() == n
    assert np.array_equal(bits(g.generators()), bits(o.generators()))
    for i in range(n):
        assert_same_cell(g.get_cell(i), o.get_cell(i), i)
    vol, cen, nf, _ = o.reduced()
    assert np.array_equal(bits(g.volumes()), bits(vol))
    assert np.array_equal(bits(g.centroids()), bits(cen))
    assert np.array_equal(g.face_counts(), nf)


@pytest.mark.parametrize("seed", range(40))
def test_random_case_bit_exact(seed):
    lo, hi, grid, pts, walls, walls_first = random_case(seed)
    b = BoundingBox(lo, hi)
    g = Tessellation(b, Algorithm3DGrid(*grid, b))
    o = OracleTess(lo, hi, *grid)
    if walls_first:
        for w in walls:
            g.add_wall(w)
            oracle_wall(o, w)
    g.set_generators(pts)
    o.set_generators(pts)
    if not walls_first:
        for w in walls:
            g.add_wall(w)
            oracle_wall(o, w)
    compare(g, o)
    if o.count_generators():
        g.relax()
        o.relax()
        compare(g, o)


@pytest.mark.parametrize("seed", range(10))
def test_random_batch_bit_exact(seed):
    # 1 to 8 random cases as members of one batch (walls given with the points), each member
    # against the oracle run on it alone, then one batch relax step
    from vorothree_amd import Batch
    rng = np.random.default_rng(5000 + seed)
    cases = [random_case(int(s)) for s in rng.integers(0, 10**6, int(rng.integers(1, 9)))]
    b = Batch(keep_cells=True)
    oracles = []
    for lo, hi, grid, pts, walls, _ in cases:
        b.add(BoundingBox(lo, hi), grid, walls, pts)
        o = OracleTess(lo, hi, *grid)
        for w in walls:
            oracle_wall(o, w)
        o.set_generators(pts)
        oracles.append(o)
    for step in range(2):
        b.calculate()
        for t, o in enumerate(oracles):
            o.calculate(threads=8)
            vol, cen, nf, _ = o.reduced()
            assert np.array_equal(bits(b.generators(t)), bits(o.generators())), (step, t)
            assert np.array_equal(bits(b.volumes(t)), bits(vol)), (step, t)
            assert np.array_equal(bits(b.centroids(t)), bits(cen)), (step, t)
            assert np.array_equal(b.face_counts(t), nf), (step, t)
            for i in range(o.count_cells()):
                assert_same_cell(b.get_cell(t, i), o.get_cell(i), (step, t, i))
        b.relax()
        for o in oracles:
            o.relax()
