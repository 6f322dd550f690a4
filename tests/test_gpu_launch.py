"""Launch shapes of the cell kernels: every tier-0 instantiation (5, 6, 7, 8 waves per SIMD), queue
grain and first tier gives bit-exact the oracle's cells, and the launch autotuners keep results
identical while picking a valid shape.
"""
import numpy as np
import pytest

from conftest import assert_same_cell, bits
from oracle import OracleTess
from vorothree_amd import Algorithm3DGrid, Batch, BoundingBox, SphereGeometry, Tessellation, Wall
from vorothree_amd._lib import VdError

pytestmark = pytest.mark.gpu


def _sphere_case(n=20_000, seed=11):
    rng = np.random.default_rng(seed)
    pts = rng.random((n, 3))
    lo, hi, grid = (0.0,) * 3, (1.0,) * 3, (27, 27, 27)
    return lo, hi, grid, pts


@pytest.fixture(scope="module")
def oracle_sphere():
    lo, hi, grid, pts = _sphere_case()
    o = OracleTess(lo, hi, *grid)
    o.add_wall_sphere(-1000, (0.5, 0.5, 0.5), 0.5)
    o.set_generators(pts)
    o.calculate(threads=8)
    return o


@pytest.mark.parametrize("waves,grain,first,probe", [(5, 1, 0, 0), (6, 1, 0, 0), (7, 1, 0, 0), (8, 1, 0, 0),
                                                     (6, 4, 0, 0), (6, 8, 0, 0), (5, 64, 0, 0), (8, 4, 0, 0),
                                                     (6, 1, 1, 0), (7, 2, 1, 0), (5, 1, 0, 1), (6, 2, 0, 1),
                                                     (7, 1, 0, 1), (8, 1, 0, 1), (6, 1, 1, 1)])
def test_launch_shape_bit_exact(oracle_sphere, waves, grain, first, probe):
    print(f"launch shape waves={waves} grain={grain} first_tier={first} probe={probe}", flush=True)
    lo, hi, grid, pts = _sphere_case()
    b = BoundingBox(lo, hi)
    g = Tessellation(b, Algorithm3DGrid(*grid, b))
    g.add_wall(Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), 0.5)))
    g.set_launch(t0_waves=waves, queue_grain=grain, first_tier=first, first_probe=probe)
    g.set_generators(pts)
    o = oracle_sphere
    for rep in range(2):   # the second calculate replays the first round from a graph
        g.calculate()
        n = o.count_cells()
        assert g.count_cells() == n
        vol, cen, nf, _ = o.reduced()
        assert np.array_equal(bits(g.volumes()), bits(vol))
        assert np.array_equal(bits(g.centroids()), bits(cen))
        assert np.array_equal(g.face_counts(), nf)
    st = g.stats()
    assert st[0] == n
    if first == 1:
        assert st[1] == 0   # every cell of this workload fits 128 vertices
    for i in range(0, n, 997):
        assert_same_cell(g.get_cell(i), o.get_cell(i), i)
    g.close()


def test_launch_options_rejected():
    b = BoundingBox((0.0,) * 3, (1.0,) * 3)
    g = Tessellation(b, Algorithm3DGrid(4, 4, 4, b))
    for kw in ({"t0_waves": 4}, {"t0_waves": 9}, {"queue_grain": 0}, {"queue_grain": 65}, {"first_tier": 2},
               {"first_tier": -1}, {"first_probe": 2}):
        with pytest.raises(VdError):
            g.set_launch(**kw)
    g.close()


def test_tess_autotune_launch_keeps_results(oracle_sphere):
    lo, hi, grid, pts = _sphere_case()
    b = BoundingBox(lo, hi)
    g = Tessellation(b, Algorithm3DGrid(*grid, b), keep_cells=False)
    g.add_wall(Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), 0.5)))
    g.set_generators(pts)
    picks = g.autotune_launch()
    print("picks", picks, flush=True)
    assert picks["t0_waves"] in (5, 6, 7, 8)
    assert picks["queue_grain"] in (1, 2, 4, 8)
    assert picks["first_tier"] in (0, 1) and picks["first_probe"] in (0, 1)
    vol, cen, nf, _ = oracle_sphere.reduced()
    assert np.array_equal(bits(g.volumes()), bits(vol))
    g.calculate()
    assert np.array_equal(bits(g.volumes()), bits(vol))
    assert np.array_equal(bits(g.centroids()), bits(cen))
    g.close()


def test_batch_autotune_launch_keeps_results():
    rng = np.random.default_rng(5)
    bt = Batch(per_bin=2.0)
    members = []
    for k in range(24):
        n = int(rng.integers(64, 3000))
        pts = rng.random((n, 3))
        bt.add(BoundingBox((0.0,) * 3, (1.0,) * 3), None, [], pts)
        members.append(pts)
    bt.calculate()
    ref = [bits(bt.volumes(t)).copy() for t in range(len(members))]
    picks = bt.autotune_launch()
    print("batch picks", picks, flush=True)
    assert picks["t0_waves"] in (5, 6, 7, 8) and picks["queue_grain"] in (1, 2, 4, 8)
    assert picks["first_tier"] in (0, 1) and picks["first_probe"] in (0, 1)
    bt.calculate()
    for t in range(len(members)):
        assert np.array_equal(bits(bt.volumes(t)), ref[t]), t
    bt.close()
