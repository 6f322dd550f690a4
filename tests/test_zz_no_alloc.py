"""vd_calculate_static (the entry point that allocates no device memory) against the handle path
(vd_tess_create + add_wall + set_generators_device + calculate): bit-identical volumes, centroids
and face counts on two calls in a row, the second smaller, both in the one static scratch array."""
import numpy as np
import pytest

from conftest import bits
from vorothree_amd import (Algorithm3DGrid, BoundingBox, PlaneGeometry, SphereGeometry, Tessellation, Wall,
                           _hip, calculate_static)
from vorothree_amd import workloads as W

pytestmark = pytest.mark.gpu


def handle_path(b, grid, walls, pts):
    t = Tessellation(b, Algorithm3DGrid(*grid, b), keep_cells=False)
    for w in walls:
        t.add_wall(w)
    dev = _hip.DeviceArray(pts)
    t.set_generators_device(dev.ptr.value, len(pts))
    t.calculate()
    out = (t.volumes(), t.centroids(), t.face_counts())
    t.close()
    dev.close()
    return out


def static_path(b, grid, walls, pts, d_vol, d_cen, d_nf):
    dev = _hip.DeviceArray(pts)
    kept = calculate_static(b, grid, walls, dev.ptr.value, len(pts), d_vol.ptr.value, d_cen.ptr.value,
                            d_nf.ptr.value)
    out = (d_vol.download(np.float64, kept), d_cen.download(np.float64, 3 * kept).reshape(-1, 3),
           d_nf.download(np.uint32, kept))
    dev.close()
    return out


def test_static_entry_equals_handle_path_twice():
    cases = []
    w = W.config2(n=200_000, grid=58)
    cases.append((BoundingBox(w.lo, w.hi), w.grid, [Wall(-1000, SphereGeometry((0.5, 0.5, 0.5), 0.5))],
                  w.points))
    pts = W.PCG32(41).uniform(30_000, (0, 0, 0), (2.0, 1.0, 1.5))
    cases.append((BoundingBox((0, 0, 0), (2.0, 1.0, 1.5)), (26, 13, 20),
                  [Wall(-1001, PlaneGeometry((1.0, 0.5, 0.7), (0.3, -1.0, 0.2)))], pts))
    n0 = max(len(c[3]) for c in cases)
    d_vol, d_cen, d_nf = (_hip.DeviceArray.empty(8 * n0), _hip.DeviceArray.empty(24 * n0),
                          _hip.DeviceArray.empty(4 * n0))
    for b, grid, walls, p in cases:
        ref = handle_path(b, grid, walls, p)
        got = static_path(b, grid, walls, p, d_vol, d_cen, d_nf)
        assert len(got[0]) == len(ref[0]) > 0
        assert np.array_equal(bits(got[0]), bits(ref[0]))
        assert np.array_equal(bits(got[1]), bits(ref[1]))
        assert np.array_equal(got[2], ref[2])
