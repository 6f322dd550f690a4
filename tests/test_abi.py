"""C ABI of libvoronoid_hip.so without a GPU: the library loads, exports every entry point that
include/vd.h declares, and its host-only shell-order builder equals the oracle's stable order."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from oracle import OracleTess
from vorothree_amd import BoundingBox, Wall, shell_prefix
from vorothree_amd import _lib


def declared_functions():
    text = open(os.path.join(ROOT, "include", "vd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vd_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), "Python binding and header disagree"


def test_last_error_and_invalid_arguments():
    lib = _lib.lib()
    out = ctypes.c_void_p()
    lo = (ctypes.c_double * 3)(0, 0, 0)
    hi = (ctypes.c_double * 3)(1, 1, 1)
    rc = lib.vd_tess_create(lo, hi, 0, 1, 1, ctypes.byref(out))   # zero bins: rejected before any device call
    assert rc == _lib.VD_EINVAL and not out.value
    assert b"grid" in lib.vd_last_error()


def test_wall_id_check():
    # src/wall/mod.rs:24-27 panics for id > WALL_ID_MAX
    from vorothree_amd import SphereGeometry
    with pytest.raises(ValueError):
        Wall(-999, SphereGeometry((0, 0, 0), 1.0))
    Wall(-1000, SphereGeometry((0, 0, 0), 1.0))


@pytest.mark.parametrize("grid,lo,hi,k", [
    ((10, 10, 10), (0, 0, 0), (10, 10, 10), 7),
    ((3, 4, 5), (0, 0, 0), (1, 2, 0.5), 3),
    ((6, 6, 6), (-10, -10, -10), (10, 10, 10), 40),
    ((20, 7, 9), (0, 0, 0), (1.3, 0.4, 0.9), 5),
    ((1, 1, 1), (0, 0, 0), (1, 1, 1), 2),
])
def test_shell_prefix_matches_oracle(grid, lo, hi, k):
    d, keys, nxt = shell_prefix(BoundingBox(lo, hi), *grid, k)
    o = OracleTess(lo, hi, *grid)
    od, ok = o.shell_order()
    m = len(keys)
    assert m >= 1
    assert np.array_equal(od[:m], d)
    assert np.array_equal(ok[:m].view(np.uint64), keys.view(np.uint64))
    if m < len(ok):
        assert ok[m] == nxt and ok[m] > ok[m - 1] or ok[m] == nxt
    else:
        assert np.isinf(nxt)


def test_shell_prefix_grows_to_the_full_table():
    b = BoundingBox((0, 0, 0), (1, 1, 1))
    sizes = [len(shell_prefix(b, 8, 8, 8, k)[1]) for k in (2, 4, 8, 16)]
    assert sizes == sorted(sizes) and sizes[-1] == 17 ** 3


def coarse_density(points, n_axis):
    """Density ratio and empty-bin fraction over a histogram of n_axis bins per axis (unit cube)."""
    idx = np.minimum((points * n_axis).astype(np.int64), n_axis - 1)
    c = np.bincount((idx[:, 2] * n_axis + idx[:, 1]) * n_axis + idx[:, 0], minlength=n_axis ** 3).astype(float)
    n = len(points)
    return (c * np.maximum(c - 1, 0)).sum() * n_axis ** 3 / (n * n), float((c == 0).mean())


def test_grid_for_points_rule():
    from vorothree_amd import grid_for, grid_for_points
    from vorothree_amd import workloads as W
    b = BoundingBox((0.0, 0.0, 0.0), (1.0, 1.0, 1.0))
    n = 200_000
    n_axis = grid_for(b, n, 8.0)[0]                                    # about 8 generators per coarse bin
    uni = W.PCG32(91).uniform(n, (0, 0, 0), (1, 1, 1))
    g_u = grid_for_points(b, uni, 0.35)
    g_ref = grid_for(b, n, 0.35)
    assert all(abs(a - r) <= 1 for a, r in zip(g_u, g_ref))           # uniform: the uniform rule
    # clusters over a uniform background (few empty coarse bins): per_bin / ratio^(2/3)
    bg = W.clustered_points(W.PCG32(92), n, 64, 0.02, 0.8)
    r, empty = coarse_density(bg, n_axis)
    assert 10 < r < 64 and empty < 0.25
    g_c = grid_for_points(b, bg, 0.35)
    g_exp = grid_for(b, n, 0.35 / r ** (2.0 / 3.0))
    assert all(abs(a - e) <= 1 for a, e in zip(g_c, g_exp)) and g_c[0] > 2 * g_ref[0]
    # clusters in empty space: three times the target, a coarser grid than uniform input gets
    void = W.clustered_points(W.PCG32(94), n, 8, 0.05, 1.0)
    assert coarse_density(void, n_axis)[1] > 0.25
    assert grid_for_points(b, void, 0.35) == grid_for(b, n, 1.05)
    small = uni[:40]
    assert grid_for_points(b, small, 0.35) == grid_for(b, 40, 0.35)    # too few points to measure
    box = BoundingBox((0.0, 0.0, 0.0), (2.0, 1.0, 0.5))
    g_a = grid_for_points(box, W.PCG32(93).uniform(100_000, (0, 0, 0), (2.0, 1.0, 0.5)), 0.35)
    assert g_a[0] > g_a[1] > g_a[2]                                    # proportional to the edges


def test_static_entry_rejects_sizes_beyond_its_maximum():
    # vd_calculate_static checks its limits before it touches a device
    from vorothree_amd import calculate_static
    from vorothree_amd._lib import VD_EINVAL, VD_STATIC_MAX_BINS, VD_STATIC_MAX_POINTS, VdError
    b = BoundingBox((0, 0, 0), (1, 1, 1))
    with pytest.raises(VdError) as e1:
        calculate_static(b, (10, 10, 10), [], 0x1000, VD_STATIC_MAX_POINTS + 1, 0x1000, 0x1000, 0x1000)
    assert e1.value.code == VD_EINVAL and "at most" in str(e1.value)
    k = int(round(VD_STATIC_MAX_BINS ** (1 / 3))) + 1
    with pytest.raises(VdError) as e2:
        calculate_static(b, (k, k, k), [], 0x1000, 100, 0x1000, 0x1000, 0x1000)
    assert e2.value.code == VD_EINVAL and "bins" in str(e2.value)
