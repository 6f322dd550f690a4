"""Seeded input generators (vorothree_amd/workloads.py)."""
import numpy as np

from vorothree_amd import workloads as W


def test_pcg32_known_answers():
    # pcg32-demo (O'Neill): seed 42, stream 54
    g = W.PCG32(42, 54)
    assert [int(x) for x in g.next_u32(6)] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293,
                                                0xBFA4784B, 0xCBED606E]


def test_pcg32_blockwise_equals_stepwise():
    a = W.PCG32(7).next_u32(10000)
    g = W.PCG32(7)
    b = np.concatenate([g.next_u32(1), g.next_u32(4095), g.next_u32(1), g.next_u32(5903)])
    assert np.array_equal(a, b)


def test_u01_range_and_uniform_points():
    g = W.PCG32(3)
    u = g.u01(100000)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    p = W.PCG32(1).uniform(1000, (0, 0, 0), (10, 10, 10))
    assert p.shape == (1000, 3) and p.min() >= 0.0 and p.max() < 10.0


def test_configs_shapes():
    w1 = W.config1()
    assert w1.points.shape == (1000, 3) and w1.grid == (10, 10, 10)
    w2 = W.config2(n=1000, grid=10)
    assert w2.walls[0][0] == "sphere"
    pts = W.clustered_points(W.PCG32(3), 5000, 8, 0.02, 0.8)
    assert pts.shape == (5000, 3) and pts.min() >= 0.0 and pts.max() <= 1.0
    batch = W.config4(count=50)
    ns = [b.points.shape[0] for b in batch]
    assert min(ns) >= 64 and max(ns) <= 65536
    for b in batch:
        assert all(x >= 1 for x in b.grid)
