"""Seeded synthetic inputs for the tessellation workloads.

The reference has no seeded generator (``random_generators`` seeds from ``thread_rng``,
src/tessellation.rs:96-97,529-538), so the workloads bring their own:

* PCG32 (O'Neill 2014, XSH-RR 64/32): multiplier 6364136223846793005, increment
  ``(stream << 1) | 1``, seeded as ``pcg32_srandom_r(initstate=seed, initseq=stream)``.
* ``u01``: two consecutive 32-bit draws ``hi, lo`` form ``u64 = hi << 32 | lo``;
  ``u01 = (u64 >> 11) * 2**-53`` in [0, 1).
* Coordinates ``min + u01 * (max - min)``, the affine form of src/tessellation.rs:110.
* Gaussians by the Marsaglia polar method; out-of-box draws are re-drawn.

Everything is vectorised with numpy: the LCG state sequence is produced block-wise by the
affine jump ``s[k + B] = A_B * s[k] + C_B (mod 2**64)``.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_MULT = 6364136223846793005
_MASK = (1 << 64) - 1


def _affine_pow(a: int, c: int, k: int):
    """(A, C) of k steps of x -> a*x + c (mod 2**64)."""
    ra, rc = 1, 0
    ba, bc = a, c
    while k:
        if k & 1:
            ra, rc = (ba * ra) & _MASK, (ba * rc + bc) & _MASK
        ba, bc = (ba * ba) & _MASK, (ba * bc + bc) & _MASK
        k >>= 1
    return ra, rc


class PCG32:
    """Vectorised PCG32 (XSH-RR). ``next_u32(n)`` returns the next n outputs in order."""

    def __init__(self, seed: int, stream: int | None = None):
        if stream is None:
            stream = seed
        self.inc = ((int(stream) << 1) | 1) & _MASK
        state = 0
        state = (state * _MULT + self.inc) & _MASK
        state = (state + int(seed)) & _MASK
        state = (state * _MULT + self.inc) & _MASK
        self.state = state

    def _states(self, n: int) -> np.ndarray:
        """The n 'old' states consumed by the next n draws; advances the generator."""
        out = np.empty(n, dtype=np.uint64)
        block = min(n, 4096)
        s = self.state
        for k in range(block):
            out[k] = s
            s = (s * _MULT + self.inc) & _MASK
        if n > block:
            a, c = _affine_pow(_MULT, self.inc, block)
            a64 = np.uint64(a)
            c64 = np.uint64(c)
            pos = block
            with np.errstate(over="ignore"):
                while pos < n:
                    m = min(block, n - pos)
                    out[pos:pos + m] = out[pos - block:pos - block + m] * a64 + c64
                    pos += m
        a, c = _affine_pow(_MULT, self.inc, n)
        self.state = (a * self.state + c) & _MASK
        return out

    def next_u32(self, n: int) -> np.ndarray:
        old = self._states(n)
        xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)) & np.uint64(0xFFFFFFFF)
        rot = (old >> np.uint64(59)).astype(np.uint64)
        left = (np.uint64(32) - rot) & np.uint64(31)
        r = ((xs >> rot) | (xs << left)) & np.uint64(0xFFFFFFFF)
        return r.astype(np.uint32)

    def u01(self, n: int) -> np.ndarray:
        d = self.next_u32(2 * n).astype(np.uint64)
        u64 = (d[0::2] << np.uint64(32)) | d[1::2]
        return (u64 >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)

    def uniform(self, n: int, lo, hi) -> np.ndarray:
        """n points, f64 xyz AoS (shape [n, 3]), each coordinate ``lo + u * (hi - lo)``."""
        lo = np.asarray(lo, dtype=np.float64)
        hi = np.asarray(hi, dtype=np.float64)
        u = self.u01(3 * n).reshape(n, 3)
        return lo + u * (hi - lo)

    def gauss(self, n: int) -> np.ndarray:
        """n standard normals by the Marsaglia polar method (pairs; rejection in batches)."""
        out = np.empty(0, dtype=np.float64)
        while out.size < n:
            need = n - out.size
            m = int(need * 0.7) + 16
            u = 2.0 * self.u01(2 * m).reshape(m, 2) - 1.0
            s = u[:, 0] * u[:, 0] + u[:, 1] * u[:, 1]
            ok = (s > 0.0) & (s < 1.0)
            u, s = u[ok], s[ok]
            f = np.sqrt(-2.0 * np.log(s) / s)
            z = np.stack([u[:, 0] * f, u[:, 1] * f], axis=1).reshape(-1)
            out = np.concatenate([out, z])
        return out[:n]


@dataclass
class Workload:
    """One tessellation: box, grid resolution, walls and generators (f64 xyz AoS)."""
    name: str
    lo: tuple
    hi: tuple
    grid: tuple
    points: np.ndarray
    walls: list = field(default_factory=list)   # ("sphere", id, centre, r) | ("plane", id, p, n)


def config1() -> Workload:
    """Box [0,10]^3, 1,000 uniform generators, seed 1, grid 10^3, no walls."""
    g = PCG32(1)
    return Workload("plain-1k-box10", (0.0, 0.0, 0.0), (10.0, 10.0, 10.0), (10, 10, 10),
                    g.uniform(1000, (0, 0, 0), (10, 10, 10)))


def config2(n: int = 1_000_000, grid: int = 100) -> Workload:
    """Unit cube, 1M uniform generators, seed 2, grid 100^3, sphere wall c=(0.5,0.5,0.5) r=0.5."""
    g = PCG32(2)
    return Workload(f"uniform-{n}-sphere", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (grid, grid, grid),
                    g.uniform(n, (0, 0, 0), (1, 1, 1)),
                    [("sphere", -1000, (0.5, 0.5, 0.5), 0.5)])


def clustered_points(g: PCG32, n: int, nclusters: int, sigma: float, frac_cluster: float,
                     lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)) -> np.ndarray:
    """frac_cluster of n from isotropic Gaussian clusters (centres uniform in the box, point k in
    cluster k % nclusters, out-of-box draws re-drawn), then the rest uniform."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    nc = int(round(n * frac_cluster))
    centres = g.uniform(nclusters, lo, hi)
    pts = np.empty((nc, 3), dtype=np.float64)
    which = np.arange(nc) % nclusters
    todo = np.arange(nc)
    while todo.size:
        z = g.gauss(3 * todo.size).reshape(-1, 3)
        cand = centres[which[todo]] + sigma * z
        ok = np.all((cand >= lo) & (cand <= hi), axis=1)
        pts[todo[ok]] = cand[ok]
        todo = todo[~ok]
    uni = g.uniform(n - nc, lo, hi)
    return np.concatenate([pts, uni], axis=0)


def config3(n: int = 4_000_000, grid=None) -> Workload:
    """Unit cube, 4M generators: 80% from 64 Gaussian clusters (sigma 0.02) + 20% uniform, seed 3."""
    g = PCG32(3)
    pts = clustered_points(g, n, 64, 0.02, 0.8)
    if grid is None:
        k = max(1, int(round((n / 2.0) ** (1.0 / 3.0))))
        grid = (k, k, k)
    return Workload(f"clustered-{n}", (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), tuple(grid), pts)


def grid_for(n: int, lo, hi, per_bin: float = 2.0) -> tuple:
    """Bins per axis proportional to the box edges, about ``per_bin`` generators per bin."""
    e = np.asarray(hi, dtype=np.float64) - np.asarray(lo, dtype=np.float64)
    vol = float(np.prod(e))
    h = (vol * per_bin / max(n, 1)) ** (1.0 / 3.0)
    return tuple(int(max(1, round(x / h))) for x in e)


def config4(count: int = 4096, seed: int = 4, nmin: int = 64, nmax: int = 65536,
            per_bin: float = 2.0) -> list:
    """Batched workload: ``count`` independent tessellations, generator counts log-uniform in
    [nmin, nmax], boxes at the origin with edges U[0.5, 2] per axis, 30% with a sphere wall
    (centre = box centre, r = 0.45 * min edge), 20% with one plane wall (point uniform in the box,
    normal uniform on S^2), the rest box-only; grid about ``per_bin`` generators per bin."""
    g = PCG32(seed)
    out = []
    ratio = math.log(nmax / nmin)
    for t in range(count):
        u = g.u01(1)[0]
        n = int(math.floor(nmin * math.exp(u * ratio)))
        n = min(max(n, nmin), nmax)
        edges = 0.5 + 1.5 * g.u01(3)
        lo = (0.0, 0.0, 0.0)
        hi = tuple(float(x) for x in edges)
        kind = g.u01(1)[0]
        walls = []
        if kind < 0.3:
            c = tuple(0.5 * x for x in hi)
            walls.append(("sphere", -1000, c, 0.45 * float(min(hi))))
        elif kind < 0.5:
            p = tuple(float(x) for x in g.uniform(1, lo, hi)[0])
            z = g.gauss(3)
            nrm = z / np.sqrt(np.sum(z * z))
            walls.append(("plane", -1000, p, tuple(float(x) for x in nrm)))
        pts = g.uniform(n, lo, hi)
        out.append(Workload(f"batch-{t}", lo, hi, grid_for(n, lo, hi, per_bin), pts, walls))
    return out


@dataclass
class Request:
    """One request of the served workload."""
    arrival: float          # seconds after the stream starts
    points: np.ndarray
    clustered: bool
    iterations: int         # 1 = plain tessellation, else Lloyd iterations


def config5(count: int = 100, rate: float = 2000.0, seed: int = 5, nmin: int = 1000,
            nmax: int = 2_000_000) -> list:
    """Served workload: Poisson arrivals at ``rate`` requests/s (inter-arrival -ln(1-u)/rate),
    generator counts log-uniform in [nmin, nmax] in the unit cube, 50% uniform / 50% from 8
    Gaussian clusters (sigma 0.05, out-of-box draws re-drawn), 30% Lloyd jobs of U{10..50}
    iterations. Only the first ``count`` requests of the stream are generated."""
    g = PCG32(seed)
    out = []
    t = 0.0
    ratio = math.log(nmax / nmin)
    for _ in range(count):
        u = g.u01(4)
        t += -math.log(1.0 - u[0]) / rate
        n = min(max(int(math.floor(nmin * math.exp(u[1] * ratio))), nmin), nmax)
        clustered = u[2] < 0.5
        iters = 1
        if u[3] < 0.3:
            iters = 10 + int(g.u01(1)[0] * 41)
        pts = clustered_points(g, n, 8, 0.05, 1.0) if clustered else g.uniform(n, (0, 0, 0), (1, 1, 1))
        out.append(Request(t, pts, clustered, iters))
    return out
