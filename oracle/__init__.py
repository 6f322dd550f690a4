"""ctypes front end of the CPU oracle (``oracle/vd_oracle.cpp``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py``, never by the ``vorothree_amd`` package. The oracle restates
``Tessellation<3, Cell3DFaces, Algorithm3DGrid>`` of mdt-re/vorothree on the CPU
(``src/tessellation.rs``, ``src/algorithm/algo_3d_grid.rs``, ``src/cell/cell_3d_faces.rs``,
``src/wall/wall_3d.rs``); every C function there cites the reference line it follows.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p
_HOST_CONTAINS = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, _dp)
_HOST_CUT = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, _dp, _dp, ctypes.c_int)


def build() -> None:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    sig = {
        "or_tess_new": (_vp, [_dp, _dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]),
        "or_tess_free": (None, [_vp]),
        "or_tess_add_wall_plane": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, _dp]),
        "or_tess_add_wall_sphere": (ctypes.c_int, [_vp, ctypes.c_int32, _dp, ctypes.c_double]),
        "or_tess_add_wall": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int, _dp, _dp, ctypes.c_double,
                                            ctypes.c_double, _dp, ctypes.c_uint32, ctypes.c_int]),
        "or_tess_add_wall_host": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, _HOST_CONTAINS, _HOST_CUT, ctypes.c_int]),
        "or_tess_set_generators": (None, [_vp, _dp, ctypes.c_size_t]),
        "or_tess_set_generator": (ctypes.c_int, [_vp, ctypes.c_size_t, _dp]),
        "or_tess_random_generators": (None, [_vp, ctypes.c_size_t, ctypes.c_uint64]),
        "or_splitmix64_at": (ctypes.c_uint64, [ctypes.c_uint64, ctypes.c_uint64]),
        "or_tess_count_generators": (ctypes.c_size_t, [_vp]),
        "or_tess_get_generators": (None, [_vp, _dp]),
        "or_tess_get_bins": (None, [_vp, _vp, _vp, _vp]),
        "or_tess_bin_index": (ctypes.c_size_t, [_vp, _dp]),
        "or_tess_shell_order": (ctypes.c_size_t, [_vp, _vp, _vp, ctypes.c_size_t]),
        "or_tess_shell_prefix": (ctypes.c_size_t, [_vp, ctypes.c_double, _vp, _vp, ctypes.c_size_t, _vp]),
        "or_tess_shell_mode": (ctypes.c_size_t, [_vp, ctypes.c_int]),
        "or_tess_shell_len": (ctypes.c_size_t, [_vp]),
        "or_tess_calculate": (None, [_vp, _vp, ctypes.c_size_t, ctypes.c_int, ctypes.c_int]),
        "or_tess_count_cells": (ctypes.c_size_t, [_vp]),
        "or_tess_relax": (None, [_vp]),
        "or_tess_calculate_sealed": (None, [_vp, ctypes.c_int]),
        "or_tess_prune_boundaries": (None, [_vp]),
        "or_tess_seal_log": (ctypes.c_size_t, [_vp, _vp, ctypes.c_size_t]),
        "or_tess_prune_log": (None, [_vp, _vp, ctypes.c_size_t, _vp, ctypes.c_size_t, _vp]),
        "or_cell_sizes": (ctypes.c_int, [_vp, ctypes.c_size_t, _vp]),
        "or_cell_get": (ctypes.c_int, [_vp, ctypes.c_size_t, _vp, _vp, _vp, _vp]),
        "or_tess_reduced": (None, [_vp, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp]),
        "or_tess_neighbors": (None, [_vp, _vp, ctypes.c_size_t, _vp, _vp, ctypes.c_size_t]),
        "or_cell_new": (_vp, [_dp, _dp]),
        "or_cell_free": (None, [_vp]),
        "or_cell_clip": (ctypes.c_int, [_vp, _dp, _dp, ctypes.c_int32, _vp, _dp]),
        "or_cell_volume": (ctypes.c_double, [_vp]),
        "or_cell_centroid": (None, [_vp, _dp]),
        "or_cell_counts": (None, [_vp, _vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _d3(x):
    a = (ctypes.c_double * 3)(*[float(v) for v in x])
    return a


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def splitmix64_at(seed: int, k: int) -> int:
    """Output k of the SplitMix64 stream of seed (the candidates of random_generators)."""
    return int(_load().or_splitmix64_at(int(seed), int(k)))


class OracleCell:
    """A cell as returned by the oracle: the four arrays of ``Cell3DFaces`` (cell_3d_faces.rs:23-34)."""

    def __init__(self, vertices, face_counts, face_indices, face_neighbors):
        self.vertices = vertices
        self.face_counts = face_counts
        self.face_indices = face_indices
        self.face_neighbors = face_neighbors

    def volume(self) -> float:
        return cell_volume(self.vertices, self.face_counts, self.face_indices)


def cell_volume(v, fc, fi) -> float:
    """Python restatement of Cell3DFaces::volume (cell_3d_faces.rs:106-143), left fold as written."""
    vol = 0.0
    off = 0
    v = np.asarray(v, dtype=np.float64).reshape(-1, 3)
    for cnt in fc:
        cnt = int(cnt)
        if cnt >= 3:
            a = v[fi[off]]
            for i in range(1, cnt - 1):
                b = v[fi[off + i]]
                c = v[fi[off + i + 1]]
                vol += (float(a[0]) * (float(b[1]) * float(c[2]) - float(b[2]) * float(c[1]))
                        + float(a[1]) * (float(b[2]) * float(c[0]) - float(b[0]) * float(c[2]))
                        + float(a[2]) * (float(b[0]) * float(c[1]) - float(b[1]) * float(c[0])))
        off += cnt
    return abs(vol / 6.0)


class OracleTess:
    """CPU restatement of ``Tessellation::<3, Cell3DFaces, Algorithm3DGrid>``."""

    def __init__(self, lo, hi, nx, ny, nz):
        self._lib = _load()
        self.lo = tuple(float(x) for x in lo)
        self.hi = tuple(float(x) for x in hi)
        self.n = (int(nx), int(ny), int(nz))
        self._h = self._lib.or_tess_new(_d3(lo), _d3(hi), int(nx), int(ny), int(nz))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.or_tess_free(h)
            self._h = None

    def add_wall_plane(self, wid, point, normal):
        if self._lib.or_tess_add_wall_plane(self._h, int(wid), _d3(point), _d3(normal)) != 0:
            raise ValueError("Wall ID must be <= -1000")

    def add_wall_sphere(self, wid, center, radius):
        if self._lib.or_tess_add_wall_sphere(self._h, int(wid), _d3(center), float(radius)) != 0:
            raise ValueError("Wall ID must be <= -1000")

    def add_wall(self, wid, kind, a=(0.0, 0.0, 0.0), b=(0.0, 0.0, 0.0), r=0.0, r2=0.0, pts=None,
                 closed=False):
        """Any wall kind (or_tess_add_wall): 2 cylinder (centre, axis, radius), 3 cone (tip, axis,
        half-angle), 4 torus (centre, axis, major, minor), 5 convex polyhedron (pts: planes as
        point + outward normal rows), 6 tube (pts: samples, r: tube radius, closed)."""
        p = np.ascontiguousarray(np.zeros(1) if pts is None else np.asarray(pts, dtype=np.float64).reshape(-1))
        per = 6 if kind == 5 else 3
        npts = 0 if pts is None else p.size // per
        rc = self._lib.or_tess_add_wall(self._h, int(wid), int(kind), _d3(a), _d3(b), float(r), float(r2),
                                        p.ctypes.data_as(_dp), npts, 1 if closed else 0)
        if rc != 0:
            raise ValueError("Wall ID must be <= -1000" if rc == -1 else "unknown wall kind")

    def add_wall_host(self, wid, geometry):
        """Wall::new(id, Box<dyn WallGeometry>) for a Python geometry with contains(p) and cut(g)
        (a list of (point, normal)); is_planar() optional. Called back from the oracle's threads."""
        def contains(_ctx, p):
            return 1 if geometry.contains((p[0], p[1], p[2])) else 0

        def cut(_ctx, g, out, cap):
            planes = list(geometry.cut((g[0], g[1], g[2])))
            for k, (pt, nm) in enumerate(planes[:cap]):
                for c in range(3):
                    out[6 * k + c] = float(pt[c])
                    out[6 * k + 3 + c] = float(nm[c])
            return len(planes)

        cbs = (_HOST_CONTAINS(contains), _HOST_CUT(cut))
        self.__dict__.setdefault("_host", []).append(cbs)
        planar = getattr(geometry, "is_planar", None)
        if self._lib.or_tess_add_wall_host(self._h, int(wid), None, cbs[0], cbs[1],
                                           1 if (planar and planar()) else 0) != 0:
            raise ValueError("Wall ID must be <= -1000")

    def calculate_sealed(self, threads=1):
        """Tessellation::calculate_sealed (tessellation.rs:234-318)."""
        self._lib.or_tess_calculate_sealed(self._h, int(threads))

    def prune_boundaries(self):
        """Tessellation::prune_boundaries (tessellation.rs:324-385)."""
        self._lib.or_tess_prune_boundaries(self._h)

    def seal_log(self):
        n = self._lib.or_tess_seal_log(self._h, None, 0)
        out = np.zeros(n, dtype=np.int32)
        self._lib.or_tess_seal_log(self._h, out.ctypes.data_as(ctypes.c_void_p), n)
        return out.reshape(-1, 3)

    def prune_log(self):
        c = np.zeros(2, dtype=np.uint64)
        self._lib.or_tess_prune_log(self._h, None, 0, None, 0, c.ctypes.data_as(ctypes.c_void_p))
        ints = np.zeros(int(c[0]), dtype=np.int32)
        pos = np.zeros(int(c[1]), dtype=np.float64)
        self._lib.or_tess_prune_log(self._h, ints.ctypes.data_as(ctypes.c_void_p), ints.size,
                                    pos.ctypes.data_as(ctypes.c_void_p), pos.size, c.ctypes.data_as(ctypes.c_void_p))
        return ints, pos.reshape(-1, 3)

    def set_generators(self, xyz):
        a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1))
        self._lib.or_tess_set_generators(self._h, a.ctypes.data_as(_dp), a.size // 3)

    def random_generators(self, count: int, seed: int) -> None:
        """random_generators (tessellation.rs:96-122) over the SplitMix64 stream of seed."""
        self._lib.or_tess_random_generators(self._h, int(count), int(seed))

    def set_generator(self, index: int, q) -> int:
        """set_generator (tessellation.rs:71-92): 0 moved, 1 index past the generators, 2 a wall
        does not contain q (the last two change nothing)."""
        return int(self._lib.or_tess_set_generator(self._h, int(index), _d3(q)))

    def count_generators(self) -> int:
        return int(self._lib.or_tess_count_generators(self._h))

    def generators(self) -> np.ndarray:
        out = np.empty(3 * self.count_generators(), dtype=np.float64)
        self._lib.or_tess_get_generators(self._h, out.ctypes.data_as(_dp))
        return out

    def bins(self):
        n = self.count_generators()
        nb = self.n[0] * self.n[1] * self.n[2]
        bs = np.empty(nb + 1, dtype=np.uint32)
        bi = np.empty(n, dtype=np.uint32)
        gb = np.empty(n, dtype=np.uint32)
        self._lib.or_tess_get_bins(self._h, _ptr(bs), _ptr(bi), _ptr(gb))
        return bs, bi, gb

    def bin_index(self, q) -> int:
        return int(self._lib.or_tess_bin_index(self._h, _d3(q)))

    def shell_order(self, cap=None):
        total = int(self._lib.or_tess_shell_order(self._h, None, None, 0))
        m = total if cap is None else min(cap, total)
        dxyz = np.empty((m, 3), dtype=np.int32)
        d2 = np.empty(m, dtype=np.float64)
        self._lib.or_tess_shell_order(self._h, _ptr(dxyz), _ptr(d2), m)
        return dxyz, d2

    def shell_prefix(self, thr):
        """The stable shell order restricted to keys <= thr, built the way the walks extend it:
        (dxyz, keys, key of the first entry left out)."""
        nxt = np.zeros(1, dtype=np.float64)
        total = int(self._lib.or_tess_shell_prefix(self._h, float(thr), None, None, 0, _ptr(nxt)))
        dxyz = np.empty((total, 3), dtype=np.int32)
        d2 = np.empty(total, dtype=np.float64)
        self._lib.or_tess_shell_prefix(self._h, float(thr), _ptr(dxyz), _ptr(d2), total, _ptr(nxt))
        return dxyz, d2, float(nxt[0])

    def shell_mode(self, full: bool) -> int:
        """Walks start from the whole table (True) or from the lazily grown prefix (False)."""
        return int(self._lib.or_tess_shell_mode(self._h, int(bool(full))))

    def shell_len(self) -> int:
        """Length of the prefix the walks use now."""
        return int(self._lib.or_tess_shell_len(self._h))

    def calculate(self, subset=None, threads=1, brute=False):
        if subset is None:
            self._lib.or_tess_calculate(self._h, None, 0, int(threads), int(bool(brute)))
        else:
            s = np.ascontiguousarray(np.asarray(subset, dtype=np.int64))
            self._lib.or_tess_calculate(self._h, _ptr(s), s.size, int(threads), int(bool(brute)))

    def relax(self):
        self._lib.or_tess_relax(self._h)

    def count_cells(self) -> int:
        return int(self._lib.or_tess_count_cells(self._h))

    def get_cell(self, i):
        sizes = np.zeros(3, dtype=np.uint64)
        if self._lib.or_cell_sizes(self._h, int(i), _ptr(sizes)) != 0:
            return None
        nv, nf, ni = (int(x) for x in sizes)
        v = np.empty(3 * nv, dtype=np.float64)
        fc = np.empty(nf, dtype=np.uint32)
        fi = np.empty(ni, dtype=np.uint32)
        fn = np.empty(nf, dtype=np.int32)
        self._lib.or_cell_get(self._h, int(i), _ptr(v), _ptr(fc), _ptr(fi), _ptr(fn))
        return OracleCell(v, fc, fi, fn)

    def reduced(self, idx=None):
        """(volumes, centroids[n,3], face_counts, work[n,9]) for the listed cells (or all).

        work columns: flops, candidates, bins, clips, cuts, shell entries, largest vertex
        count, largest face-index count, largest face count while the cell was built."""
        if idx is None:
            m = self.count_cells()
            ip = None
        else:
            idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int64))
            m = idx.size
            ip = _ptr(idx)
        vol = np.empty(m, dtype=np.float64)
        cen = np.empty((m, 3), dtype=np.float64)
        nf = np.empty(m, dtype=np.uint32)
        work = np.empty((m, 9), dtype=np.uint64)
        self._lib.or_tess_reduced(self._h, ip, m, _ptr(vol), _ptr(cen), _ptr(nf), _ptr(work))
        return vol, cen, nf, work


    def neighbors(self, idx=None):
        """Neighbour ids of the listed cells (or all) in face order, as CSR (offsets, ids)."""
        if idx is None:
            m, ip = self.count_cells(), None
        else:
            idx = np.ascontiguousarray(np.asarray(idx, dtype=np.int64))
            m, ip = idx.size, _ptr(idx)
        off = np.empty(m + 1, dtype=np.uint64)
        self._lib.or_tess_neighbors(self._h, ip, m, _ptr(off), None, 0)
        ids = np.empty(int(off[-1]), dtype=np.int32)
        self._lib.or_tess_neighbors(self._h, ip, m, _ptr(off), _ptr(ids), ids.size)
        return off, ids


class OracleStandaloneCell:
    """A lone ``Cell3DFaces`` for the reference's cell unit tests (cell_3d_faces.rs:571-594)."""

    def __init__(self, lo, hi):
        self._lib = _load()
        self._h = self._lib.or_cell_new(_d3(lo), _d3(hi))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.or_cell_free(h)
            self._h = None

    def clip(self, point, normal, nid, generator=None):
        r = ctypes.c_double(0.0)
        g = _d3(generator) if generator is not None else None
        m = self._lib.or_cell_clip(self._h, _d3(point), _d3(normal), int(nid),
                                   ctypes.cast(g, ctypes.c_void_p) if g is not None else None,
                                   ctypes.byref(r))
        return bool(m), r.value

    def volume(self) -> float:
        return float(self._lib.or_cell_volume(self._h))

    def centroid(self):
        out = (ctypes.c_double * 3)()
        self._lib.or_cell_centroid(self._h, out)
        return [out[0], out[1], out[2]]

    def counts(self):
        s = np.zeros(3, dtype=np.uint64)
        self._lib.or_cell_counts(self._h, _ptr(s))
        return tuple(int(x) for x in s)
