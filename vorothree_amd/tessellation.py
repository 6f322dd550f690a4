"""Python mirror of voronoid's public API for the 3D grid path, over the HIP C ABI.

Names, argument meaning and error behaviour follow the reference crate (mdt-re/vorothree):

* ``BoundingBox``                  src/bounds/mod.rs:2-12, ``box_side`` :21-23
* ``Algorithm3DGrid(nx, ny, nz, bounds)``  src/algorithm/algo_3d_grid.rs:47-95
* ``Wall(id, geometry)``           src/wall/mod.rs:24-32 (id > WALL_ID_MAX raises, the panic)
* ``PlaneGeometry``, ``SphereGeometry``   src/wall/wall_3d.rs:8-101
* ``Tessellation(bounds, algorithm)``     src/tessellation.rs:13-36 with ``set_generators`` :43-67,
  ``add_wall`` :197-200, ``clear_walls`` :203-205, ``calculate`` :213-230, ``map`` :391-412,
  ``relax`` :482-495, ``count_generators``/``count_cells`` :498-505, ``get_cell`` :514-516
  (None when out of range), ``generators`` :519-521
* ``Cell3DFaces`` accessors        src/cell/cell_3d_faces.rs:76-99, ``volume`` :106-143,
  ``centroid`` :145-205, ``face_area`` :207-254, ``is_empty`` :528-530

Every computation runs in libvoronoid_hip.so on the GPU; the Cell3DFaces methods evaluate the
returned cell on the host with the reference formulas.
"""
from __future__ import annotations

import ctypes
import math
import re
import secrets
from typing import Optional

import numpy as np

from . import _lib
from ._lib import check, lib

WALL_ID_MAX = -1000


def box_side(axis: int, is_max: bool) -> int:
    """Id of a bounding-box face: x- -1, x+ -2, y- -3, y+ -4, z- -5, z+ -6."""
    return -1 - (axis * 2 + (1 if is_max else 0))


def _d3(v):
    return (ctypes.c_double * 3)(*[float(x) for x in v])


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


class BoundingBox:
    def __init__(self, min, max):  # noqa: A002 - reference field names
        self.min = [float(x) for x in min]
        self.max = [float(x) for x in max]


class Algorithm3DGrid:
    """Uniform bin grid; the index itself is built on the device by ``set_generators``."""

    def __init__(self, nx: int, ny: int, nz: int, bounds: BoundingBox):
        self.grid_res_x, self.grid_res_y, self.grid_res_z = int(nx), int(ny), int(nz)
        self.bounds = bounds

    def get_bin_index(self, x: float, y: float, z: float, bounds: BoundingBox) -> int:
        """algo_3d_grid.rs:98-107 (host evaluation, same arithmetic as the device binning)."""
        n = (self.grid_res_x, self.grid_res_y, self.grid_res_z)
        idx = []
        for a, v in enumerate((x, y, z)):
            s = n[a] / (bounds.max[a] - bounds.min[a])
            t = (v - bounds.min[a]) * s
            lim = n[a] - 1e-5
            if t < 0.0:
                t = 0.0
            if t > lim:
                t = lim
            idx.append(0 if t != t else int(t))
        return idx[0] + idx[1] * n[0] + idx[2] * n[0] * n[1]


def _f3(v):
    return [float(x) for x in v]


class PlaneGeometry:
    """PlaneGeometry::new (src/wall/wall_3d.rs:23-27); the library normalises the normal."""
    kind = _lib.VD_WALL_PLANE

    def __init__(self, point, normal):
        self.point = _f3(point)
        self.normal_raw = _f3(normal)

    def _fill(self, w):
        w.a[:] = self.point
        w.b[:] = self.normal_raw


class SphereGeometry:
    """SphereGeometry::new (wall_3d.rs:67-69)."""
    kind = _lib.VD_WALL_SPHERE

    def __init__(self, center, radius):
        self.center = _f3(center)
        self.radius = float(radius)

    def _fill(self, w):
        w.a[:] = self.center
        w.r = self.radius


class CylinderGeometry:
    """CylinderGeometry::new (wall_3d.rs:124-128): infinite cylinder, valid region inside."""
    kind = _lib.VD_WALL_CYLINDER

    def __init__(self, center, axis, radius):
        self.center, self.axis_raw, self.radius = _f3(center), _f3(axis), float(radius)

    def _fill(self, w):
        w.a[:] = self.center
        w.b[:] = self.axis_raw
        w.r = self.radius


class ConeGeometry:
    """ConeGeometry::new (wall_3d.rs:194-198): infinite cone from `tip` along `axis` with
    half-angle `angle` (radians), valid region inside."""
    kind = _lib.VD_WALL_CONE

    def __init__(self, tip, axis, angle):
        self.tip, self.axis_raw, self.angle = _f3(tip), _f3(axis), float(angle)

    def _fill(self, w):
        w.a[:] = self.tip
        w.b[:] = self.axis_raw
        w.r = self.angle


class TorusGeometry:
    """TorusGeometry::new (wall_3d.rs:285-289): valid region inside the tube."""
    kind = _lib.VD_WALL_TORUS

    def __init__(self, center, axis, major_radius, minor_radius):
        self.center, self.axis_raw = _f3(center), _f3(axis)
        self.major_radius, self.minor_radius = float(major_radius), float(minor_radius)

    def _fill(self, w):
        w.a[:] = self.center
        w.b[:] = self.axis_raw
        w.r = self.major_radius
        w.r2 = self.minor_radius


class ConvexPolyhedronGeometry:
    """ConvexPolyhedronGeometry::new(points, normals) (wall_3d.rs:476-491): planes (point,
    outward normal), used as given."""
    kind = _lib.VD_WALL_POLYHEDRON

    def __init__(self, points, normals):
        p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        n = np.asarray(normals, dtype=np.float64).reshape(-1, 3)
        if p.shape != n.shape:
            raise ValueError("Points and normals must have same length and be multiple of 3")
        self.planes = np.ascontiguousarray(np.concatenate([p, n], axis=1))

    @classmethod
    def new_tetrahedron(cls, center, radius):
        """wall_3d.rs:499-529: faces at the inradius R / 3 with normals (+-1, +-1, +-1)/sqrt(3)."""
        dist = float(radius) / 3.0
        pts, nrm = [], []
        for n in ([-1.0, -1.0, -1.0], [-1.0, 1.0, 1.0], [1.0, -1.0, 1.0], [1.0, 1.0, -1.0]):
            ln = math.sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2])
            u = [n[0] / ln, n[1] / ln, n[2] / ln]
            nrm.append(u)
            pts.append([center[0] + u[0] * dist, center[1] + u[1] * dist, center[2] + u[2] * dist])
        return cls(pts, nrm)

    @classmethod
    def new_hexahedron(cls, center, radius):
        """wall_3d.rs:534-560: axis-aligned cube with inradius R / sqrt(3)."""
        dist = float(radius) / math.sqrt(3.0)
        pts, nrm = [], []
        for n in ([1.0, 0.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, -1.0, 0.0],
                  [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]):
            nrm.append(n)
            pts.append([center[0] + n[0] * dist, center[1] + n[1] * dist, center[2] + n[2] * dist])
        return cls(pts, nrm)

    def _fill(self, w):
        w.pts = self.planes.ctypes.data
        w.npts = len(self.planes)


class _Tube:
    """Tube of radius tube_radius around a sampled curve (the closest point on the sample
    polyline defines contains and cut: wall_3d.rs:408-456, 762-815, 908-961)."""
    kind = _lib.VD_WALL_TUBE

    def _fill(self, w):
        w.pts = self.samples.ctypes.data if len(self.samples) else None
        w.npts = len(self.samples)
        w.r = self.tube_radius
        w.closed = 1 if self.closed else 0


class TrefoilKnotGeometry(_Tube):
    """TrefoilKnotGeometry::new (wall_3d.rs:390-407); the polyline is closed."""

    def __init__(self, center, scale, tube_radius, resolution):
        c = _f3(center)
        s = []
        for i in range(int(resolution)):
            t = (i / int(resolution)) * math.tau
            x = math.sin(t) + 2.0 * math.sin(2.0 * t)
            y = math.cos(t) - 2.0 * math.cos(2.0 * t)
            z = -math.sin(3.0 * t)
            s.append([c[0] + x * scale, c[1] + y * scale, c[2] + z * scale])
        self.samples = np.ascontiguousarray(np.asarray(s, dtype=np.float64).reshape(-1, 3))
        self.tube_radius, self.closed = float(tube_radius), True


class CubicBezierGeometry(_Tube):
    """CubicBezierGeometry::new (wall_3d.rs:739-760): resolution + 1 samples of the curve."""

    def __init__(self, p0, p1, p2, p3, tube_radius, resolution, closed):
        p0, p1, p2, p3 = _f3(p0), _f3(p1), _f3(p2), _f3(p3)
        s = []
        for i in range(int(resolution) + 1):
            t = i / int(resolution)
            t2 = t * t
            t3 = t2 * t
            mt = 1.0 - t
            mt2 = mt * mt
            mt3 = mt2 * mt
            s.append([mt3 * p0[k] + 3.0 * mt2 * t * p1[k] + 3.0 * mt * t2 * p2[k] + t3 * p3[k] for k in range(3)])
        self.samples = np.ascontiguousarray(np.asarray(s, dtype=np.float64).reshape(-1, 3))
        self.tube_radius, self.closed = float(tube_radius), bool(closed)


def _rem(a: int, b: int) -> int:
    """Rust's % on signed integers (the sign of the dividend)."""
    r = abs(a) % b
    return -r if a < 0 else r


class CatmullRomGeometry(_Tube):
    """CatmullRomGeometry::new (wall_3d.rs:831-906, 963-995): centripetal Catmull-Rom samples."""

    def __init__(self, points, tube_radius, resolution, closed):
        pts = [_f3(p) for p in points]
        s = []
        if len(pts) >= 2:
            for i in range(int(resolution) + 1):
                s.append(self._point(i / int(resolution), pts, bool(closed)))
        self.samples = np.ascontiguousarray(np.asarray(s, dtype=np.float64).reshape(-1, 3))
        self.tube_radius, self.closed = float(tube_radius), bool(closed)

    @staticmethod
    def _point(t, points, closed):
        l = len(points)
        p = (l - (0.0 if closed else 1.0)) * t
        ip = math.floor(p)
        weight = p - ip
        if closed:
            if ip <= 0:
                ip += (abs(ip) // l + 1) * l
        elif weight == 0.0 and ip == l - 1:
            ip = l - 2
        if closed or ip > 0:
            p0 = points[(_rem(ip - 1, l) + l) % l]
        else:
            p0 = [points[0][k] - (points[1][k] - points[0][k]) for k in range(3)]
        p1 = points[ip % l]
        p2 = points[(ip + 1) % l]
        if closed or ip + 2 < l:
            p3 = points[(ip + 2) % l]
        else:
            last, prev = points[l - 1], points[l - 2]
            p3 = [last[k] - (prev[k] - last[k]) for k in range(3)]

        def dsq(a, b):
            dx, dy, dz = a[0] - b[0], a[1] - b[1], a[2] - b[2]
            return dx * dx + dy * dy + dz * dz

        dt0 = math.pow(dsq(p0, p1), 0.25)
        dt1 = math.pow(dsq(p1, p2), 0.25)
        dt2 = math.pow(dsq(p2, p3), 0.25)
        if dt1 < 1e-4:
            dt1 = 1.0
        if dt0 < 1e-4:
            dt0 = dt1
        if dt2 < 1e-4:
            dt2 = dt1
        out = []
        for k in range(3):
            x0, x1, x2, x3 = p0[k], p1[k], p2[k], p3[k]
            t1 = (x1 - x0) / dt0 - (x2 - x0) / (dt0 + dt1) + (x2 - x1) / dt1
            t2 = (x2 - x1) / dt1 - (x3 - x1) / (dt1 + dt2) + (x3 - x2) / dt2
            t1 *= dt1
            t2 *= dt1
            c0, c1 = x1, t1
            c2 = -3.0 * x1 + 3.0 * x2 - 2.0 * t1 - t2
            c3 = 2.0 * x1 - 2.0 * x2 + t1 + t2
            w2 = weight * weight
            w3 = w2 * weight
            out.append(c0 + c1 * weight + c2 * w2 + c3 * w3)
        return out


_GEOMETRIES = (PlaneGeometry, SphereGeometry, CylinderGeometry, ConeGeometry, TorusGeometry,
               ConvexPolyhedronGeometry, _Tube)


# White_Space characters of Unicode, which Rust's str::split_whitespace and str::trim split on (no
# \x1c-\x1f, which Python's str.split also takes)
_RUST_WS = re.compile("[\t\n\x0b\x0c\r \x85\xa0\u1680\u2000-\u200a\u2028\u2029\u202f\u205f\u3000]+")
_RUST_F64 = re.compile(r"[+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf|infinity|nan)", re.IGNORECASE)


def _rust_f64(tok: str) -> float:
    """str::parse::<f64> or 0.0 where it fails (tessellation.rs:163: .ok().unwrap_or(0.0)):
    decimal and exponent forms, inf, infinity and nan in any case, with an optional sign; no
    underscores or hexadecimal, which Python's float() would take or reject differently."""
    return float(tok) if _RUST_F64.fullmatch(tok) else 0.0


def parse_generators(text: str) -> np.ndarray:
    """read_generators' parser (src/tessellation.rs:153-171): one point per line (lines split at
    \n, a trailing \r dropped, as str::lines does) that is not blank; its first token, the id, is
    skipped and the next three are x, y and z, 0.0 where missing or unparsable. Returns (n, 3)."""
    pts = []
    for line in text.split("\n"):
        if line.endswith("\r"):
            line = line[:-1]
        parts = [p for p in _RUST_WS.split(line) if p]
        if not parts:
            continue
        pts.append([_rust_f64(parts[k]) if k < len(parts) else 0.0 for k in (1, 2, 3)])
    return np.asarray(pts, dtype=np.float64).reshape(-1, 3)


def _is_host_geometry(g) -> bool:
    """A geometry of the caller's (the WallGeometry trait, src/wall/mod.rs:53-71): ``contains(p)``
    and ``cut(g)`` returning the (point, normal) planes for generator g, normals pointing out of
    the valid region; ``is_planar()`` is optional (False by default, as in the trait)."""
    return callable(getattr(g, "contains", None)) and callable(getattr(g, "cut", None))


class HostWallCallbacks:
    """ctypes callbacks (struct vd_wall_host) around a geometry of the caller's. An exception raised
    by the geometry is kept and re-raised by ``check_error`` after the library call returns."""

    def __init__(self, geometry):
        self.geometry = geometry
        self.error = None

        def contains(_ctx, p):
            try:
                return 1 if self.geometry.contains((p[0], p[1], p[2])) else 0
            except Exception as e:   # noqa: BLE001 - re-raised by check_error
                self.error = self.error or e
                return 0

        def cut(_ctx, g, out, cap):
            try:
                planes = list(self.geometry.cut((g[0], g[1], g[2])))
            except Exception as e:   # noqa: BLE001 - re-raised by check_error
                self.error = self.error or e
                return 0
            for k, (pt, nm) in enumerate(planes[:cap]):
                for c in range(3):
                    out[6 * k + c] = float(pt[c])
                    out[6 * k + 3 + c] = float(nm[c])
            return len(planes)

        self._contains = _lib.HOST_CONTAINS(contains)
        self._cut = _lib.HOST_CUT(cut)
        planar = getattr(geometry, "is_planar", None)
        self.struct = _lib.VdWallHost(None, self._contains, self._cut, 1 if (planar and planar()) else 0)

    def check_error(self):
        if self.error is not None:
            e, self.error = self.error, None
            raise e


class Wall:
    def __init__(self, id: int, geometry):  # noqa: A002
        if id > WALL_ID_MAX:
            raise ValueError(f"Wall ID must be <= {WALL_ID_MAX}")
        if not isinstance(geometry, _GEOMETRIES) and not _is_host_geometry(geometry):
            raise TypeError("a wall geometry is a built-in one or has contains(p) and cut(g)")
        self._id = int(id)
        self.geometry = geometry

    def id(self) -> int:
        return self._id


class Cell3DFaces:
    """A computed cell: the four arrays of the reference struct (cell_3d_faces.rs:23-34)."""

    def __init__(self, id, vertices, face_counts, face_indices, face_neighbors):  # noqa: A002
        self._id = id
        self._v = vertices
        self._fc = face_counts
        self._fi = face_indices
        self._fn = face_neighbors

    def id(self) -> int:
        return self._id

    def vertices(self) -> np.ndarray:
        return self._v.copy()

    def face_counts(self) -> np.ndarray:
        return self._fc.copy()

    def face_indices(self) -> np.ndarray:
        return self._fi.copy()

    def face_neighbors(self) -> np.ndarray:
        return self._fn.copy()

    def neighbors(self) -> np.ndarray:
        return self._fn

    def is_empty(self) -> bool:
        return self._v.size == 0

    def _fan(self):
        v = self._v.reshape(-1, 3)
        off = 0
        for cnt in self._fc:
            cnt = int(cnt)
            if cnt >= 3:
                a = v[self._fi[off]]
                for i in range(1, cnt - 1):
                    yield a, v[self._fi[off + i]], v[self._fi[off + i + 1]]
            off += cnt

    def volume(self) -> float:
        vol = 0.0
        for a, b, c in self._fan():
            a0, a1, a2 = float(a[0]), float(a[1]), float(a[2])
            b0, b1, b2 = float(b[0]), float(b[1]), float(b[2])
            c0, c1, c2 = float(c[0]), float(c[1]), float(c[2])
            vol += a0 * (b1 * c2 - b2 * c1) + a1 * (b2 * c0 - b0 * c2) + a2 * (b0 * c1 - b1 * c0)
        return abs(vol / 6.0)

    def centroid(self):
        cx = cy = cz = tv = 0.0
        for a, b, c in self._fan():
            a0, a1, a2 = float(a[0]), float(a[1]), float(a[2])
            b0, b1, b2 = float(b[0]), float(b[1]), float(b[2])
            c0, c1, c2 = float(c[0]), float(c[1]), float(c[2])
            crx = b1 * c2 - b2 * c1
            cry = b2 * c0 - b0 * c2
            crz = b0 * c1 - b1 * c0
            det = a0 * crx + a1 * cry + a2 * crz
            tv += det
            cx += det * (a0 + b0 + c0)
            cy += det * (a1 + b1 + c1)
            cz += det * (a2 + b2 + c2)
        if abs(tv) < 1e-9:
            return [0.0, 0.0, 0.0]
        f = 1.0 / (4.0 * tv)
        return [cx * f, cy * f, cz * f]

    def face_area(self, face_index: int) -> float:
        if face_index >= len(self._fc):
            return 0.0
        off = int(sum(int(c) for c in self._fc[:face_index]))
        cnt = int(self._fc[face_index])
        if cnt < 3:
            return 0.0
        v = self._v.reshape(-1, 3)
        p0 = v[self._fi[off]]
        area = 0.0
        for i in range(1, cnt - 1):
            p1 = v[self._fi[off + i]]
            p2 = v[self._fi[off + i + 1]]
            u = [float(p1[k]) - float(p0[k]) for k in range(3)]
            w = [float(p2[k]) - float(p0[k]) for k in range(3)]
            cx = u[1] * w[2] - u[2] * w[1]
            cy = u[2] * w[0] - u[0] * w[2]
            cz = u[0] * w[1] - u[1] * w[0]
            area += 0.5 * math.sqrt(cx * cx + cy * cy + cz * cz)
        return area


class Tessellation:
    """``Tessellation::<3, Cell3DFaces, Algorithm3DGrid>`` on the GPU."""

    def __init__(self, bounds: BoundingBox, algorithm: Algorithm3DGrid, keep_cells: bool = True,
                 graphs: bool = True):
        self.bounds = bounds
        self.algorithm = algorithm
        self._walls = []
        self._host = []   # callbacks of the caller's wall geometries, alive while the handle holds them
        h = ctypes.c_void_p()
        check(lib().vd_tess_create(_d3(bounds.min), _d3(bounds.max), algorithm.grid_res_x,
                                   algorithm.grid_res_y, algorithm.grid_res_z, ctypes.byref(h)))
        self._h = h
        if not keep_cells:
            check(lib().vd_tess_set_option(self._h, _lib.VD_OPT_KEEP_CELLS, 0))
        if not graphs:
            check(lib().vd_tess_set_option(self._h, _lib.VD_OPT_GRAPHS, 0))

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().vd_tess_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------------------------
    def set_generators(self, generators) -> None:
        a = np.ascontiguousarray(np.asarray(generators, dtype=np.float64).reshape(-1))
        n = a.size // 3
        rc = lib().vd_tess_set_generators(self._h, _ptr(a), n)
        self._host_errors()
        check(rc)

    def set_generators_device(self, ptr: int, n: int) -> None:
        """Points already in device memory (e.g. a torch CUDA tensor's data_ptr())."""
        rc = lib().vd_tess_set_generators_device(self._h, ctypes.c_void_p(int(ptr)), int(n))
        self._host_errors()
        check(rc)

    def read_generators(self, text: str) -> None:
        """read_generators (src/tessellation.rs:153-171): `id x y z` lines (parse_generators),
        then set_generators."""
        self.set_generators(parse_generators(text))

    def import_generators(self, path) -> None:
        """import_generators (src/tessellation.rs:126-148): read_generators of a UTF-8 text file;
        an unreadable file raises OSError (the reference returns its io::Error)."""
        with open(path, "r", encoding="utf-8", newline="") as fh:
            self.read_generators(fh.read())

    def random_generators(self, count: int, seed: Optional[int] = None) -> None:
        """random_generators (src/tessellation.rs:96-122): `count` points uniform in the bounding
        box that every wall contains (fewer if count * 1000 candidates do not yield them). Without
        a seed, one is drawn from the operating system, as the reference seeds from thread_rng;
        with one, the candidates are the SplitMix64 stream of that seed (vd_tess_random_generators)."""
        if seed is None:
            seed = secrets.randbits(64)
        rc = lib().vd_tess_random_generators(self._h, int(count), int(seed) & 0xFFFFFFFFFFFFFFFF)
        self._host_errors()
        check(rc)

    def set_generator(self, index: int, generator) -> None:
        """set_generator (src/tessellation.rs:71-92): moves one generator if every wall contains
        the new position; the bins follow update_generator (swap_remove, push). An index past the
        generators is ignored, as in the reference."""
        g = np.ascontiguousarray(np.asarray(generator, dtype=np.float64).reshape(3))
        rc = lib().vd_tess_set_generator(self._h, int(index), _ptr(g))
        self._host_errors()
        if rc != _lib.VD_ERANGE:
            check(rc)

    def add_wall(self, wall: Wall) -> None:
        g = wall.geometry
        if not isinstance(g, _GEOMETRIES):
            # any other geometry: contains and cut evaluated here, on the host, through callbacks
            cb = HostWallCallbacks(g)
            rc = lib().vd_tess_add_wall_planes_per_generator(self._h, wall.id(), ctypes.byref(cb.struct))
            cb.check_error()
            check(rc)
            self._host.append(cb)
        elif isinstance(g, PlaneGeometry):
            check(lib().vd_tess_add_wall_plane(self._h, wall.id(), _d3(g.point), _d3(g.normal_raw)))
        elif isinstance(g, SphereGeometry):
            check(lib().vd_tess_add_wall_sphere(self._h, wall.id(), _d3(g.center), g.radius))
        else:
            w = _vd_wall(wall)
            check(lib().vd_tess_add_wall(self._h, ctypes.byref(w)))
        self._walls.append(wall)

    def clear_walls(self) -> None:
        check(lib().vd_tess_clear_walls(self._h))
        self._walls = []
        self._host = []

    def _host_errors(self):
        for cb in self._host:
            cb.check_error()

    def calculate(self) -> None:
        check(lib().vd_tess_calculate(self._h))

    def relax(self) -> None:
        rc = lib().vd_tess_relax(self._h)
        self._host_errors()
        check(rc)

    def calculate_sealed(self) -> None:
        """calculate(), then the curved-wall sealing pass (vd_tess_calculate_sealed)."""
        check(lib().vd_tess_calculate_sealed(self._h))

    def prune_boundaries(self) -> None:
        """The boundary-vertex log of the current cells (vd_tess_prune_boundaries)."""
        check(lib().vd_tess_prune_boundaries(self._h))

    def seal_log(self) -> np.ndarray:
        """(cell, neighbour, wall id) rows, one per sealing clip that changed a cell."""
        n = ctypes.c_size_t(0)
        check(lib().vd_tess_get_seal_log(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.int32)
        check(lib().vd_tess_get_seal_log(self._h, _ptr(out), out.size, ctypes.byref(n)))
        return out.reshape(-1, 3)

    def prune_log(self):
        """(ints, positions): the flat prune log ((cell, neighbour, wall id, count) per entry) and
        the vertices of the entries (count rows of xyz each)."""
        c = np.zeros(2, dtype=np.uint64)
        check(lib().vd_tess_get_prune_log(self._h, None, 0, None, 0, _ptr(c)))
        ints = np.zeros(int(c[0]), dtype=np.int32)
        pos = np.zeros(int(c[1]), dtype=np.float64)
        check(lib().vd_tess_get_prune_log(self._h, _ptr(ints), ints.size, _ptr(pos), pos.size, _ptr(c)))
        return ints, pos.reshape(-1, 3)

    def set_keep_cells(self, keep: bool) -> None:
        check(lib().vd_tess_set_option(self._h, _lib.VD_OPT_KEEP_CELLS, 1 if keep else 0))

    def set_graphs(self, on: bool) -> None:
        """Replay repeated first-round launch sequences from a captured hipGraph (VD_OPT_GRAPHS)."""
        check(lib().vd_tess_set_option(self._h, _lib.VD_OPT_GRAPHS, 1 if on else 0))

    def set_arena(self, vertices_per_cell: int) -> None:
        """Initial full-cell arena per cell (VD_OPT_ARENA); a calculate that fills it grows it."""
        check(lib().vd_tess_set_option(self._h, _lib.VD_OPT_ARENA, int(vertices_per_cell)))

    def set_far_list(self, min_entries: int) -> None:
        """Walk length (shell entries) from which cells reaching across empty space rerun over a
        list of the occupied entries only (VD_OPT_FAR_LIST; 0: never). Results are identical."""
        check(lib().vd_tess_set_option(self._h, _lib.VD_OPT_FAR_LIST, int(min_entries)))

    def set_attempts(self, attempts: int) -> None:
        """Launches, each with twice the arena of the one before, after which calculate() gives up
        with VD_ECAPACITY and keeps no results (VD_OPT_ATTEMPTS, default 8)."""
        check(lib().vd_tess_set_option(self._h, _lib.VD_OPT_ATTEMPTS, int(attempts)))

    def set_launch(self, t0_waves: int | None = None, queue_grain: int | None = None,
                   first_tier: int | None = None, first_probe: int | None = None) -> None:
        """Launch shape of the cell kernels (VD_OPT_T0_WAVES 5/6/7/8, VD_OPT_QUEUE_GRAIN 1-64,
        VD_OPT_FIRST_TIER 0/1, VD_OPT_FIRST_PROBE 0/1). Results are identical under every shape."""
        for key, v in ((_lib.VD_OPT_T0_WAVES, t0_waves), (_lib.VD_OPT_QUEUE_GRAIN, queue_grain),
                       (_lib.VD_OPT_FIRST_TIER, first_tier), (_lib.VD_OPT_FIRST_PROBE, first_probe)):
            if v is not None:
                check(lib().vd_tess_set_option(self._h, key, int(v)))

    def autotune_launch(self) -> dict:
        """Times calculate() over the launch shapes and keeps the fastest (vd_tess_autotune_launch);
        returns the picks. Generators must be set."""
        picks = (ctypes.c_int * 4)()
        check(lib().vd_tess_autotune_launch(self._h, picks))
        return {"t0_waves": picks[0], "queue_grain": picks[1], "first_tier": picks[2], "first_probe": picks[3]}

    # -- queries -------------------------------------------------------------------------------
    def count_generators(self) -> int:
        return int(lib().vd_tess_count_generators(self._h))

    def count_cells(self) -> int:
        return int(lib().vd_tess_count_cells(self._h))

    def generators(self) -> np.ndarray:
        n = self.count_generators()
        out = np.empty(3 * n, dtype=np.float64)
        check(lib().vd_tess_get_generators(self._h, _ptr(out), n))
        return out

    def get_generator(self, index: int):
        return self.generators()[3 * index:3 * index + 3]

    def volumes(self) -> np.ndarray:
        n = self.count_cells()
        out = np.empty(n, dtype=np.float64)
        check(lib().vd_tess_get_volumes(self._h, _ptr(out), n))
        return out

    def centroids(self) -> np.ndarray:
        n = self.count_cells()
        out = np.empty((n, 3), dtype=np.float64)
        check(lib().vd_tess_get_centroids(self._h, _ptr(out), n))
        return out

    def face_counts(self) -> np.ndarray:
        """Number of faces per cell."""
        n = self.count_cells()
        out = np.empty(n, dtype=np.uint32)
        check(lib().vd_tess_get_face_counts(self._h, _ptr(out), n))
        return out

    def neighbors(self):
        """(offsets[n+1], ids) CSR of every cell's face_neighbors in face order."""
        n = self.count_cells()
        off = np.empty(n + 1, dtype=np.uint64)
        check(lib().vd_tess_get_neighbors(self._h, _ptr(off), n, None, 0))
        ids = np.empty(int(off[-1]), dtype=np.int32)
        check(lib().vd_tess_get_neighbors(self._h, _ptr(off), n, _ptr(ids), ids.size))
        return off, ids

    def get_cell(self, index: int):
        if index < 0 or index >= self.count_cells():
            return None
        sizes = np.zeros(3, dtype=np.uint64)
        check(lib().vd_tess_get_cell(self._h, int(index), _ptr(sizes), None, None, None, None))
        nv, nf, ni = (int(x) for x in sizes)
        v = np.empty(3 * nv, dtype=np.float64)
        fc = np.empty(nf, dtype=np.uint32)
        fi = np.empty(ni, dtype=np.uint32)
        fn = np.empty(nf, dtype=np.int32)
        check(lib().vd_tess_get_cell(self._h, int(index), _ptr(sizes), _ptr(v), _ptr(fc), _ptr(fi), _ptr(fn)))
        return Cell3DFaces(int(index), v, fc, fi, fn)

    def cells(self):
        return [self.get_cell(i) for i in range(self.count_cells())]

    def map(self, f):
        """Tessellation::map: compute every cell and apply f to it."""
        self.calculate()
        return [f(self.get_cell(i)) for i in range(self.count_cells())]

    def stats(self) -> np.ndarray:
        out = np.zeros(10, dtype=np.float64)
        check(lib().vd_tess_get_stats(self._h, _ptr(out)))
        return out


def shell_prefix(bounds: BoundingBox, nx: int, ny: int, nz: int, kmax: int, device: bool = False):
    """Stable shell-order prefix (vd_shell_prefix in include/vd.h), built on the host, or with
    ``device=True`` on the GPU the way calculate() builds it (vd_shell_prefix_device)."""
    fn = lib().vd_shell_prefix_device if device else lib().vd_shell_prefix
    count = ctypes.c_size_t(0)
    nxt = ctypes.c_double(0.0)
    check(fn(_d3(bounds.min), _d3(bounds.max), nx, ny, nz, kmax, None, None, 0,
             ctypes.byref(count), ctypes.byref(nxt)))
    m = count.value
    dxyz = np.empty((m, 3), dtype=np.int32)
    keys = np.empty(m, dtype=np.float64)
    check(fn(_d3(bounds.min), _d3(bounds.max), nx, ny, nz, kmax, _ptr(dxyz),
             _ptr(keys), m, ctypes.byref(count), ctypes.byref(nxt)))
    return dxyz, keys, nxt.value


def grid_for_points(bounds: BoundingBox, points, per_bin: float):
    """The grid rule for uneven input (vd_grid_for_points): per_bin generators per bin as seen by
    the average generator."""
    a = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1))
    out = np.zeros(3, dtype=np.uint32)
    check(lib().vd_grid_for_points(_d3(bounds.min), _d3(bounds.max), _ptr(a), a.size // 3, float(per_bin), _ptr(out)))
    return tuple(int(x) for x in out)


def grid_for(bounds: BoundingBox, n: int, per_bin: float):
    """The grid rule of the batch autotuner (vd_grid_for)."""
    out = np.zeros(3, dtype=np.uint32)
    check(lib().vd_grid_for(_d3(bounds.min), _d3(bounds.max), int(n), float(per_bin), _ptr(out)))
    return tuple(int(x) for x in out)


def _vd_wall(wall: Wall):
    """struct vd_wall of a Wall; its point arrays stay owned by the geometry object."""
    w = _lib.VdWall()
    w.id = wall.id()
    w.kind = wall.geometry.kind
    wall.geometry._fill(w)
    return w


def calculate_static(bounds: BoundingBox, grid, walls, d_points: int, n: int, d_vol: int, d_cen: int,
                     d_nface: int, stream: int = 0) -> int:
    """vd_calculate_static: one tessellation (walls in order, then the n points at device address
    d_points) computed without device allocation, in the library's static scratch array; volumes,
    centroids and face counts go to the caller's device buffers d_vol, d_cen and d_nface. Returns
    the number of cells kept by the walls."""
    ws = (_lib.VdWall * max(1, len(walls)))()
    for k, w in enumerate(walls):
        ws[k] = _vd_wall(w)
    kept = ctypes.c_size_t(0)
    nx, ny, nz = (int(g) for g in grid)
    check(lib().vd_calculate_static(_d3(bounds.min), _d3(bounds.max), nx, ny, nz, ws, len(walls),
                                    ctypes.c_void_p(d_points), int(n), ctypes.c_void_p(d_vol),
                                    ctypes.c_void_p(d_cen), ctypes.c_void_p(d_nface), ctypes.byref(kept),
                                    ctypes.c_void_p(stream) if stream else None))
    return int(kept.value)


class Batch:
    """Many independent tessellations computed in one persistent launch (vd_batch_* in vd.h).

    ``add(bounds, grid, walls, points)`` adds one ``Tessellation::<3, Cell3DFaces,
    Algorithm3DGrid>`` (grid ``None`` lets the batch's generators-per-bin rule pick it);
    ``calculate()`` runs every member; results are read per member.
    """

    def __init__(self, keep_cells: bool = False, per_bin: float = 0.35, graphs: bool = True,
                 far_list: int | None = None):
        h = ctypes.c_void_p()
        check(lib().vd_batch_create(ctypes.byref(h)))
        self._h = h
        if keep_cells:
            check(lib().vd_batch_set_option(self._h, _lib.VD_BOPT_KEEP_CELLS, 1))
        if not graphs:
            check(lib().vd_batch_set_option(self._h, _lib.VD_BOPT_GRAPHS, 0))
        check(lib().vd_batch_set_option(self._h, _lib.VD_BOPT_PER_BIN, int(round(per_bin * 1000))))
        if far_list is not None:   # VD_BOPT_FAR_LIST: as Tessellation.set_far_list
            check(lib().vd_batch_set_option(self._h, _lib.VD_BOPT_FAR_LIST, int(far_list)))

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().vd_batch_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, bounds: BoundingBox, grid, walls, points) -> int:
        a = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1))
        nx, ny, nz = (0, 0, 0) if grid is None else (int(g) for g in grid)
        ws = (_lib.VdWall * max(1, len(walls)))()
        for k, w in enumerate(walls):
            ws[k] = _vd_wall(w)
        idx = ctypes.c_uint32(0)
        check(lib().vd_batch_add(self._h, _d3(bounds.min), _d3(bounds.max), nx, ny, nz, ws, len(walls),
                                 _ptr(a), a.size // 3, ctypes.byref(idx)))
        return int(idx.value)

    def upload(self):
        check(lib().vd_batch_upload(self._h))

    def calculate(self):
        check(lib().vd_batch_calculate(self._h))

    def relax(self):
        check(lib().vd_batch_relax(self._h))

    def autotune(self, candidates):
        c = np.ascontiguousarray(np.asarray(candidates, dtype=np.float64))
        best = ctypes.c_double(0.0)
        check(lib().vd_batch_autotune(self._h, _ptr(c), c.size, ctypes.byref(best)))
        return best.value

    def set_launch(self, t0_waves: int | None = None, queue_grain: int | None = None,
                   first_tier: int | None = None, first_probe: int | None = None) -> None:
        """As Tessellation.set_launch (VD_BOPT_T0_WAVES, VD_BOPT_QUEUE_GRAIN, VD_BOPT_FIRST_TIER,
        VD_BOPT_FIRST_PROBE)."""
        for key, v in ((_lib.VD_BOPT_T0_WAVES, t0_waves), (_lib.VD_BOPT_QUEUE_GRAIN, queue_grain),
                       (_lib.VD_BOPT_FIRST_TIER, first_tier), (_lib.VD_BOPT_FIRST_PROBE, first_probe)):
            if v is not None:
                check(lib().vd_batch_set_option(self._h, key, int(v)))

    def autotune_launch(self) -> dict:
        """vd_batch_autotune_launch: the launch shape for the batch's current grids."""
        picks = (ctypes.c_int * 4)()
        check(lib().vd_batch_autotune_launch(self._h, picks))
        return {"t0_waves": picks[0], "queue_grain": picks[1], "first_tier": picks[2], "first_probe": picks[3]}

    def __len__(self):
        return int(lib().vd_batch_count(self._h))

    def info(self, t: int):
        out = np.zeros(5, dtype=np.uint64)
        check(lib().vd_batch_member_info(self._h, int(t), _ptr(out)))
        return {"cells": int(out[0]), "grid": (int(out[1]), int(out[2]), int(out[3])), "offset": int(out[4])}

    def volumes(self, t: int) -> np.ndarray:
        n = self.info(t)["cells"]
        out = np.empty(n, dtype=np.float64)
        check(lib().vd_batch_get_volumes(self._h, int(t), _ptr(out), n))
        return out

    def centroids(self, t: int) -> np.ndarray:
        n = self.info(t)["cells"]
        out = np.empty((n, 3), dtype=np.float64)
        check(lib().vd_batch_get_centroids(self._h, int(t), _ptr(out), n))
        return out

    def face_counts(self, t: int) -> np.ndarray:
        n = self.info(t)["cells"]
        out = np.empty(n, dtype=np.uint32)
        check(lib().vd_batch_get_face_counts(self._h, int(t), _ptr(out), n))
        return out

    def generators(self, t: int) -> np.ndarray:
        n = self.info(t)["cells"]
        out = np.empty(3 * n, dtype=np.float64)
        check(lib().vd_batch_get_generators(self._h, int(t), _ptr(out), n))
        return out

    def neighbors(self, t: int):
        n = self.info(t)["cells"]
        off = np.empty(n + 1, dtype=np.uint64)
        check(lib().vd_batch_get_neighbors(self._h, int(t), _ptr(off), n, None, 0))
        ids = np.empty(int(off[-1]), dtype=np.int32)
        check(lib().vd_batch_get_neighbors(self._h, int(t), _ptr(off), n, _ptr(ids), ids.size))
        return off, ids

    def get_cell(self, t: int, i: int):
        sizes = np.zeros(3, dtype=np.uint64)
        rc = lib().vd_batch_get_cell(self._h, int(t), int(i), _ptr(sizes), None, None, None, None)
        if rc == _lib.VD_ERANGE:
            return None
        check(rc)
        nv, nf, ni = (int(x) for x in sizes)
        v = np.empty(3 * nv, dtype=np.float64)
        fc = np.empty(nf, dtype=np.uint32)
        fi = np.empty(ni, dtype=np.uint32)
        fn = np.empty(nf, dtype=np.int32)
        check(lib().vd_batch_get_cell(self._h, int(t), int(i), _ptr(sizes), _ptr(v), _ptr(fc), _ptr(fi), _ptr(fn)))
        return Cell3DFaces(int(i), v, fc, fi, fn)

    def stats(self) -> np.ndarray:
        out = np.zeros(10, dtype=np.float64)
        check(lib().vd_batch_get_stats(self._h, _ptr(out)))
        return out


class Server:
    """Continuous-batching server (vd_server_* in vd.h): tessellation and Lloyd jobs admitted by
    HBM budget, each of two lanes (small jobs, and large ones on a worker thread) running one
    batched launch over its admitted jobs per step, Lloyd jobs relaxed on the device between their
    iterations."""

    QUEUED, RUNNING, FINISHED, REFUSED = 0, 1, 2, 3

    def __init__(self, hbm_budget: int, max_slots: int = 64, per_bin: float | None = None,
                 small_job: int | None = None):
        h = ctypes.c_void_p()
        check(lib().vd_server_create(int(hbm_budget), int(max_slots), ctypes.byref(h)))
        self._h = h
        if per_bin is not None:
            check(lib().vd_server_set_option(h, _lib.VD_SOPT_PER_BIN, int(round(per_bin * 1000))))
        if small_job is not None:
            check(lib().vd_server_set_option(h, _lib.VD_SOPT_SMALL_JOB, int(small_job)))

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            lib().vd_server_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, bounds: BoundingBox, grid, walls, points, iterations: int = 1) -> int:
        """Queue a job; raises VdError (VD_ENOMEM) for a job larger than the whole budget."""
        a = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1))
        nx, ny, nz = (0, 0, 0) if grid is None else (int(g) for g in grid)
        ws = (_lib.VdWall * max(1, len(walls)))()
        for k, w in enumerate(walls):
            ws[k] = _vd_wall(w)
        job = ctypes.c_uint64(0)
        check(lib().vd_server_submit(self._h, _d3(bounds.min), _d3(bounds.max), nx, ny, nz, ws, len(walls),
                                     _ptr(a), a.size // 3, int(iterations), ctypes.byref(job)))
        return int(job.value)

    def step(self) -> int:
        fin = ctypes.c_uint32(0)
        check(lib().vd_server_step(self._h, ctypes.byref(fin)))
        return int(fin.value)

    def pending(self):
        c = np.zeros(3, dtype=np.uint64)
        check(lib().vd_server_pending(self._h, _ptr(c)))
        return {"queued": int(c[0]), "running": int(c[1]), "bytes": int(c[2])}

    def info(self, job: int):
        st = ctypes.c_int32(0)
        info = np.zeros(3, dtype=np.uint64)
        times = np.zeros(3, dtype=np.float64)
        check(lib().vd_server_job_info(self._h, int(job), ctypes.byref(st), _ptr(info), _ptr(times)))
        return {"state": int(st.value), "cells": int(info[0]), "runs": int(info[1]),
                "total_cells": int(info[2]), "submit": float(times[0]), "admit": float(times[1]),
                "finish": float(times[2])}

    def grid(self, job: int):
        """The grid the job runs on (vd_server_job_grid)."""
        out = np.zeros(3, dtype=np.uint32)
        check(lib().vd_server_job_grid(self._h, int(job), _ptr(out)))
        return tuple(int(x) for x in out)

    def result(self, job: int):
        n = self.info(job)["cells"]
        vol = np.empty(n, dtype=np.float64)
        cen = np.empty((n, 3), dtype=np.float64)
        gen = np.empty((n, 3), dtype=np.float64)
        check(lib().vd_server_result(self._h, int(job), _ptr(vol), _ptr(cen), _ptr(gen), n))
        return vol, cen, gen

    def release(self, job: int):
        check(lib().vd_server_release(self._h, int(job)))

    def stats(self):
        out = np.zeros(4, dtype=np.float64)
        check(lib().vd_server_stats(self._h, _ptr(out)))
        return {"steps": int(out[0]), "cells": int(out[1]), "bytes": int(out[2]), "budget": int(out[3])}

    def memory(self):
        """Device memory the server holds now, the most it held at once, its fixed reserve and its
        budget, in bytes (vd_server_memory)."""
        out = np.zeros(4, dtype=np.uint64)
        check(lib().vd_server_memory(self._h, _ptr(out)))
        return {"held": int(out[0]), "peak": int(out[1]), "reserve": int(out[2]), "budget": int(out[3])}
