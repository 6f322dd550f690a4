"""Wall geometries written in Python, the way a user of the reference implements WallGeometry
(src/wall/mod.rs:53-71) for a shape the crate does not ship: contains(p), cut(g) -> [(point,
normal)] with normals out of the valid region, and is_planar(). Python floats are IEEE doubles
evaluated left to right without contraction, and math.sqrt is correctly rounded, so a restatement
of a built-in geometry's expressions gives the built-in's bits."""
import math


class PySphere:
    """SphereGeometry (wall_3d.rs:66-100), expression for expression."""

    def __init__(self, center, radius):
        self.c = tuple(float(x) for x in center)
        self.r = float(radius)

    def contains(self, p):
        dx, dy, dz = p[0] - self.c[0], p[1] - self.c[1], p[2] - self.c[2]
        return dx * dx + dy * dy + dz * dz <= self.r * self.r

    def cut(self, g):
        dx, dy, dz = g[0] - self.c[0], g[1] - self.c[1], g[2] - self.c[2]
        dist = math.sqrt(dx * dx + dy * dy + dz * dz)
        if dist == 0.0:
            return []
        sc = self.r / dist
        return [((self.c[0] + dx * sc, self.c[1] + dy * sc, self.c[2] + dz * sc),
                 (dx / dist, dy / dist, dz / dist))]

    def is_planar(self):
        return False


def octahedron_planes(center, radius):
    """The 8 facets of |x - c|_1 <= radius as (point, unit outward normal)."""
    k = 1.0 / math.sqrt(3.0)
    out = []
    for sx in (1.0, -1.0):
        for sy in (1.0, -1.0):
            for sz in (1.0, -1.0):
                out.append(((center[0] + radius * sx, center[1], center[2]), (sx * k, sy * k, sz * k)))
    return out


class PyOctahedron:
    """A convex polyhedron (planar: cut emits every facet, contains tests them in order), the
    expressions of ConvexPolyhedronGeometry (wall_3d.rs:476-508) on the octahedron's facets."""

    def __init__(self, center, radius):
        self.planes = octahedron_planes(center, radius)

    def contains(self, p):
        for q, n in self.planes:
            if (p[0] - q[0]) * n[0] + (p[1] - q[1]) * n[1] + (p[2] - q[2]) * n[2] > 0.0:
                return False
        return True

    def cut(self, g):
        return list(self.planes)

    def is_planar(self):
        return True


class PyNearFacets:
    """A curved-wall stand-in whose cut emits a number of planes that depends on the generator:
    the octahedron's facets that face it (signed distance above -radius/2), farthest first. Not a
    shape of the reference; it exercises variable plane counts per generator (1 to 8)."""

    def __init__(self, center, radius):
        self.c = tuple(float(x) for x in center)
        self.r = float(radius)
        self.planes = octahedron_planes(center, radius)

    def contains(self, p):
        return abs(p[0] - self.c[0]) + abs(p[1] - self.c[1]) + abs(p[2] - self.c[2]) <= self.r

    def cut(self, g):
        sd = []
        for k, (q, n) in enumerate(self.planes):
            d = (g[0] - q[0]) * n[0] + (g[1] - q[1]) * n[1] + (g[2] - q[2]) * n[2]
            if d > -0.5 * self.r:
                sd.append((-d, k))
        return [self.planes[k] for _, k in sorted(sd)]

    def is_planar(self):
        return False
