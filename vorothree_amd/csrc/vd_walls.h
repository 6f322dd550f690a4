// Wall geometry on the device: `contains` (the wall filter of set_generators) and `cut` (the
// planes a generator's cell is clipped by), restating src/wall/wall_3d.rs expression for
// expression, in the same evaluation order (no contraction). powi(2) is a plain square.
#pragma once

#include <cfloat>

#include "vd_internal.h"

namespace vd {

enum : int32_t {
    kWallPlane = 0,
    kWallSphere = 1,
    kWallCylinder = 2,
    kWallCone = 3,
    kWallTorus = 4,
    kWallPolyhedron = 5,
    kWallTube = 6,
    kWallHost = 7,   // the caller's geometry: contains and cut evaluated on the host
};

// WallGeometry::is_planar (wall_3d.rs:44, 709; mod.rs:66-70 for the caller's geometry, whose flag
// is in `closed`).
__host__ __device__ inline bool wall_planar(const WallDev &w) {
    return w.kind == kWallPlane || w.kind == kWallPolyhedron || (w.kind == kWallHost && w.closed);
}

// Segments of a tube's sample polyline: n in a closed loop, n - 1 open (wall_3d.rs:412, 769, 915).
__host__ __device__ inline uint32_t tube_segments(const WallDev &w) {
    if (w.npts == 0) return 0u;
    return w.closed ? w.npts : w.npts - 1u;
}

// Projection of q on segment i (samples i and (i + 1) % n) and its squared distance
// (get_closest_point, wall_3d.rs:415-436, 771-795, 917-941).
__host__ __device__ inline double tube_segment(const WallDev &w, uint32_t i, double qx, double qy, double qz,
                                               double &px, double &py, double &pz) {
    const uint32_t j = (i + 1u) % w.npts;
    const double *p0 = w.pts + 3 * (size_t)i, *p1 = w.pts + 3 * (size_t)j;
    const double v0 = p1[0] - p0[0], v1 = p1[1] - p0[1], v2 = p1[2] - p0[2];
    const double w0 = qx - p0[0], w1 = qy - p0[1], w2 = qz - p0[2];
    const double c1 = w0 * v0 + w1 * v1 + w2 * v2;
    const double c2 = v0 * v0 + v1 * v1 + v2 * v2;
    double t = 0.0;
    if (!(c2 <= 0.0)) {
        t = c1 / c2;
        if (t < 0.0) t = 0.0;   // f64::clamp(0, 1)
        if (t > 1.0) t = 1.0;
    }
    px = p0[0] + v0 * t;
    py = p0[1] + v1 * t;
    pz = p0[2] + v2 * t;
    const double dx = qx - px, dy = qy - py, dz = qz - pz;
    return dx * dx + dy * dy + dz * dz;
}

// get_closest_point, one thread: the first segment with the smallest distance (strict <).
__host__ __device__ inline void tube_closest(const WallDev &w, double qx, double qy, double qz, double &cx,
                                             double &cy, double &cz) {
    cx = cy = cz = 0.0;
    if (w.npts == 0) return;
    cx = w.pts[0]; cy = w.pts[1]; cz = w.pts[2];
    double best = DBL_MAX;
    const uint32_t lim = tube_segments(w);
    for (uint32_t i = 0; i < lim; ++i) {
        double px, py, pz;
        const double d2 = tube_segment(w, i, qx, qy, qz, px, py, pz);
        if (d2 < best) { best = d2; cx = px; cy = py; cz = pz; }
    }
}

// WallGeometry::contains (wall_3d.rs:31-36, 73-78, 132-143, 202-214, 293-309, 442-446, 691-701,
// 801-805, 947-951). Also evaluated on the host (set_generator), with pts in host memory.
__host__ __device__ inline bool wall_contains(const WallDev &w, double x, double y, double z) {
    const double dx = x - w.a[0], dy = y - w.a[1], dz = z - w.a[2];
    switch (w.kind) {
    case kWallPlane:
        return (dx * w.b[0] + dy * w.b[1] + dz * w.b[2]) >= 0.0;
    case kWallSphere:
        return (dx * dx + dy * dy + dz * dz) <= w.r * w.r;
    case kWallCylinder: {
        const double dot = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
        const double px = dx - dot * w.b[0], py = dy - dot * w.b[1], pz = dz - dot * w.b[2];
        return (px * px + py * py + pz * pz) <= w.r * w.r;
    }
    case kWallCone: {
        const double h = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
        const double px = dx - h * w.b[0], py = dy - h * w.b[1], pz = dz - h * w.b[2];
        const double r = sqrt(px * px + py * py + pz * pz);
        return h >= 0.0 && r <= h * w.tn;
    }
    case kWallTorus: {
        const double dot = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
        const double px = dx - dot * w.b[0], py = dy - dot * w.b[1], pz = dz - dot * w.b[2];
        const double dist_perp = sqrt(px * px + py * py + pz * pz);
        const double t = dist_perp - w.r;
        const double dist_tube = sqrt(t * t + dot * dot);
        return dist_tube <= w.r2;
    }
    case kWallPolyhedron:
        for (uint32_t k = 0; k < w.npts; ++k) {
            const double *p = w.pts + 6 * (size_t)k;
            const double ex = x - p[0], ey = y - p[1], ez = z - p[2];
            if (ex * p[3] + ey * p[4] + ez * p[5] > 0.0) return false;
        }
        return true;
    case kWallTube: {
        double cx, cy, cz;
        tube_closest(w, x, y, z, cx, cy, cz);
        const double ex = x - cx, ey = y - cy, ez = z - cz;
        return ex * ex + ey * ey + ez * ez <= w.r * w.r;
    }
    default:
        return true;
    }
}

// A cut plane: a point on it and the normal pointing out of the valid region. Named fields rather
// than arrays keep it in registers where it is filled branch by branch.
struct WallPlane { double px, py, pz, nx, ny, nz; };

// Plane j of WallGeometry::cut for generator g (point p, normal n pointing out of the valid
// region); false when the wall emits no plane j. Every kind but the polyhedron emits at most one
// plane; the tube's plane is built by the caller from its closest point (tube_plane).
// wall_3d.rs:38-42, 80-100, 145-170, 216-258, 311-363, 703-707.
__device__ inline bool wall_plane(const WallDev &w, uint32_t j, double gx, double gy, double gz,
                                  WallPlane &out) {
    double qpx, qpy, qpz, qnx, qny, qnz;   // set case by case, written to out once
    if (w.kind == kWallHost) return false;   // its planes come from the host, per generator
    if (w.kind == kWallPolyhedron) {
        if (j >= w.npts) return false;
        const double *q = w.pts + 6 * (size_t)j;
        qpx = q[0]; qpy = q[1]; qpz = q[2];
        qnx = q[3]; qny = q[4]; qnz = q[5];
    } else {
        if (j > 0) return false;
        const double dx = gx - w.a[0], dy = gy - w.a[1], dz = gz - w.a[2];
        switch (w.kind) {
        case kWallPlane:
            qpx = w.a[0]; qpy = w.a[1]; qpz = w.a[2];
            qnx = -w.b[0]; qny = -w.b[1]; qnz = -w.b[2];
            break;
        case kWallSphere: {
            const double dist = sqrt(dx * dx + dy * dy + dz * dz);
            if (dist == 0.0) return false;
            const double sc = w.r / dist;
            qpx = w.a[0] + dx * sc; qpy = w.a[1] + dy * sc; qpz = w.a[2] + dz * sc;
            qnx = dx / dist; qny = dy / dist; qnz = dz / dist;
            break;
        }
        case kWallCylinder: {
            const double dot = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
            const double px = dx - dot * w.b[0], py = dy - dot * w.b[1], pz = dz - dot * w.b[2];
            const double dist = sqrt(px * px + py * py + pz * pz);
            if (dist == 0.0) return false;
            const double sc = w.r / dist;
            qpx = w.a[0] + dot * w.b[0] + px * sc;
            qpy = w.a[1] + dot * w.b[1] + py * sc;
            qpz = w.a[2] + dot * w.b[2] + pz * sc;
            qnx = px / dist; qny = py / dist; qnz = pz / dist;
            break;
        }
        case kWallCone: {
            const double h = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
            const double px = dx - h * w.b[0], py = dy - h * w.b[1], pz = dz - h * w.b[2];
            const double r = sqrt(px * px + py * py + pz * pz);
            if (r == 0.0) return false;
            const double rx = px / r, ry = py / r, rz = pz / r;
            const double dist = r * w.cs - h * w.sn;
            const double p2r = r - dist * w.cs;
            const double p2h = h + dist * w.sn;
            if (p2h < 0.0) {
                const double dt = sqrt(dx * dx + dy * dy + dz * dz);
                if (dt == 0.0) return false;
                qpx = w.a[0]; qpy = w.a[1]; qpz = w.a[2];
                qnx = dx / dt; qny = dy / dt; qnz = dz / dt;
                break;
            }
            qpx = w.a[0] + p2h * w.b[0] + p2r * rx;
            qpy = w.a[1] + p2h * w.b[1] + p2r * ry;
            qpz = w.a[2] + p2h * w.b[2] + p2r * rz;
            qnx = w.cs * rx - w.sn * w.b[0];
            qny = w.cs * ry - w.sn * w.b[1];
            qnz = w.cs * rz - w.sn * w.b[2];
            break;
        }
        case kWallTorus: {
            const double dot = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
            const double px = dx - dot * w.b[0], py = dy - dot * w.b[1], pz = dz - dot * w.b[2];
            const double dist_perp = sqrt(px * px + py * py + pz * pz);
            double ux, uy, uz;
            if (dist_perp < 1e-9) {
                // on the axis: any direction perpendicular to it
                double tx = 1.0, ty = 0.0;
                const double tz = 0.0;
                if (fabs(w.b[0]) > 0.9) { tx = 0.0; ty = 1.0; }
                const double td = tx * w.b[0] + ty * w.b[1] + tz * w.b[2];
                const double ax = tx - td * w.b[0], ay = ty - td * w.b[1], az = tz - td * w.b[2];
                const double len = sqrt(ax * ax + ay * ay + az * az);
                if (len == 0.0) return false;
                ux = ax / len; uy = ay / len; uz = az / len;
            } else {
                ux = px / dist_perp; uy = py / dist_perp; uz = pz / dist_perp;
            }
            const double cx = w.a[0] + ux * w.r, cy = w.a[1] + uy * w.r, cz = w.a[2] + uz * w.r;
            const double vx = gx - cx, vy = gy - cy, vz = gz - cz;
            const double dc = sqrt(vx * vx + vy * vy + vz * vz);
            if (dc == 0.0) return false;
            qnx = vx / dc; qny = vy / dc; qnz = vz / dc;
            qpx = cx + qnx * w.r2; qpy = cy + qny * w.r2; qpz = cz + qnz * w.r2;
            break;
        }
        default:
            return false;
        }
    }
    out = WallPlane{qpx, qpy, qpz, qnx, qny, qnz};
    return true;
}

// The tube's cut plane from the closest point c of generator g (wall_3d.rs:448-456, 807-815,
// 953-961); false when g lies on the curve.
__device__ inline bool tube_plane(const WallDev &w, double gx, double gy, double gz, double cx, double cy,
                                  double cz, double p[3], double n[3]) {
    const double ex = gx - cx, ey = gy - cy, ez = gz - cz;
    const double dist = sqrt(ex * ex + ey * ey + ez * ez);
    if (dist == 0.0) return false;
    n[0] = ex / dist; n[1] = ey / dist; n[2] = ez / dist;
    p[0] = cx + n[0] * w.r; p[1] = cy + n[1] * w.r; p[2] = cz + n[2] * w.r;
    return true;
}

}  // namespace vd
