// CPU restatement of voronoid's Tessellation<3, Cell3DFaces, Algorithm3DGrid>::calculate() path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP library is compared against. It
// is loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by nothing
// that ships. It restates the reference algorithm step by step; every function cites the reference
// file:line it follows (paths relative to the mdt-re/vorothree repository root).
//
// Numerics policy: compiled with -ffp-contract=off and without -ffast-math, so every expression is
// evaluated as written, left to right, with IEEE-correct division and sqrt -- the same semantics as
// the Rust reference, which never contracts a*b+c.
//
// One documented deviation from the reference: the bin shell order is sorted STABLY by its
// min-distance key (ties keep the z-outer / y / x-inner generation order). The reference uses
// sort_unstable_by (src/algorithm/algo_3d_grid.rs:76), whose tie order is that of Rust's ipnsort.
// The order is built as a prefix that grows as the walks need it (the whole table of a 474^3 grid
// would hold 8.6e8 entries); walks over it visit exactly the entries a walk over the whole table
// visits.
//
// A second one, in random_generators: the reference draws from a StdRng seeded by thread_rng, so
// no run of it can be reproduced; the candidates here are the SplitMix64 stream of a given seed.
//
// Pinned by: the Voro++ fixtures (vertex counts and positions, face counts), closed forms (box,
// lattices, plane-clipped boxes), a brute-force all-pairs clipping on small inputs, topology
// invariants (reciprocity, Euler characteristic), the grid unit-test values of the reference, the
// bin orders Vec::swap_remove and Vec::push fix for set_generator, and SplitMix64's published
// outputs with a direct rejection check for random_generators. See tests/test_oracle.py.

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

namespace {

// ---------------------------------------------------------------------------------------------
// Polyhedral cell. Restates src/cell/cell_3d_faces.rs:23-74 (layout and box initialisation).
// ---------------------------------------------------------------------------------------------
struct Poly {
    std::vector<double> v;        // xyz per vertex
    std::vector<uint8_t> fcnt;    // vertex count per face (u8, as the reference)
    std::vector<uint16_t> fidx;   // flattened face vertex indices (u16, as the reference)
    std::vector<int32_t> fnb;     // neighbour id per face; negative = box side or wall
    size_t nv() const { return v.size() / 3; }
    bool empty() const { return v.empty(); }
};

// Algorithmic-work counters (flop accounting used as the roofline unit). Rules: each + - * / sqrt
// counts 1; compares, clamps and max() count 0.
struct Work {
    uint64_t flops = 0;
    uint64_t candidates = 0;   // visitor calls (tessellation.rs:446)
    uint64_t bins = 0;         // in-range bins whose distance was evaluated (algo_3d_grid.rs:182)
    uint64_t clips = 0;        // clip calls after the distance test
    uint64_t cuts = 0;         // clips that modified the cell
    uint64_t entries = 0;      // shell entries examined
    uint64_t maxv = 0, maxs = 0, maxf = 0;   // largest polytope met while the cell was built
    void add(const Work &o) {
        flops += o.flops; candidates += o.candidates; bins += o.bins;
        clips += o.clips; cuts += o.cuts; entries += o.entries;
    }
    void size(const Poly &c) {
        maxv = std::max<uint64_t>(maxv, c.nv());
        maxs = std::max<uint64_t>(maxs, c.fidx.size());
        maxf = std::max<uint64_t>(maxf, c.fcnt.size());
    }
};

inline int32_t box_side(int axis, bool is_max) {  // src/bounds/mod.rs:21-23
    return -1 - (axis * 2 + (is_max ? 1 : 0));
}

void poly_box(Poly &c, const double lo[3], const double hi[3]) {  // cell_3d_faces.rs:37-74
    const double xs[2] = {lo[0], hi[0]}, ys[2] = {lo[1], hi[1]}, zs[2] = {lo[2], hi[2]};
    // vertex k: x from bit pattern of the reference's listing (0:(lo,lo,lo) 1:(hi,lo,lo)
    // 2:(hi,hi,lo) 3:(lo,hi,lo) 4..7 the same at z=hi)
    static const int xb[8] = {0, 1, 1, 0, 0, 1, 1, 0};
    static const int yb[8] = {0, 0, 1, 1, 0, 0, 1, 1};
    c.v.resize(24);
    for (int k = 0; k < 8; ++k) {
        c.v[3 * k + 0] = xs[xb[k]];
        c.v[3 * k + 1] = ys[yb[k]];
        c.v[3 * k + 2] = zs[k >> 2];
    }
    c.fcnt.assign(6, 4);
    static const uint16_t faces[24] = {3, 2, 1, 0, 4, 5, 6, 7, 0, 1, 5, 4,
                                       2, 3, 7, 6, 0, 4, 7, 3, 1, 2, 6, 5};
    c.fidx.assign(faces, faces + 24);
    c.fnb = {box_side(2, false), box_side(2, true), box_side(1, false),
             box_side(1, true),  box_side(0, false), box_side(0, true)};
}

struct ClipScratch {  // per-thread buffers, the role of Cell3DFacesScratch (cell_3d_faces.rs:6-20)
    Poly out;
    std::vector<double> d;
    std::vector<uint8_t> isx;
    std::vector<int32_t> remap;
    std::vector<std::pair<uint32_t, uint16_t>> xmap;
    std::vector<std::pair<uint16_t, uint16_t>> seg;
    std::vector<uint16_t> fbuf, lid, lmap;
};

constexpr double kTol = 1e-9;

// Half-space clip. Restates clip_with_scratch, src/cell/cell_3d_faces.rs:288-503.
// Keeps the side (x - p).n <= 1e-9. Returns {modified, max |v-g|^2 over the new vertices}; the
// radius is 0.0 when no generator is given.
std::pair<bool, double> poly_clip(Poly &c, const double p[3], const double n[3], int32_t nid,
                                  ClipScratch &s, const double *g, Work *w) {
    const size_t nv = c.nv();
    s.d.resize(nv);
    bool all_in = true, all_out = true;
    for (size_t i = 0; i < nv; ++i) {  // :303-315
        const double *q = &c.v[3 * i];
        double di = (q[0] - p[0]) * n[0] + (q[1] - p[1]) * n[1] + (q[2] - p[2]) * n[2];
        s.d[i] = di;
        if (di > kTol) all_in = false;
        else if (di < -kTol) all_out = false;
    }
    if (w) w->flops += 8 * nv;
    if (all_in) return {false, 0.0};  // :317-319
    if (all_out) {                    // :320-326
        c.v.clear(); c.fcnt.clear(); c.fidx.clear(); c.fnb.clear();
        return {true, 0.0};
    }

    Poly &o = s.out;
    o.v.clear(); o.fcnt.clear(); o.fidx.clear(); o.fnb.clear();
    s.isx.clear();
    s.remap.assign(nv, -1);
    s.xmap.clear();
    s.seg.clear();
    double max_d2 = 0.0;

    for (size_t i = 0; i < nv; ++i) {  // keep inside vertices, :345-362
        if (s.d[i] <= kTol) {
            const double *q = &c.v[3 * i];
            s.remap[i] = (int32_t)(o.v.size() / 3);
            o.v.insert(o.v.end(), q, q + 3);
            s.isx.push_back(0);
            if (g) {
                double dx = q[0] - g[0], dy = q[1] - g[1], dz = q[2] - g[2];
                double d2 = dx * dx + dy * dy + dz * dz;
                if (d2 > max_d2) max_d2 = d2;
                if (w) w->flops += 8;
            }
        }
    }

    // Intersection vertex on edge (a -> b) in traversal direction, deduplicated by the unordered
    // edge key (:391-417 and :422-448).
    auto cross = [&](size_t a, size_t b) -> uint16_t {
        uint32_t key = a < b ? ((uint32_t)a << 16 | (uint32_t)b) : ((uint32_t)b << 16 | (uint32_t)a);
        for (auto &kv : s.xmap)
            if (kv.first == key) return kv.second;
        double t = s.d[a] / (s.d[a] - s.d[b]);
        if (t < 0.0) t = 0.0;
        if (t > 1.0) t = 1.0;
        const double *A = &c.v[3 * a], *B = &c.v[3 * b];
        double x = A[0] + t * (B[0] - A[0]);
        double y = A[1] + t * (B[1] - A[1]);
        double z = A[2] + t * (B[2] - A[2]);
        uint16_t id = (uint16_t)(o.v.size() / 3);
        o.v.push_back(x); o.v.push_back(y); o.v.push_back(z);
        if (w) w->flops += 11;
        if (g) {
            double dx = x - g[0], dy = y - g[1], dz = z - g[2];
            double d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > max_d2) max_d2 = d2;
            if (w) w->flops += 8;
        }
        s.isx.push_back(1);
        s.xmap.push_back({key, id});
        return id;
    };

    size_t off = 0;
    for (size_t f = 0; f < c.fcnt.size(); ++f) {  // face walk, :370-468
        const size_t cnt = c.fcnt[f];
        const uint16_t *fi = &c.fidx[off];
        off += cnt;
        s.fbuf.clear();
        for (size_t i = 0; i < cnt; ++i) {
            size_t a = fi[i], b = fi[(i + 1) % cnt];
            bool a_in = s.d[a] <= kTol, b_in = s.d[b] <= kTol;
            if (a_in) {
                if (b_in) {
                    if (s.remap[b] >= 0) s.fbuf.push_back((uint16_t)s.remap[b]);
                } else {
                    s.fbuf.push_back(cross(a, b));
                }
            } else if (b_in) {
                s.fbuf.push_back(cross(a, b));
                if (s.remap[b] >= 0) s.fbuf.push_back((uint16_t)s.remap[b]);
            }
        }
        const size_t m = s.fbuf.size();
        if (m >= 3) {  // :454-467
            o.fcnt.push_back((uint8_t)m);
            o.fnb.push_back(c.fnb[f]);
            for (size_t i = 0; i < m; ++i) {
                uint16_t u = s.fbuf[i], v = s.fbuf[(i + 1) % m];
                if (s.isx[u] && s.isx[v]) s.seg.push_back({v, u});
            }
            o.fidx.insert(o.fidx.end(), s.fbuf.begin(), s.fbuf.end());
        }
    }

    if (!s.seg.empty()) {  // lid polygon, :471-495
        const uint16_t NONE = std::numeric_limits<uint16_t>::max();
        s.lmap.assign(o.v.size() / 3, NONE);
        for (auto &sg : s.seg) s.lmap[sg.first] = sg.second;
        s.lid.clear();
        uint16_t start = s.seg[0].first, cur = s.seg[0].second;
        s.lid.push_back(start);
        while (cur != start && s.lid.size() <= s.seg.size()) {
            s.lid.push_back(cur);
            cur = s.lmap[cur];
            if (cur == NONE) break;
        }
        if (s.lid.size() >= 3) {
            o.fcnt.push_back((uint8_t)s.lid.size());
            o.fidx.insert(o.fidx.end(), s.lid.begin(), s.lid.end());
            o.fnb.push_back(nid);
        }
    }
    std::swap(c.v, o.v);
    std::swap(c.fcnt, o.fcnt);
    std::swap(c.fidx, o.fidx);
    std::swap(c.fnb, o.fnb);
    if (w) w->size(c);
    return {true, max_d2};
}

double poly_max_radius_sq(const Poly &c, const double g[3]) {  // cell_3d_faces.rs:271-286
    double m = 0.0;
    for (size_t k = 0; k < c.nv(); ++k) {
        double dx = c.v[3 * k] - g[0], dy = c.v[3 * k + 1] - g[1], dz = c.v[3 * k + 2] - g[2];
        double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 > m) m = d2;
    }
    return m;
}

// Fan triangulation over faces with >= 3 vertices, pivot at the face's first vertex. Calls
// fn(det, v0, v1, v2) per triangle in face order. Shared traversal of volume (:106-143) and
// centroid (:145-205).
template <class F>
void poly_fan(const Poly &c, F &&fn) {
    size_t off = 0;
    for (size_t f = 0; f < c.fcnt.size(); ++f) {
        const size_t cnt = c.fcnt[f];
        if (cnt >= 3) {
            const double *a = &c.v[3 * c.fidx[off]];
            for (size_t i = 1; i + 1 < cnt; ++i) {
                const double *b = &c.v[3 * c.fidx[off + i]];
                const double *q = &c.v[3 * c.fidx[off + i + 1]];
                double cx = b[1] * q[2] - b[2] * q[1];
                double cy = b[2] * q[0] - b[0] * q[2];
                double cz = b[0] * q[1] - b[1] * q[0];
                double det = a[0] * cx + a[1] * cy + a[2] * cz;
                fn(det, a, b, q);
            }
        }
        off += cnt;
    }
}

double poly_volume(const Poly &c) {  // cell_3d_faces.rs:106-143
    double vol = 0.0;
    poly_fan(c, [&](double det, const double *, const double *, const double *) { vol += det; });
    return std::fabs(vol / 6.0);
}

void poly_centroid(const Poly &c, double out[3]) {  // cell_3d_faces.rs:145-205
    double sx = 0.0, sy = 0.0, sz = 0.0, tv = 0.0;
    poly_fan(c, [&](double det, const double *a, const double *b, const double *q) {
        tv += det;
        sx += det * (a[0] + b[0] + q[0]);
        sy += det * (a[1] + b[1] + q[1]);
        sz += det * (a[2] + b[2] + q[2]);
    });
    if (std::fabs(tv) < 1e-9) { out[0] = out[1] = out[2] = 0.0; return; }
    double f = 1.0 / (4.0 * tv);
    out[0] = sx * f; out[1] = sy * f; out[2] = sz * f;
}

uint64_t poly_triangles(const Poly &c) {
    uint64_t t = 0;
    for (uint8_t k : c.fcnt) if (k >= 3) t += k - 2;
    return t;
}

// ---------------------------------------------------------------------------------------------
// Walls. Restates src/wall/mod.rs:15-49 and src/wall/wall_3d.rs:8-101 (Plane and Sphere only).
// ---------------------------------------------------------------------------------------------
// A wall whose geometry lives with the caller (kind 7): WallGeometry::contains and ::cut
// (src/wall/mod.rs:53-71) called back through these pointers; planar is is_planar().
using HostContains = int (*)(void *ctx, const double p[3]);
using HostCut = int (*)(void *ctx, const double g[3], double *planes, int cap);

struct Wall {
    int32_t id;
    int kind;       // 0 plane, 1 sphere, 2 cylinder, 3 cone, 4 torus, 5 convex polyhedron, 6 tube,
                    // 7 host callbacks
    double a[3];    // plane point / sphere, cylinder, torus centre / cone tip
    double b[3];    // plane unit normal (pointing into the valid region) / unit axis
    double r;       // sphere, cylinder, tube radius / torus major radius / cone half-angle
    double r2;      // torus minor radius
    std::vector<double> pts;   // polyhedron planes (point, outward normal) / tube samples (xyz)
    bool closed;               // tube: the last sample joins the first
    void *hctx = nullptr;      // host callbacks (kind 7)
    HostContains hcontains = nullptr;
    HostCut hcut = nullptr;
    bool planar = false;
};

// get_closest_point of the sampled-curve walls: TrefoilKnotGeometry (wall_3d.rs:408-438, always
// closed), CubicBezierGeometry (:762-797) and CatmullRomGeometry (:908-943) share it.
void tube_closest(const Wall &w, const double q[3], double c[3]) {
    const size_t n = w.pts.size() / 3;
    c[0] = c[1] = c[2] = 0.0;
    if (n == 0) return;                             // :763-765 (Bezier, Catmull-Rom)
    double min_dist_sq = DBL_MAX;                   // f64::MAX
    for (int k = 0; k < 3; ++k) c[k] = w.pts[k];    // samples[0]
    const size_t limit = w.closed ? n : n - 1;
    for (size_t i = 0; i < limit; ++i) {
        const double *p0 = &w.pts[3 * i], *p1 = &w.pts[3 * ((i + 1) % n)];
        const double v[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
        const double ww[3] = {q[0] - p0[0], q[1] - p0[1], q[2] - p0[2]};
        const double c1 = ww[0] * v[0] + ww[1] * v[1] + ww[2] * v[2];
        const double c2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        double t = 0.0;
        if (!(c2 <= 0.0)) {
            t = c1 / c2;
            if (t < 0.0) t = 0.0;   // f64::clamp(0.0, 1.0)
            if (t > 1.0) t = 1.0;
        }
        const double proj[3] = {p0[0] + v[0] * t, p0[1] + v[1] * t, p0[2] + v[2] * t};
        const double dx = q[0] - proj[0], dy = q[1] - proj[1], dz = q[2] - proj[2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (d2 < min_dist_sq) {
            min_dist_sq = d2;
            for (int k = 0; k < 3; ++k) c[k] = proj[k];
        }
    }
}

// WallGeometry::contains, src/wall/wall_3d.rs.
bool wall_contains(const Wall &w, const double q[3]) {
    if (w.kind == 7) return w.hcontains(w.hctx, q) != 0;   // wall/mod.rs:34-36
    double dx = q[0] - w.a[0], dy = q[1] - w.a[1], dz = q[2] - w.a[2];
    switch (w.kind) {
    case 0: return (dx * w.b[0] + dy * w.b[1] + dz * w.b[2]) >= 0.0;     // :31-36
    case 1: return (dx * dx + dy * dy + dz * dz) <= w.r * w.r;           // :73-78
    case 2: {                                                            // :132-143
        double dot = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
        double px = dx - dot * w.b[0], py = dy - dot * w.b[1], pz = dz - dot * w.b[2];
        return (px * px + py * py + pz * pz) <= w.r * w.r;
    }
    case 3: {                                                            // :202-214
        double h = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
        double px = dx - h * w.b[0], py = dy - h * w.b[1], pz = dz - h * w.b[2];
        double r = std::sqrt(px * px + py * py + pz * pz);
        return h >= 0.0 && r <= h * std::tan(w.r);
    }
    case 4: {                                                            // :293-309
        double dot = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
        double px = dx - dot * w.b[0], py = dy - dot * w.b[1], pz = dz - dot * w.b[2];
        double dist_perp = std::sqrt(px * px + py * py + pz * pz);
        double t = dist_perp - w.r;
        double dist_tube = std::sqrt(t * t + dot * dot);                 // powi(2) = square
        return dist_tube <= w.r2;
    }
    case 5:                                                              // :691-701
        for (size_t k = 0; k + 6 <= w.pts.size(); k += 6) {
            double ex = q[0] - w.pts[k], ey = q[1] - w.pts[k + 1], ez = q[2] - w.pts[k + 2];
            if (ex * w.pts[k + 3] + ey * w.pts[k + 4] + ez * w.pts[k + 5] > 0.0) return false;
        }
        return true;
    case 6: {                                                            // :442-446, 801-805, 947-951
        double c[3];
        tube_closest(w, q, c);
        double ex = q[0] - c[0], ey = q[1] - c[1], ez = q[2] - c[2];
        return ex * ex + ey * ey + ez * ez <= w.r * w.r;
    }
    }
    return true;
}

// WallGeometry::cut: every plane (point, outward normal) the wall emits for generator g, in order.
template <class Emit>
void wall_cut(const Wall &w, const double g[3], Emit emit, Work *wk) {
    if (w.kind == 7) {   // wall/mod.rs:38-40: every plane the caller's cut emits, in order
        std::vector<double> buf(6 * 16);
        int m = w.hcut(w.hctx, g, buf.data(), 16);
        if (m > 16) {
            buf.resize(6 * (size_t)m);
            m = w.hcut(w.hctx, g, buf.data(), m);
        }
        for (int k = 0; k < m; ++k) emit(&buf[6 * (size_t)k], &buf[6 * (size_t)k + 3]);
        return;
    }
    double dx = g[0] - w.a[0], dy = g[1] - w.a[1], dz = g[2] - w.a[2];
    switch (w.kind) {
    case 0: {  // :38-42
        double p[3] = {w.a[0], w.a[1], w.a[2]}, n[3] = {-w.b[0], -w.b[1], -w.b[2]};
        emit(p, n);
        return;
    }
    case 1: {  // :80-100
        double dist = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (dist == 0.0) return;
        double sc = w.r / dist;
        double p[3] = {w.a[0] + dx * sc, w.a[1] + dy * sc, w.a[2] + dz * sc};
        double n[3] = {dx / dist, dy / dist, dz / dist};
        if (wk) wk->flops += 19;
        emit(p, n);
        return;
    }
    case 2: {  // :145-170
        double dot = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
        double px = dx - dot * w.b[0], py = dy - dot * w.b[1], pz = dz - dot * w.b[2];
        double dist = std::sqrt(px * px + py * py + pz * pz);
        if (dist == 0.0) return;
        double sc = w.r / dist;
        double p[3] = {w.a[0] + dot * w.b[0] + px * sc, w.a[1] + dot * w.b[1] + py * sc,
                       w.a[2] + dot * w.b[2] + pz * sc};
        double n[3] = {px / dist, py / dist, pz / dist};
        emit(p, n);
        return;
    }
    case 3: {  // :216-258
        double h = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
        double px = dx - h * w.b[0], py = dy - h * w.b[1], pz = dz - h * w.b[2];
        double r = std::sqrt(px * px + py * py + pz * pz);
        if (r == 0.0) return;
        double rx = px / r, ry = py / r, rz = pz / r;
        double cos_a = std::cos(w.r), sin_a = std::sin(w.r);
        double dist = r * cos_a - h * sin_a;
        double p2d_r = r - dist * cos_a;
        double p2d_h = h + dist * sin_a;
        if (p2d_h < 0.0) {
            double dist_tip = std::sqrt(dx * dx + dy * dy + dz * dz);
            if (dist_tip == 0.0) return;
            double p[3] = {w.a[0], w.a[1], w.a[2]}, n[3] = {dx / dist_tip, dy / dist_tip, dz / dist_tip};
            emit(p, n);
            return;
        }
        double p[3] = {w.a[0] + p2d_h * w.b[0] + p2d_r * rx, w.a[1] + p2d_h * w.b[1] + p2d_r * ry,
                       w.a[2] + p2d_h * w.b[2] + p2d_r * rz};
        double n[3] = {cos_a * rx - sin_a * w.b[0], cos_a * ry - sin_a * w.b[1], cos_a * rz - sin_a * w.b[2]};
        emit(p, n);
        return;
    }
    case 4: {  // :311-363
        double dot = dx * w.b[0] + dy * w.b[1] + dz * w.b[2];
        double px = dx - dot * w.b[0], py = dy - dot * w.b[1], pz = dz - dot * w.b[2];
        double dist_perp = std::sqrt(px * px + py * py + pz * pz);
        double ux, uy, uz;
        if (dist_perp < 1e-9) {
            double tx = 1.0, ty = 0.0, tz = 0.0;
            if (std::fabs(w.b[0]) > 0.9) { tx = 0.0; ty = 1.0; }
            double t_dot = tx * w.b[0] + ty * w.b[1] + tz * w.b[2];
            double ax = tx - t_dot * w.b[0], ay = ty - t_dot * w.b[1], az = tz - t_dot * w.b[2];
            double len = std::sqrt(ax * ax + ay * ay + az * az);
            if (len == 0.0) return;
            ux = ax / len; uy = ay / len; uz = az / len;
        } else {
            ux = px / dist_perp; uy = py / dist_perp; uz = pz / dist_perp;
        }
        double cx = w.a[0] + ux * w.r, cy = w.a[1] + uy * w.r, cz = w.a[2] + uz * w.r;
        double vx = g[0] - cx, vy = g[1] - cy, vz = g[2] - cz;
        double dist_c = std::sqrt(vx * vx + vy * vy + vz * vz);
        if (dist_c == 0.0) return;
        double n[3] = {vx / dist_c, vy / dist_c, vz / dist_c};
        double p[3] = {cx + n[0] * w.r2, cy + n[1] * w.r2, cz + n[2] * w.r2};
        emit(p, n);
        return;
    }
    case 5:    // :703-707
        for (size_t k = 0; k + 6 <= w.pts.size(); k += 6) emit(&w.pts[k], &w.pts[k + 3]);
        return;
    case 6: {  // :448-456, 807-815, 953-961
        double c[3];
        tube_closest(w, g, c);
        double ex = g[0] - c[0], ey = g[1] - c[1], ez = g[2] - c[2];
        double dist = std::sqrt(ex * ex + ey * ey + ez * ez);
        if (dist == 0.0) return;
        double n[3] = {ex / dist, ey / dist, ez / dist};
        double p[3] = {c[0] + n[0] * w.r, c[1] + n[1] * w.r, c[2] + n[2] * w.r};
        emit(p, n);
        return;
    }
    }
}

// ---------------------------------------------------------------------------------------------
// Uniform grid. Restates src/algorithm/algo_3d_grid.rs:47-202.
// ---------------------------------------------------------------------------------------------
struct ShellEntry { int32_t dx, dy, dz; double d2; };

// A prefix of the stable shell order: every entry whose key is <= thr, in order, and the key of
// the first entry after them (+inf when the prefix is the whole table). The whole table has
// (2n+1)^3 entries (8.6e8 at 474^3), so only the prefix the walks reach is built: a walk that
// reaches the end of the prefix while the next key is still within its limit asks for a longer
// one. Keys do not decrease along the stable order, so every prefix is the start of the table
// and a longer one starts with the same entries. A built prefix is immutable; calculate()'s
// threads share it and swap in the longer one under the grid's lock.
struct ShellPrefix {
    std::vector<ShellEntry> e;
    double thr = 0.0;
    double next = std::numeric_limits<double>::infinity();
};

struct Grid {
    int64_t n[3];
    double s[3], lim[3], lo[3], cs[3];
    std::vector<uint32_t> bin_start;  // CSR of the bin lists (each in ascending generator index)
    std::vector<uint32_t> bin_items;
    std::vector<uint32_t> gen_bin;
    mutable std::shared_ptr<const ShellPrefix> pre;   // guarded by mu
    std::unique_ptr<std::mutex> mu{new std::mutex};
};

// Key of shell offset (dx, dy, dz), algo_3d_grid.rs:61-66,72: the squared distance between the
// nearest faces of two bins that far apart (0 along an axis with d = 0).
inline double shell_key(const double cs[3], int64_t x, int64_t y, int64_t z) {
    auto axis = [](int64_t d, double c) -> double {
        if (d > 0) return (double)(d - 1) * c;
        if (d < 0) return (double)(-d - 1) * c;
        return 0.0;
    };
    const double mx = axis(x, cs[0]), my = axis(y, cs[1]), mz = axis(z, cs[2]);
    return mx * mx + my * my + mz * mz;
}

// The stable shell order restricted to keys <= thr (thr = +inf: the whole table), :61-76 with a
// stable sort. Offsets with |d_a| > m_a, m_a = floor(sqrt(thr) / cs_a) + 3, have an axis term of
// at least ((floor + 3) cs_a)^2 > thr, so the box |d_a| <= m_a holds every kept entry; it also
// holds the axis entry d_a = m_a, whose key is no larger than any key outside the box, so the
// smallest key > thr inside the box is the next key of the whole table. The box is enumerated
// z-outer, y, x-inner like the table, so the stable sort keeps the table's tie order.
std::shared_ptr<const ShellPrefix> shell_build(const Grid &G, double thr) {
    int64_t m[3];
    for (int k = 0; k < 3; ++k) {
        const double q = std::sqrt(thr) / G.cs[k];
        m[k] = !(q < (double)G.n[k]) ? G.n[k] : std::min<int64_t>(G.n[k], (int64_t)q + 3);
    }
    auto P = std::make_shared<ShellPrefix>();
    P->thr = thr;
    P->e.reserve((size_t)(2 * m[0] + 1) * (2 * m[1] + 1) * (2 * m[2] + 1));
    for (int64_t z = -m[2]; z <= m[2]; ++z)
        for (int64_t y = -m[1]; y <= m[1]; ++y)
            for (int64_t x = -m[0]; x <= m[0]; ++x) {
                const double d2 = shell_key(G.cs, x, y, z);
                if (d2 <= thr) P->e.push_back({(int32_t)x, (int32_t)y, (int32_t)z, d2});
                else if (d2 < P->next) P->next = d2;
            }
    // :76 sorts unstably; this restatement sorts stably (ties keep generation order).
    std::stable_sort(P->e.begin(), P->e.end(),
                     [](const ShellEntry &a, const ShellEntry &b) { return a.d2 < b.d2; });
    return P;
}

std::shared_ptr<const ShellPrefix> grid_prefix(const Grid &G) {
    std::lock_guard<std::mutex> lk(*G.mu);
    return G.pre;
}

// A prefix holding every entry with key <= need (at least four times the old threshold, so a
// cell's walk extends it a few times at most).
std::shared_ptr<const ShellPrefix> grid_extend(const Grid &G, double need) {
    std::lock_guard<std::mutex> lk(*G.mu);
    if (G.pre->thr >= need || G.pre->next == std::numeric_limits<double>::infinity()) return G.pre;
    G.pre = shell_build(G, std::max(need, 4.0 * G.pre->thr));
    return G.pre;
}

void grid_new(Grid &G, int64_t nx, int64_t ny, int64_t nz, const double lo[3], const double hi[3]) {
    G.n[0] = nx; G.n[1] = ny; G.n[2] = nz;
    double csmax = 0.0;
    for (int k = 0; k < 3; ++k) {
        G.s[k] = (double)G.n[k] / (hi[k] - lo[k]);   // :48-50
        G.cs[k] = 1.0 / G.s[k];                      // :52-54
        G.lim[k] = (double)G.n[k] - 1e-5;            // :85-87
        G.lo[k] = lo[k];
        csmax = std::max(csmax, G.cs[k]);
    }
    // the first prefix: keys up to (4 * the largest bin edge)^2, extended as walks need
    G.pre = shell_build(G, 16.0 * csmax * csmax);
}

inline size_t axis_index(double x, double lo, double s, double lim) {  // :102-104
    double v = (x - lo) * s;
    if (v < 0.0) v = 0.0;       // f64::clamp; NaN passes through
    if (v > lim) v = lim;
    if (!(v == v)) return 0;    // `NaN as usize` is 0
    return (size_t)v;           // truncation
}

size_t grid_bin_index(const Grid &G, const double q[3]) {  // :98-107
    size_t ix = axis_index(q[0], G.lo[0], G.s[0], G.lim[0]);
    size_t iy = axis_index(q[1], G.lo[1], G.s[1], G.lim[1]);
    size_t iz = axis_index(q[2], G.lo[2], G.s[2], G.lim[2]);
    return ix + iy * (size_t)G.n[0] + iz * (size_t)G.n[0] * (size_t)G.n[1];
}

void grid_set_generators(Grid &G, const std::vector<double> &gen) {  // :111-130
    const size_t nb = (size_t)(G.n[0] * G.n[1] * G.n[2]);
    const size_t cnt = gen.size() / 3;
    G.gen_bin.resize(cnt);
    G.bin_start.assign(nb + 1, 0);
    for (size_t i = 0; i < cnt; ++i) {
        G.gen_bin[i] = (uint32_t)grid_bin_index(G, &gen[3 * i]);
        G.bin_start[G.gen_bin[i] + 1]++;
    }
    for (size_t b = 0; b < nb; ++b) G.bin_start[b + 1] += G.bin_start[b];
    std::vector<uint32_t> fill(G.bin_start.begin(), G.bin_start.end() - 1);
    G.bin_items.resize(cnt);
    for (size_t i = 0; i < cnt; ++i) G.bin_items[fill[G.gen_bin[i]]++] = (uint32_t)i;  // push in i order
}

// visit_neighbors, :145-202. The visitor returns the new squared search radius.
template <class V>
void grid_visit(const Grid &G, const std::vector<double> &gen, size_t index, const double pos[3],
                double &r2, V &&visitor, Work *w) {
    const size_t b = G.gen_bin[index];
    const int64_t nxy = G.n[0] * G.n[1];
    const int64_t iz = (int64_t)b / nxy, rem = (int64_t)b % nxy;
    const int64_t iy = rem / G.n[0], ix = rem % G.n[0];
    const double csx = 1.0 / G.s[0], csy = 1.0 / G.s[1], csz = 1.0 / G.s[2];
    const double rx = (pos[0] - G.lo[0]) * G.s[0] - (double)ix;
    const double ry = (pos[1] - G.lo[1]) * G.s[1] - (double)iy;
    const double rz = (pos[2] - G.lo[2]) * G.s[2] - (double)iz;
    auto bd = [](int64_t d, double rel, double c) -> double {  // :182-184
        if (d > 0) return ((double)d - rel) * c;
        if (d < 0) return ((double)(-(d + 1)) + rel) * c;
        return 0.0;
    };
    std::shared_ptr<const ShellPrefix> pre = grid_prefix(G);
    for (size_t k = 0;;) {
        if (k == pre->e.size()) {
            // end of the prefix: the table's next entry ends the walk, or a longer prefix is needed
            if (pre->next > 4.0 * r2) {
                if (w && pre->next != std::numeric_limits<double>::infinity()) w->entries++;
                break;
            }
            pre = grid_extend(G, 4.0 * r2);
            continue;
        }
        const ShellEntry &e = pre->e[k++];
        if (w) w->entries++;
        if (e.d2 > 4.0 * r2) break;  // :170
        const int64_t bx = ix + e.dx, by = iy + e.dy, bz = iz + e.dz;
        if (bx < 0 || bx >= G.n[0] || by < 0 || by >= G.n[1] || bz < 0 || bz >= G.n[2]) continue;
        double ex = std::max(bd(e.dx, rx, csx), 0.0);
        double ey = std::max(bd(e.dy, ry, csy), 0.0);
        double ez = std::max(bd(e.dz, rz, csz), 0.0);
        if (w) { w->bins++; w->flops += 12; }
        if (ex * ex + ey * ey + ez * ez <= 4.0 * r2) {  // :190
            const size_t bi = (size_t)(bx + by * G.n[0] + bz * nxy);
            for (uint32_t k = G.bin_start[bi]; k < G.bin_start[bi + 1]; ++k) {
                const uint32_t j = G.bin_items[k];
                if (j == index) continue;
                r2 = visitor(j, &gen[3 * j], r2);
            }
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Tessellation. Restates src/tessellation.rs:13-67, 175-230, 414-475, 482-495.
// ---------------------------------------------------------------------------------------------
struct Tess {
    double lo[3], hi[3];
    std::vector<double> gen;
    std::vector<Wall> walls;
    Grid grid;
    std::vector<Poly> cells;
    std::vector<uint8_t> computed;
    std::vector<Work> work;   // per cell
    std::vector<int32_t> seal_log, prune_log;   // tessellation.rs:19-21
    std::vector<double> prune_pos_log;
};

void tess_set_generators(Tess &T, const double *xyz, size_t n) {  // tessellation.rs:43-67
    std::vector<double> keep;
    keep.reserve(3 * n);
    for (size_t i = 0; i < n; ++i) {
        const double *q = xyz + 3 * i;
        bool in = true;
        for (const Wall &w : T.walls)
            if (!wall_contains(w, q)) { in = false; break; }
        if (in) keep.insert(keep.end(), q, q + 3);
    }
    T.gen.swap(keep);
    grid_set_generators(T.grid, T.gen);
    T.cells.clear();
    T.computed.clear();
}

// Output k of the SplitMix64 stream of `seed` (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014): the state advances by 0x9e3779b97f4a7c15 per
// output and each output is the finaliser of the new state.
uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

// random_generators, tessellation.rs:96-122, one candidate at a time: the reference draws
// rng.gen::<f64>() per axis from a StdRng seeded by thread_rng; here the doubles are
// (x >> 11) * 2^-53 of consecutive SplitMix64 outputs of `seed`, three per candidate.
void tess_random_generators(Tess &T, size_t count, uint64_t seed) {
    std::vector<double> points;
    points.reserve(3 * count);
    size_t found = 0;
    const size_t max_attempts = count * 1000;                                // :101
    size_t attempts = 0;
    uint64_t k = 0;
    while (found < count && attempts < max_attempts) {                       // :104
        attempts += 1;
        double point[3];
        for (int i = 0; i < 3; ++i) {                                        // :107-111
            const double u = (double)(splitmix64_at(seed, k++) >> 11) * 0x1.0p-53;
            point[i] = T.lo[i] + u * (T.hi[i] - T.lo[i]);
        }
        bool in = true;                                                      // :113
        for (const Wall &w : T.walls)
            if (!wall_contains(w, point)) { in = false; break; }
        if (in) {
            points.insert(points.end(), point, point + 3);
            found += 1;
        }
    }
    T.gen.swap(points);                                                      // :119
    grid_set_generators(T.grid, T.gen);                                      // :120
    T.cells.clear();
    T.computed.clear();
}

void tess_add_wall(Tess &T, const Wall &w) {  // tessellation.rs:175-200
    T.walls.push_back(w);
    std::vector<double> keep;
    const size_t n = T.gen.size() / 3;
    for (size_t i = 0; i < n; ++i) {
        const double *q = &T.gen[3 * i];
        bool in = true;
        for (const Wall &x : T.walls)
            if (!wall_contains(x, q)) { in = false; break; }
        if (in) keep.insert(keep.end(), q, q + 3);
    }
    if (keep.size() != T.gen.size()) {
        T.gen.swap(keep);
        grid_set_generators(T.grid, T.gen);
    }
}

// set_generator, tessellation.rs:71-92, with update_generator, algo_3d_grid.rs:132-143. Returns 0
// when the generator moved, 1 for an index past the generators and 2 for a position a wall does
// not contain (the reference ignores both calls).
int tess_set_generator(Tess &T, size_t index, const double q[3]) {
    const size_t offset = index * 3;
    if (offset + 3 > T.gen.size()) return 1;                  // :73-75
    for (const Wall &w : T.walls)
        if (!wall_contains(w, q)) return 2;                    // :77-81
    Grid &G = T.grid;
    const uint32_t nb = (uint32_t)grid_bin_index(G, q);       // :133
    const uint32_t ob = G.gen_bin[index];                      // :134
    if (nb != ob) {                                            // :136
        const size_t nbins = G.bin_start.size() - 1;
        std::vector<std::vector<uint32_t>> bins(nbins);        // grid_bins
        for (size_t b = 0; b < nbins; ++b)
            bins[b].assign(G.bin_items.begin() + G.bin_start[b], G.bin_items.begin() + G.bin_start[b + 1]);
        std::vector<uint32_t> &old = bins[ob];
        const auto it = std::find(old.begin(), old.end(), (uint32_t)index);   // :137
        if (it != old.end()) {                                 // :138 swap_remove
            *it = old.back();
            old.pop_back();
        }
        bins[nb].push_back((uint32_t)index);                   // :140
        G.gen_bin[index] = nb;                                 // :141
        uint32_t k = 0;
        for (size_t b = 0; b < nbins; ++b) {
            G.bin_start[b] = k;
            for (uint32_t v : bins[b]) G.bin_items[k++] = v;
        }
        G.bin_start[nbins] = k;
    }
    std::memcpy(&T.gen[offset], q, 3 * sizeof(double));       // :89-91
    return 0;
}

// compute_cell, tessellation.rs:414-475. brute != 0 visits every other generator in index order
// with no grid and no early exit (used only to pin the grid walk).
void compute_cell(const Tess &T, size_t i, Poly &cell, ClipScratch &s, Work *w, int brute) {
    const double *g = &T.gen[3 * i];
    poly_box(cell, T.lo, T.hi);
    for (const Wall &wl : T.walls) {  // :429-436
        wall_cut(wl, g, [&](const double *p, const double *n) {
            if (w) w->clips++;
            auto r = poly_clip(cell, p, n, wl.id, s, nullptr, w);
            if (w && r.first) w->cuts++;
        }, w);
        if (cell.empty()) return;
    }
    double r2 = poly_max_radius_sq(cell, g);  // :438
    if (w) w->flops += 8 * cell.nv();
    auto visitor = [&](uint32_t j, const double *q, double cur) -> double {  // :446-471
        double dist_sq = 0.0, mid[3], nrm[3];
        for (int k = 0; k < 3; ++k) {
            double d = q[k] - g[k];
            dist_sq += d * d;
            mid[k] = g[k] + d * 0.5;
            nrm[k] = d;
        }
        if (w) { w->candidates++; w->flops += 16; }
        if (dist_sq > 4.0 * cur) return cur;
        if (w) w->clips++;
        auto r = poly_clip(cell, mid, nrm, (int32_t)j, s, g, w);
        if (r.first) {
            if (w) w->cuts++;
            return cell.empty() ? 0.0 : r.second;
        }
        return cur;
    };
    if (!brute) {
        grid_visit(T.grid, T.gen, i, g, r2, visitor, w);
    } else {
        const size_t n = T.gen.size() / 3;
        for (size_t j = 0; j < n; ++j)
            if (j != i) r2 = visitor((uint32_t)j, &T.gen[3 * j], 1e300);
    }
}

void tess_calculate(Tess &T, const int64_t *subset, size_t m, int threads, int brute) {
    const size_t n = T.gen.size() / 3;
    T.seal_log.clear();   // tessellation.rs:214-216
    T.prune_log.clear();
    T.prune_pos_log.clear();
    T.cells.assign(n, Poly());
    T.computed.assign(n, 0);
    T.work.assign(n, Work());
    const size_t total = subset ? m : n;
    std::atomic<size_t> next{0};
    auto run = [&]() {
        ClipScratch s;
        for (;;) {
            size_t a = next.fetch_add(64);
            if (a >= total) break;
            size_t b = std::min(total, a + 64);
            for (size_t k = a; k < b; ++k) {
                size_t i = subset ? (size_t)subset[k] : k;
                if (i >= n) continue;
                Work w;
                compute_cell(T, i, T.cells[i], s, &w, brute);
                // reductions done by the caller side of the reference (volume/centroid) are part
                // of the measured output here, as they are in the library's reduced mode
                uint64_t tri = poly_triangles(T.cells[i]);
                w.flops += 15 * tri + 1 + 27 * tri + 5;
                T.work[i] = w;
                T.computed[i] = 1;
            }
        }
    };
    if (threads <= 1) {
        run();
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t) pool.emplace_back(run);
        for (auto &th : pool) th.join();
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// C ABI (ctypes) for the tests and the CPU baseline. Prefix "or_" = oracle.
// ---------------------------------------------------------------------------------------------
extern "C" {

void *or_tess_new(const double lo[3], const double hi[3], int64_t nx, int64_t ny, int64_t nz) {
    Tess *T = new Tess();
    for (int k = 0; k < 3; ++k) { T->lo[k] = lo[k]; T->hi[k] = hi[k]; }
    grid_new(T->grid, nx, ny, nz, lo, hi);
    grid_set_generators(T->grid, T->gen);
    return T;
}

void or_tess_free(void *t) { delete (Tess *)t; }

int or_tess_add_wall_plane(void *t, int32_t id, const double p[3], const double n[3]) {
    if (id > -1000) return -1;  // wall/mod.rs:25-27 panics
    Wall w{};
    w.id = id; w.kind = 0;
    double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);  // wall_3d.rs:23-27
    if (len == 0.0) { w.b[0] = 0.0; w.b[1] = 0.0; w.b[2] = 1.0; }
    else { w.b[0] = n[0] / len; w.b[1] = n[1] / len; w.b[2] = n[2] / len; }
    for (int k = 0; k < 3; ++k) w.a[k] = p[k];
    tess_add_wall(*(Tess *)t, w);
    return 0;
}

int or_tess_add_wall_sphere(void *t, int32_t id, const double c[3], double r) {
    if (id > -1000) return -1;
    Wall w{};
    w.id = id; w.kind = 1; w.r = r;
    for (int k = 0; k < 3; ++k) w.a[k] = c[k];
    tess_add_wall(*(Tess *)t, w);
    return 0;
}

// Any wall kind: 2 CylinderGeometry::new(a, b, r) (wall_3d.rs:124-128), 3 ConeGeometry::new(a, b,
// angle r) (:194-198), 4 TorusGeometry::new(a, b, r, r2) (:285-289), 5 ConvexPolyhedronGeometry
// with npts planes in pts (point, outward normal; :476-491), 6 a sampled-curve tube with npts
// samples in pts, radius r, closed as given; 0 and 1 as the two calls above.
int or_tess_add_wall(void *t, int32_t id, int kind, const double a[3], const double b[3], double r, double r2,
                     const double *pts, uint32_t npts, int closed) {
    if (id > -1000) return -1;
    if (kind < 0 || kind > 6) return -2;
    Wall w{};
    w.id = id; w.kind = kind; w.r = r; w.r2 = r2; w.closed = closed != 0;
    for (int k = 0; k < 3; ++k) w.a[k] = a[k];
    if (kind == 0 || kind == 2 || kind == 3 || kind == 4) {
        double len = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
        if (len == 0.0) { w.b[0] = 0.0; w.b[1] = 0.0; w.b[2] = 1.0; }
        else { w.b[0] = b[0] / len; w.b[1] = b[1] / len; w.b[2] = b[2] / len; }
    }
    if (kind == 5) w.pts.assign(pts, pts + 6 * (size_t)npts);
    if (kind == 6) w.pts.assign(pts, pts + 3 * (size_t)npts);
    tess_add_wall(*(Tess *)t, w);
    return 0;
}

// Wall::new(id, Box<dyn WallGeometry>) for a geometry of the caller's, evaluated through callbacks.
int or_tess_add_wall_host(void *t, int32_t id, void *ctx, HostContains contains, HostCut cut, int planar) {
    if (id > -1000) return -1;
    Wall w{};
    w.id = id; w.kind = 7;
    w.hctx = ctx; w.hcontains = contains; w.hcut = cut; w.planar = planar != 0;
    tess_add_wall(*(Tess *)t, w);
    return 0;
}

void or_tess_set_generators(void *t, const double *xyz, size_t n) {
    tess_set_generators(*(Tess *)t, xyz, n);
}

void or_tess_random_generators(void *t, size_t count, uint64_t seed) {
    tess_random_generators(*(Tess *)t, count, seed);
}

uint64_t or_splitmix64_at(uint64_t seed, uint64_t k) { return splitmix64_at(seed, k); }

int or_tess_set_generator(void *t, size_t index, const double q[3]) {
    return tess_set_generator(*(Tess *)t, index, q);
}

size_t or_tess_count_generators(void *t) { return ((Tess *)t)->gen.size() / 3; }

void or_tess_get_generators(void *t, double *out) {
    Tess &T = *(Tess *)t;
    std::memcpy(out, T.gen.data(), T.gen.size() * sizeof(double));
}

void or_tess_get_bins(void *t, uint32_t *bin_start, uint32_t *bin_items, uint32_t *gen_bin) {
    Tess &T = *(Tess *)t;
    if (bin_start) std::memcpy(bin_start, T.grid.bin_start.data(), T.grid.bin_start.size() * 4);
    if (bin_items) std::memcpy(bin_items, T.grid.bin_items.data(), T.grid.bin_items.size() * 4);
    if (gen_bin) std::memcpy(gen_bin, T.grid.gen_bin.data(), T.grid.gen_bin.size() * 4);
}

size_t or_tess_bin_index(void *t, const double q[3]) { return grid_bin_index(((Tess *)t)->grid, q); }

// Stable shell order restricted to keys <= thr (thr = +inf: the whole table) as (dx, dy, dz,
// min_d2) -> caller buffers of size cap; *next receives the key of the first entry left out.
// Returns the prefix length.
size_t or_tess_shell_prefix(void *t, double thr, int32_t *dxyz, double *d2, size_t cap, double *next) {
    Tess &T = *(Tess *)t;
    auto P = shell_build(T.grid, thr);
    size_t m = std::min(cap, P->e.size());
    for (size_t k = 0; k < m; ++k) {
        dxyz[3 * k] = P->e[k].dx;
        dxyz[3 * k + 1] = P->e[k].dy;
        dxyz[3 * k + 2] = P->e[k].dz;
        d2[k] = P->e[k].d2;
    }
    if (next) *next = P->next;
    return P->e.size();
}

// Full stable shell order (dx, dy, dz, min_d2) -> caller buffers of size cap.
size_t or_tess_shell_order(void *t, int32_t *dxyz, double *d2, size_t cap) {
    return or_tess_shell_prefix(t, std::numeric_limits<double>::infinity(), dxyz, d2, cap, nullptr);
}

// Walks start from the whole table (full = 1) instead of the lazily extended prefix (full = 0),
// the construction the tests compare the lazy one against. Returns the length of the prefix the
// walks now start from.
size_t or_tess_shell_mode(void *t, int full) {
    Tess &T = *(Tess *)t;
    std::lock_guard<std::mutex> lk(*T.grid.mu);
    double csmax = std::max(T.grid.cs[0], std::max(T.grid.cs[1], T.grid.cs[2]));
    T.grid.pre = shell_build(T.grid, full ? std::numeric_limits<double>::infinity() : 16.0 * csmax * csmax);
    return T.grid.pre->e.size();
}

// Length of the prefix the walks use now (it grows as walks need).
size_t or_tess_shell_len(void *t) {
    Tess &T = *(Tess *)t;
    std::lock_guard<std::mutex> lk(*T.grid.mu);
    return T.grid.pre->e.size();
}

// ---------------------------------------------------------------------------------------------
// Post-passes. Restate src/tessellation.rs:234-385 and Cell3DFaces::shared_vertices
// (src/cell/cell_3d_faces.rs:536-568).
// ---------------------------------------------------------------------------------------------
bool wall_planar(const Wall &w) {   // wall_3d.rs:44, 709; mod.rs:70 for the caller's geometry
    return w.kind == 0 || w.kind == 5 || (w.kind == 7 && w.planar);
}

// Neighbour generators per cell in face order (deduplicated for prune_boundaries) and the ids of
// the curved walls it touches (tessellation.rs:244-264, 330-350).
void cell_topology(const Tess &T, bool dedupe, std::vector<std::vector<size_t>> &topo,
                   std::vector<std::vector<int32_t>> &cw) {
    const size_t count = T.gen.size() / 3;
    topo.assign(count, {});
    cw.assign(count, {});
    for (size_t i = 0; i < count; ++i) {
        for (int32_t fn : T.cells[i].fnb) {
            if (fn >= 0 && (size_t)fn < count) {
                if (!dedupe || std::find(topo[i].begin(), topo[i].end(), (size_t)fn) == topo[i].end())
                    topo[i].push_back((size_t)fn);
            } else {
                const Wall *w = nullptr;
                for (const Wall &x : T.walls)
                    if (x.id == fn) { w = &x; break; }
                if (w && !wall_planar(*w) && std::find(cw[i].begin(), cw[i].end(), fn) == cw[i].end())
                    cw[i].push_back(fn);
            }
        }
    }
}

// calculate_sealed after calculate (tessellation.rs:238-317); cells are independent, so the
// sequential loop gives what the reference's parallel one gives.
void tess_seal(Tess &T) {
    std::vector<std::vector<size_t>> topo;
    std::vector<std::vector<int32_t>> cw;
    cell_topology(T, false, topo, cw);
    T.seal_log.clear();
    ClipScratch s;
    auto has = [](const std::vector<int32_t> &v, int32_t x) { return std::find(v.begin(), v.end(), x) != v.end(); };
    for (size_t i = 0; i < T.cells.size(); ++i) {
        if (cw[i].empty()) continue;
        const double *g = &T.gen[3 * i];
        for (size_t nb : topo[i]) {
            const double *q = &T.gen[3 * nb];
            for (const Wall &wall : T.walls) {
                if (wall_planar(wall) || !has(cw[i], wall.id) || !has(cw[nb], wall.id)) continue;
                wall_cut(wall, q, [&](const double *p, const double *n) {
                    double dot = 0.0;
                    for (int k = 0; k < 3; ++k) dot += (g[k] - p[k]) * n[k];
                    if (dot <= 1e-9) {
                        auto r = poly_clip(T.cells[i], p, n, wall.id, s, nullptr, nullptr);
                        if (r.first) {
                            T.seal_log.push_back((int32_t)i);
                            T.seal_log.push_back((int32_t)nb);
                            T.seal_log.push_back(wall.id);
                        }
                    }
                }, nullptr);
            }
        }
    }
}

// Cell3DFaces::shared_vertices (cell_3d_faces.rs:536-568).
std::vector<double> shared_vertices(const Poly &c, int32_t a, int32_t b) {
    std::vector<size_t> av, bv;
    size_t off = 0;
    for (size_t f = 0; f < c.fcnt.size(); ++f) {
        const size_t cnt = c.fcnt[f];
        for (size_t j = 0; j < cnt; ++j) {
            const size_t v = c.fidx[off + j];
            if (c.fnb[f] == a && std::find(av.begin(), av.end(), v) == av.end()) av.push_back(v);
            if (c.fnb[f] == b && std::find(bv.begin(), bv.end(), v) == bv.end()) bv.push_back(v);
        }
        off += cnt;
    }
    std::vector<double> out;
    for (size_t v : av)
        if (std::find(bv.begin(), bv.end(), v) != bv.end()) out.insert(out.end(), &c.v[3 * v], &c.v[3 * v] + 3);
    return out;
}

// prune_boundaries (tessellation.rs:324-385).
void tess_prune(Tess &T) {
    std::vector<std::vector<size_t>> topo;
    std::vector<std::vector<int32_t>> cw;
    cell_topology(T, true, topo, cw);
    T.prune_log.clear();
    T.prune_pos_log.clear();
    auto has = [](const std::vector<int32_t> &v, int32_t x) { return std::find(v.begin(), v.end(), x) != v.end(); };
    for (size_t i = 0; i < T.cells.size(); ++i) {
        if (cw[i].empty()) continue;
        for (size_t nb : topo[i])
            for (const Wall &wall : T.walls) {
                if (wall_planar(wall) || !has(cw[i], wall.id) || !has(cw[nb], wall.id)) continue;
                const std::vector<double> sh = shared_vertices(T.cells[i], wall.id, (int32_t)nb);
                if (sh.empty()) continue;
                T.prune_log.push_back((int32_t)i);
                T.prune_log.push_back((int32_t)nb);
                T.prune_log.push_back(wall.id);
                T.prune_log.push_back((int32_t)(sh.size() / 3));
                T.prune_pos_log.insert(T.prune_pos_log.end(), sh.begin(), sh.end());
            }
    }
}

// Compute all cells (subset == NULL) or only the listed cell indices.
void or_tess_calculate(void *t, const int64_t *subset, size_t m, int threads, int brute) {
    tess_calculate(*(Tess *)t, subset, m, threads, brute);
}

size_t or_tess_count_cells(void *t) { return ((Tess *)t)->cells.size(); }

void or_tess_calculate_sealed(void *t, int threads) {
    Tess &T = *(Tess *)t;
    tess_calculate(T, nullptr, 0, threads, 0);
    tess_seal(T);
}

void or_tess_prune_boundaries(void *t) { tess_prune(*(Tess *)t); }

size_t or_tess_seal_log(void *t, int32_t *out, size_t cap) {
    const Tess &T = *(Tess *)t;
    for (size_t k = 0; k < std::min(cap, T.seal_log.size()); ++k) out[k] = T.seal_log[k];
    return T.seal_log.size();
}

void or_tess_prune_log(void *t, int32_t *ints, size_t icap, double *pos, size_t pcap, uint64_t counts[2]) {
    const Tess &T = *(Tess *)t;
    for (size_t k = 0; k < std::min(icap, T.prune_log.size()); ++k) ints[k] = T.prune_log[k];
    for (size_t k = 0; k < std::min(pcap, T.prune_pos_log.size()); ++k) pos[k] = T.prune_pos_log[k];
    counts[0] = T.prune_log.size();
    counts[1] = T.prune_pos_log.size();
}

// Tessellation::relax, tessellation.rs:482-495 (requires a full calculate first).
void or_tess_relax(void *t) {
    Tess &T = *(Tess *)t;
    const size_t n = T.gen.size() / 3;
    std::vector<double> ng(3 * n);
    for (size_t i = 0; i < n; ++i) {
        if (T.cells[i].empty()) {
            for (int k = 0; k < 3; ++k) ng[3 * i + k] = T.gen[3 * i + k];
        } else {
            poly_centroid(T.cells[i], &ng[3 * i]);
        }
    }
    tess_set_generators(T, ng.data(), n);
}

// sizes[0] = vertices, sizes[1] = faces, sizes[2] = face index entries. Returns 0 when the cell
// was computed, -1 otherwise.
int or_cell_sizes(void *t, size_t i, uint64_t sizes[3]) {
    Tess &T = *(Tess *)t;
    if (i >= T.cells.size() || !T.computed[i]) return -1;
    const Poly &c = T.cells[i];
    sizes[0] = c.nv(); sizes[1] = c.fcnt.size(); sizes[2] = c.fidx.size();
    return 0;
}

int or_cell_get(void *t, size_t i, double *v, uint32_t *fc, uint32_t *fi, int32_t *fn) {
    Tess &T = *(Tess *)t;
    if (i >= T.cells.size() || !T.computed[i]) return -1;
    const Poly &c = T.cells[i];
    if (v) std::memcpy(v, c.v.data(), c.v.size() * 8);
    if (fc) for (size_t k = 0; k < c.fcnt.size(); ++k) fc[k] = c.fcnt[k];
    if (fi) for (size_t k = 0; k < c.fidx.size(); ++k) fi[k] = c.fidx[k];
    if (fn) std::memcpy(fn, c.fnb.data(), c.fnb.size() * 4);
    return 0;
}

// Neighbour ids of the listed cells (or all when idx == NULL) in face order, as CSR: off has
// m + 1 entries; ids (room for cap) may be NULL to query the offsets only.
void or_tess_neighbors(void *t, const int64_t *idx, size_t m, uint64_t *off, int32_t *ids, size_t cap) {
    Tess &T = *(Tess *)t;
    const size_t total = idx ? m : T.cells.size();
    uint64_t o = 0;
    for (size_t k = 0; k < total; ++k) {
        const Poly &c = T.cells[idx ? (size_t)idx[k] : k];
        off[k] = o;
        if (ids)
            for (size_t f = 0; f < c.fnb.size() && o + f < cap; ++f) ids[o + f] = c.fnb[f];
        o += c.fnb.size();
    }
    off[total] = o;
}

// Per-cell reductions for the listed cells (or all when idx == NULL): volume, centroid, face
// count; and the algorithmic-work counters [flops, candidates, bins, clips, cuts, entries,
// largest vertex count, largest face-index count, largest face count].
void or_tess_reduced(void *t, const int64_t *idx, size_t m, double *vol, double *cen,
                     uint32_t *nf, uint64_t *work6) {
    Tess &T = *(Tess *)t;
    const size_t total = idx ? m : T.cells.size();
    for (size_t k = 0; k < total; ++k) {
        size_t i = idx ? (size_t)idx[k] : k;
        const Poly &c = T.cells[i];
        if (vol) vol[k] = poly_volume(c);
        if (cen) poly_centroid(c, &cen[3 * k]);
        if (nf) nf[k] = (uint32_t)c.fcnt.size();
        if (work6) {
            const Work &w = T.work[i];
            uint64_t *o = work6 + 9 * k;
            o[0] = w.flops; o[1] = w.candidates; o[2] = w.bins; o[3] = w.clips; o[4] = w.cuts;
            o[5] = w.entries; o[6] = w.maxv; o[7] = w.maxs; o[8] = w.maxf;
        }
    }
}

// Standalone cell (for the reference's unit tests on Cell3DFaces).
void *or_cell_new(const double lo[3], const double hi[3]) {
    Poly *c = new Poly();
    poly_box(*c, lo, hi);
    return c;
}
void or_cell_free(void *c) { delete (Poly *)c; }
int or_cell_clip(void *c, const double p[3], const double n[3], int32_t nid, const double *g,
                 double *r_out) {
    ClipScratch s;
    auto r = poly_clip(*(Poly *)c, p, n, nid, s, g, nullptr);
    if (r_out) *r_out = r.second;
    return r.first ? 1 : 0;
}
double or_cell_volume(void *c) { return poly_volume(*(Poly *)c); }
void or_cell_centroid(void *c, double out[3]) { poly_centroid(*(Poly *)c, out); }
void or_cell_counts(void *c, uint64_t sizes[3]) {
    Poly &p = *(Poly *)c;
    sizes[0] = p.nv(); sizes[1] = p.fcnt.size(); sizes[2] = p.fidx.size();
}

}  // extern "C"
