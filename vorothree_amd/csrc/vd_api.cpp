// C ABI of libvoronoid_hip.so (declared in include/vd.h).
//
// An Engine owns the device state of a set of tessellations ("members"): their input points, the
// wall filter and compaction, the concatenated bins, the shell-order prefixes (one per distinct
// grid geometry), the outputs, and the cell launches with their retry tiers. vd_tess (the
// reference's Tessellation) is an Engine with one member; vd_batch holds many and computes them in
// one persistent launch.
#include "vd.h"
#include "vd_internal.h"
#include "vd_walls.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <limits>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace {

constexpr int kTiers = 5;   // polytope capacity tiers of the cell kernel (vd_cells.hip)
constexpr int kGlobalTier = 4;      // the tier whose polytopes live in global memory
constexpr int kGlobalSlots = 32;    // its waves (one polytope slot each)
constexpr uint64_t kArenaChunk[3] = {512, 256, 1536};   // vertices, faces, face indices per wave chunk
// Far cells whose rerun walks about this many shell entries or more walk a per-cell list of the
// entries that hold generators (k_far_lists); at most kFarListCells cells per launch and
// kFarListEntries list entries in all (the cells with the most entries to walk first), beyond
// which cells walk the shell prefix as before.
constexpr double kFarListMinDefault = 262144.0;
constexpr uint32_t kFarListCells = 256;
constexpr uint64_t kFarListEntries = uint64_t(1) << 27;

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define VD_HIP(call)                                                                    \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(e_ == hipErrorOutOfMemory ? VD_ENOMEM : VD_EDEVICE,             \
                        std::string(#call) + ": " + hipGetErrorString(e_));             \
    } while (0)

#define VD_TRY(expr)                  \
    do {                              \
        int rc_ = (expr);             \
        if (rc_ != VD_OK) return rc_; \
    } while (0)

// Device bytes held by one owner (a server: its lanes' engines, its jobs' buffers), and their
// high-water mark. Every hipMalloc'd DevBuf remembers the account that was current when it was
// allocated and gives its bytes back when it is freed, on whatever thread that happens.
struct AllocAccount {
    std::atomic<int64_t> bytes{0}, peak{0};
    void add(int64_t d) {
        const int64_t v = bytes.fetch_add(d) + d;
        int64_t p = peak.load();
        while (v > p && !peak.compare_exchange_weak(p, v)) {}
    }
};
thread_local AllocAccount *g_acct = nullptr;

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool carved = false;   // carved from a static arena: never passed to hipFree
    AllocAccount *acct = nullptr;
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// While g_defer is set, released device memory goes to that list instead of to hipFree, which waits
// for the whole device, work that other threads have in flight included.
thread_local std::vector<DevBuf> *g_defer = nullptr;

// While g_static is set, ensure() carves buffers from that device range (a bump allocator) instead
// of calling hipMalloc, and release() returns only the most recently carved buffer to it.
struct StaticArena {
    char *base;
    size_t cap, top;
};
thread_local StaticArena *g_static = nullptr;

void release(DevBuf &b) {
    if (b.p) {
        if (b.carved) {
            char *end = static_cast<char *>(b.p) + b.bytes;
            if (g_static && end == g_static->base + g_static->top)
                g_static->top = (size_t)(static_cast<char *>(b.p) - g_static->base);
        } else if (g_defer) {
            g_defer->push_back(b);   // still held: its account is charged until it is freed
        } else {
            (void)hipFree(b.p);
            if (b.acct) b.acct->add(-(int64_t)b.bytes);
        }
    }
    b.p = nullptr;
    b.bytes = 0;
    b.carved = false;
    b.acct = nullptr;
}

// Releases n buffers when the scope ends, on every return path (release is idempotent, so an
// explicit release before the end is fine).
struct ScopedBufs {
    DevBuf *b;
    size_t n;
    ~ScopedBufs() { for (size_t k = 0; k < n; ++k) release(b[k]); }
};

int ensure(DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return VD_OK;
    release(b);
    if (g_static) {
        const size_t at = (g_static->top + 255) & ~(size_t)255;
        if (at > g_static->cap || bytes > g_static->cap - at)
            return fail(VD_ENOMEM, "static scratch exhausted: " + std::to_string(bytes) + " more bytes needed at " +
                                       std::to_string(at) + " of " + std::to_string(g_static->cap));
        b.p = g_static->base + at;
        b.bytes = bytes;
        b.carved = true;
        g_static->top = at + bytes;
        return VD_OK;
    }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(VD_ENOMEM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    b.bytes = bytes;
    b.acct = g_acct;
    if (g_acct) g_acct->add((int64_t)bytes);
    return VD_OK;
}

// ------------------------------------------------------------------------------------------
// Shell order prefix (Algorithm3DGrid::new, algo_3d_grid.rs:47-95, with a stable sort).
// ------------------------------------------------------------------------------------------
struct ShellEntryHost {
    double d2;
    int16_t dx, dy, dz;
};

struct ShellPrefix {
    std::vector<ShellEntryHost> e;
    double next = std::numeric_limits<double>::infinity();
    bool complete = false;
};


// The enumeration box of a prefix: all entries with key <= T, T = ((kmax - 1) * max cs)^2, lie in
// [-K, K] per axis; `outside` is the smallest key of an offset beyond the box (+inf if the box
// covers the whole table).
struct ShellBox {
    int64_t K[3];
    double T;
    double outside;
};

ShellBox shell_box(const int64_t n[3], const double cs[3], uint32_t kmax) {
    ShellBox b;
    const double csm = std::max(cs[0], std::max(cs[1], cs[2]));
    const double tm = (double)(kmax > 0 ? kmax - 1 : 0) * csm;
    b.T = tm * tm;
    b.outside = std::numeric_limits<double>::infinity();
    for (int a = 0; a < 3; ++a) {
        int64_t k = 1;
        while (k < n[a]) {
            const double m = vd::shell_axis(k + 1, cs[a]);
            if (m * m > b.T) break;
            ++k;
        }
        b.K[a] = std::min<int64_t>(k, n[a]);
        // the smallest key beyond the box has one axis at K+1 and the others at 0
        if (b.K[a] < n[a]) {
            const double m = vd::shell_axis(b.K[a] + 1, cs[a]);
            b.outside = std::min(b.outside, m * m);
        }
    }
    return b;
}

// Host construction of the same prefix (used by vd_shell_prefix; the engine builds on the device).
void build_shell(const int64_t n[3], const double cs[3], uint32_t kmax, ShellPrefix &out) {
    const ShellBox b = shell_box(n, cs, kmax);
    const int64_t *K = b.K;
    struct Tmp { double d2; uint64_t gen; int16_t dx, dy, dz; };
    std::vector<Tmp> all;
    all.reserve((size_t)(2 * K[0] + 1) * (2 * K[1] + 1) * (2 * K[2] + 1));
    double next = b.outside;
    uint64_t g = 0;
    for (int64_t z = -K[2]; z <= K[2]; ++z)
        for (int64_t y = -K[1]; y <= K[1]; ++y)
            for (int64_t x = -K[0]; x <= K[0]; ++x, ++g) {
                const double d2 = vd::shell_key((int)x, (int)y, (int)z, cs);
                if (d2 <= b.T) all.push_back({d2, g, (int16_t)x, (int16_t)y, (int16_t)z});
                else next = std::min(next, d2);
            }
    std::sort(all.begin(), all.end(), [](const Tmp &p, const Tmp &q) {
        return p.d2 < q.d2 || (p.d2 == q.d2 && p.gen < q.gen);
    });
    out.e.resize(all.size());
    for (size_t k = 0; k < all.size(); ++k) out.e[k] = ShellEntryHost{all[k].d2, all[k].dx, all[k].dy, all[k].dz};
    out.next = next;
    out.complete = std::isinf(next);
}

// Algorithm3DGrid::new's scale factors (algo_3d_grid.rs:48-54, 85-87).
vd::GridGeom make_grid(const double lo[3], const double hi[3], const int64_t n[3]) {
    vd::GridGeom g{};
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = lo[a];
        g.s[a] = (double)n[a] / (hi[a] - lo[a]);
        g.cs[a] = 1.0 / g.s[a];
        g.lim[a] = (double)n[a] - 1e-5;
        g.n[a] = (int32_t)n[a];
    }
    return g;
}

// Default generators-per-bin target of grids the library picks (batch members and served jobs
// without a grid): the fastest for uniform input in a sweep of 0.25 to 2 on MI355X. Smaller bins
// cost more shell entries, which the walk tests 64 at a time, and save candidate clips, which run
// one after another.
constexpr double kPerBinDefault = 0.35;

void grid_for(const double lo[3], const double hi[3], size_t n, double per_bin, int64_t out[3]) {
    double e[3], vol = 1.0;
    for (int a = 0; a < 3; ++a) { e[a] = hi[a] - lo[a]; vol *= e[a]; }
    const double h = std::cbrt(vol * per_bin / (double)std::max<size_t>(n, 1));
    for (int a = 0; a < 3; ++a)
        out[a] = std::min<int64_t>(16384, std::max<int64_t>(1, (int64_t)std::llround(e[a] / h)));
}

// Grid for uneven input. The density seen by the average generator is the mean over generators of
// the other generators per unit volume in its coarse bin, sum(c (c - 1)) / n / v over a histogram
// of about 8 generators per bin (at most 64 bins per axis): about n / volume for uniform input.
// Its ratio to n / volume (clamped to [1, kDensityCap]) refines the grid by ratio^(2/3): dense
// clusters get finer bins without pushing a uniform background into mostly empty ones. When more
// than kVoidFrac of the coarse bins are empty, cells next to the empty space reach across it and
// the walk over its bins grows with the cube of the bin count per axis, so such input gets a
// coarse grid instead, kVoidScale times the target. A Lloyd job is the exception: its first
// relaxation already moves the generators into the empty space, and its later rounds, most of its
// work, run fastest at the plain target (a 1M-point, 8-cluster job: 20 rounds in 6.1 s at the
// target against 10.0 s at three times it on MI355X).
constexpr double kDensityCap = 64.0;
constexpr double kVoidFrac = 0.25;
constexpr double kVoidScale = 3.0;

struct Density {
    double ratio;   // clamped to [1, kDensityCap]
    bool voids;     // more than kVoidFrac of the coarse bins are empty
};

Density density_of(const double lo[3], const double hi[3], const double *xyz, size_t n) {
    if (n < 64) return {1.0, false};
    int64_t cg[3];
    grid_for(lo, hi, n, 8.0, cg);
    double e[3];
    for (int a = 0; a < 3; ++a) { cg[a] = std::min<int64_t>(cg[a], 64); e[a] = hi[a] - lo[a]; }
    std::vector<uint32_t> cnt((size_t)(cg[0] * cg[1] * cg[2]), 0u);
    for (size_t i = 0; i < n; ++i) {
        int64_t b[3];
        for (int a = 0; a < 3; ++a) {
            const double u = (xyz[3 * i + a] - lo[a]) / e[a] * (double)cg[a];
            b[a] = u >= 0.0 ? std::min<int64_t>((int64_t)u, cg[a] - 1) : 0;
        }
        cnt[(size_t)((b[2] * cg[1] + b[1]) * cg[0] + b[0])] += 1u;
    }
    double s2 = 0.0;
    size_t empty = 0;
    for (uint32_t c : cnt) {
        s2 += (double)c * (double)(c > 0 ? c - 1 : 0);
        empty += c == 0;
    }
    // (s2 / n / v_bin) / (n / volume) with v_bin = volume / bins
    const double ratio = s2 * (double)cnt.size() / ((double)n * (double)n);
    return {std::min(kDensityCap, std::max(1.0, ratio)), (double)empty > kVoidFrac * (double)cnt.size()};
}

void density_grid(const double lo[3], const double hi[3], const double *xyz, size_t n, double per_bin,
                  bool lloyd, int64_t out[3]) {
    const Density d = density_of(lo, hi, xyz, n);
    const double target = d.voids ? (lloyd ? per_bin : per_bin * kVoidScale) : per_bin / std::pow(d.ratio, 2.0 / 3.0);
    grid_for(lo, hi, n, target, out);
}

// Normalises an axis as the geometry constructors do (wall_3d.rs:23-27, 124-128, 194-198, 285-289).
void unit_axis(const double v[3], double out[3]) {
    const double len = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (len == 0.0) { out[0] = 0.0; out[1] = 0.0; out[2] = 1.0; }
    else { out[0] = v[0] / len; out[1] = v[1] / len; out[2] = v[2] / len; }
}

// vd_wall -> device wall; the polyhedron planes or tube samples go to data (copied).
int make_wall(const vd_wall &w, vd::WallDev &out, std::vector<double> &data) {
    if (w.id > VD_WALL_ID_MAX) return fail(VD_EINVAL, "Wall ID must be <= -1000");   // wall/mod.rs:25-27
    out = vd::WallDev{};
    data.clear();
    out.id = w.id;
    out.kind = w.kind;
    for (int k = 0; k < 3; ++k) out.a[k] = w.a[k];
    switch (w.kind) {
    case VD_WALL_PLANE:     // PlaneGeometry::new, wall_3d.rs:23-27
    case VD_WALL_CYLINDER:  // CylinderGeometry::new, :124-128
        unit_axis(w.b, out.b);
        out.r = w.r;
        break;
    case VD_WALL_SPHERE:
        out.r = w.r;
        break;
    case VD_WALL_CONE:      // ConeGeometry::new :194-198; tan (:212), cos and sin (:232-233) of the angle
        unit_axis(w.b, out.b);
        out.r = w.r;
        out.cs = std::cos(w.r);
        out.sn = std::sin(w.r);
        out.tn = std::tan(w.r);
        break;
    case VD_WALL_TORUS:     // TorusGeometry::new, :285-289
        unit_axis(w.b, out.b);
        out.r = w.r;
        out.r2 = w.r2;
        break;
    case VD_WALL_POLYHEDRON:
    case VD_WALL_TUBE: {
        const size_t per = w.kind == VD_WALL_POLYHEDRON ? 6 : 3;
        if (w.npts > VD_WALL_MAX_POINTS) return fail(VD_EINVAL, "too many wall planes or samples");
        if (w.npts && !w.pts) return fail(VD_EINVAL, "NULL wall points");
        data.assign(w.pts, w.pts + per * w.npts);
        out.npts = w.npts;
        out.r = w.r;
        out.closed = w.closed ? 1 : 0;
        break;
    }
    default:
        return fail(VD_EINVAL, "unknown wall kind");
    }
    return VD_OK;
}

// ------------------------------------------------------------------------------------------
// Engine
// ------------------------------------------------------------------------------------------
struct Member {
    double lo[3], hi[3];
    int64_t n[3];
    bool auto_grid = false;
    vd::GridGeom grid{};
    std::vector<vd::WallDev> walls;
    std::vector<std::vector<double>> wdata;   // per wall: polyhedron planes or tube samples
    std::vector<vd_wall_host> host;           // the caller's walls (kind 7), by their npts index
    int shape = -1;
    uint64_t raw_off = 0, nraw = 0, gen_off = 0, ngen = 0, bin_off = 0, nbins = 0;
};

// A shell order prefix resident in device memory, one per distinct grid geometry. Its radius
// parameter k starts small and doubles only for geometries whose cells walked past the prefix.
struct ShellTable {
    DevBuf buf;                 // ShellCode[count]
    vd::ShellPack pack{};
    uint32_t k = 0;
    uint32_t count = 0;
    double next = std::numeric_limits<double>::infinity();
    bool complete = false;
    uint64_t epoch = 0;         // last engine pass that used it
    double far_frac = 0.0;      // share of a launch's cells that walked past the prefix when it grew
    uint32_t count0 = 0;        // the first prefix (radius kShellK0): its entries lead every longer one
    double next0 = std::numeric_limits<double>::infinity();
};

// Shares of a launch's cells walking past the shell prefix between which a geometry counts as one
// with empty regions between clusters: a few cells per thousand cross them. Later first launches on
// such a geometry see only its first prefix (radius kShellK0), so those cells stop early instead
// of walking the long prefix in tier 0, which they outgrow, and rerun in the skip-ahead
// instantiation of tier 1 on the whole prefix. A larger share means many ordinary cells in fine
// bins (a uniform background under dense clusters) whose walks are short: they keep the whole
// prefix from the start.
constexpr double kFarFracMin = 1.0 / 1024.0;
constexpr double kFarFracMax = 1.0 / 16.0;
constexpr int kFarTier = 1;   // lowest tier in which cells that walked past their prefix rerun

bool far_geometry(const ShellTable &t) { return t.far_frac > kFarFracMin && t.far_frac <= kFarFracMax; }

using GeomKey = std::tuple<int64_t, int64_t, int64_t, uint64_t, uint64_t, uint64_t>;

struct Shape {   // one distinct grid geometry of the current members
    int64_t n[3];
    double cs[3];
    ShellTable *tab = nullptr;
    bool trunc = false;   // the uploaded descriptors show only its first prefix
};

constexpr uint32_t kShellK0 = 7;                       // initial prefix radius (bins)
constexpr size_t kShellCacheBytes = size_t(4) << 30;   // cached tables beyond this are evicted (LRU)

// Prefix radius k (bins) for the next extension of shape s: twice the current one, or enough for
// keys up to 4 r2 (the prefix holds every key up to ((k - 1) max cell size)^2), whichever is
// larger, and no larger than the complete prefix (one spanning the grid's diagonal). The direct
// jump is limited to an enumeration box of kShellJumpBox offsets: a cell that runs out of prefix
// before its first cut still has the box-sized radius, which would ask for the whole grid.
constexpr uint64_t kShellJumpBox = uint64_t(1) << 26;

uint64_t shell_box_size(const Shape &s, uint32_t k) {
    uint64_t w = 1;
    for (int a = 0; a < 3; ++a) w *= (uint64_t)(2 * std::min<int64_t>(k, s.n[a]) + 1);
    return w;
}

uint32_t shell_k_for(const Shape &s, double r2) {
    const double csm = std::max(s.cs[0], std::max(s.cs[1], s.cs[2]));
    const int64_t nmax = std::max(s.n[0], std::max(s.n[1], s.n[2]));
    const uint32_t twice = 2 * s.tab->k;
    const uint32_t full = (uint32_t)std::ceil(std::sqrt(3.0) * (double)nmax) + 3;
    double want = (double)twice;
    if (r2 > 0.0 && csm > 0.0) want = std::max(want, std::floor(2.0 * std::sqrt(r2) / csm) + 3.0);
    uint32_t k = (uint32_t)std::min<double>(want, (double)full);
    while (k > twice && shell_box_size(s, k) > kShellJumpBox) --k;
    return std::max(k, std::min(twice, full));
}

struct Engine {
    std::vector<Member> mem;
    std::vector<Shape> shapes;
    std::map<GeomKey, ShellTable> tables;   // device shell prefixes by grid geometry
    uint64_t epoch = 0;
    int device = 0, cus = 1;
    int occ[kTiers] = {1, 1, 1, 1, 1};
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;     // side stream for far-cell retries (shared: engine_stream2)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int keep_cells = 1;

    DevBuf d_desc, d_offs, d_raw, d_raw_tess, d_flags, d_pos, d_scan, d_gat;
    DevBuf d_gen, d_gen_tess, d_gen_bin, d_counts, d_bin_start, d_fill, d_items, d_spos, d_shell_tmp, d_mflags,
        d_mrmax;
    DevBuf d_vol, d_cen, d_nface, d_nvert, d_status, d_cell_off;
    DevBuf d_vtx, d_fcnt, d_fnb, d_fidx, d_top, d_queue, d_counts4, d_next;
    DevBuf d_out0, d_out1;             // first-round retry lists of tier 0
    DevBuf d_wdata;                    // wall planes and samples of every member
    DevBuf d_tube;                     // per generator: planes of its member's tube walls
    // post-passes: cell topology of the previous pass, seal log, prune log
    int seal = 0;                      // cell launches run the seal mode
    bool trunc0 = false;               // the first launch sees only the first prefix of far geometries
    int reserve_cus = 0;               // CUs the persistent cell launches leave to other engines
    // far-cell reruns on a side stream beside the rest of their round, and over per-cell lists
    // (both off in the server's lanes: with two lanes sharing the GPU each cost more than it gained
    // on the served mix, measured)
    bool side_streams = true;
    double far_list_min = kFarListMinDefault;   // walk length for a per-cell list (0: never)
    bool keep_nb = true;               // keep every cell's neighbour ids (reduced mode)
    size_t shell_cache_bytes = kShellCacheBytes;   // shell tables of other geometries kept up to this
    int global_slots = kGlobalSlots;   // polytope slots (waves) of the global-memory tier
    // vd_calculate_static: buffers carved from the static scratch, the caller's stream (never
    // destroyed), no events, no pinned memory, no graphs, no side streams
    bool static_mode = false;
    uint32_t h_local[64] = {};         // the outcome counts' host copy in static mode
    uint64_t arena_hint = 40;          // initial arena vertices per cell (VD_OPT_ARENA)
    // launch shape of the cell kernels (VD_OPT_T0_WAVES, VD_OPT_QUEUE_GRAIN, VD_OPT_FIRST_TIER,
    // VD_OPT_FIRST_PROBE): the tier-0 instantiation by resident waves per SIMD, list entries per
    // queue atomic, the tier the first round runs every cell in, and whether that launch uses the
    // skip-ahead instantiation. Any choice gives the same results.
    int t0_waves = vd::kTier0WavesDefault;
    uint32_t qgrain = 1;
    int first_tier = 0;
    int first_probe = 0;               // the first launch skips inert shell batches (VD_OPT_FIRST_PROBE)
    int max_attempts = 8;              // launches with doubled arenas before giving up (VD_OPT_ATTEMPTS)
    uint64_t cap_cells = 0;            // cells the arena capacities were sized for
    DevBuf d_smask, d_scnt, d_soff, d_snb, d_slog, d_slogcnt, d_pcnt, d_poff, d_pi, d_pp;
    uint32_t ntube = 0;                // tube-wall slots per generator
    // the caller's walls (vd_tess_add_wall_planes_per_generator): their contains verdicts per raw
    // point and cut planes per generator, evaluated on the host at every set_generators
    uint32_t nhost = 0;                // such walls per generator (slots of d_host_off)
    DevBuf d_hflags, d_host_off, d_host_planes;
    DevBuf d_gmem;                     // polytope slots of the global-memory tier
    std::vector<uint32_t> host_max;    // per such wall: most planes it emits for one generator
    std::vector<uint64_t> h_host_off;  // host copies of d_host_off and d_host_planes
    std::vector<double> h_host_planes;
    uint32_t seal_slots = vd::kMaxWalls;   // seal-log entries per neighbour
    uint32_t *h_counts = nullptr;      // pinned: per-tier outcome counts of a round
    uint64_t nraw = 0, ngen = 0, nbins = 0;
    bool prepared = false, calculated = false;
    bool binned = false;               // bins of the current generators are built
    // set_generator (tessellation.rs:71-92) on a single tessellation: its bins as update_generator
    // leaves them (algo_3d_grid.rs:132-143: swap_remove from the old bin, push to the new one), not
    // in the sorted order a binning gives. Kept on the host: a CSR (h_bin_start, h_items) with the
    // bins moves have changed as lists beside it (h_over), and the positions (h_gen); a move touches
    // only those, and the next calculate merges the lists into the CSR and uploads bins and
    // positions once (bins_dirty, pos_dirty). set_generators and relax end it (they bin afresh), as
    // does an add_wall that removes generators; one that removes none keeps it, as the reference
    // does.
    bool custom_bins = false, bins_dirty = false, pos_dirty = false;
    std::vector<uint32_t> h_bin_start, h_items, h_gen_bin;
    std::unordered_map<uint32_t, std::vector<uint32_t>> h_over;
    std::vector<double> h_gen;
    // First-round launch sequence (binning, tier-0 launch, collect, counts readback) of a shape
    // seen twice in a row is captured into a graph and replayed while the shape repeats.
    int graphs = 1;
    std::vector<uint64_t> last_key, gkey;
    hipGraphExec_t gexec = nullptr;
    uint64_t graph_launches = 0, graph_captures = 0;
    uint64_t cap[3] = {0, 0, 0};
    double stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    std::vector<vd::TessDesc> desc;
};

int engine_init(Engine &E) {
    hipError_t e = hipGetDevice(&E.device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&E.cus, hipDeviceAttributeMultiprocessorCount, E.device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&E.stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&E.ev0);
    if (e == hipSuccess) e = hipEventCreate(&E.ev1);
    for (int k = 0; k < kTiers && e == hipSuccess; ++k) e = vd::cells_occupancy(k, &E.occ[k], E.t0_waves);
    if (e == hipSuccess) e = hipHostMalloc((void **)&E.h_counts, 64 * sizeof(uint32_t), hipHostMallocDefault);
    if (e != hipSuccess) return fail(VD_EDEVICE, std::string("device setup: ") + hipGetErrorString(e));
    for (int k = 0; k < kTiers; ++k) E.occ[k] = std::max(1, E.occ[k]);
    return VD_OK;
}

// Releases every device buffer of the engine (its shell tables and graph too); the engine stays
// usable: the next set_generators or calculate allocates again. The caller has synchronised.
void engine_release_buffers(Engine &E) {
    DevBuf *bufs[] = {&E.d_desc, &E.d_offs, &E.d_raw, &E.d_raw_tess, &E.d_flags, &E.d_pos, &E.d_scan,
                      &E.d_gat, &E.d_gen, &E.d_gen_tess, &E.d_gen_bin, &E.d_counts, &E.d_bin_start,
                      &E.d_fill, &E.d_items, &E.d_spos, &E.d_shell_tmp, &E.d_mflags, &E.d_mrmax, &E.d_vol,
                      &E.d_cen, &E.d_nface, &E.d_nvert, &E.d_status, &E.d_cell_off, &E.d_vtx, &E.d_fcnt,
                      &E.d_fnb, &E.d_fidx,
                      &E.d_top, &E.d_queue, &E.d_counts4, &E.d_next};
    for (DevBuf *b : bufs) release(*b);
    release(E.d_out0);
    release(E.d_out1);
    release(E.d_wdata);
    release(E.d_tube);
    release(E.d_hflags);
    release(E.d_gmem);
    release(E.d_host_off);
    release(E.d_host_planes);
    DevBuf *post[] = {&E.d_smask, &E.d_scnt, &E.d_soff, &E.d_snb, &E.d_slog, &E.d_slogcnt, &E.d_pcnt, &E.d_poff,
                      &E.d_pi, &E.d_pp};
    for (DevBuf *b : post) release(*b);
    if (E.gexec) (void)hipGraphExecDestroy(E.gexec);
    E.gexec = nullptr;
    E.last_key.clear();
    E.gkey.clear();
    for (auto &kv : E.tables) release(kv.second.buf);
    E.tables.clear();
    E.shapes.clear();
    E.cap_cells = 0;
    E.prepared = E.calculated = E.binned = false;
}

void engine_free(Engine &E) {
    // the side stream may hold this engine's far-cell launches when a calculate failed before
    // its stream waited for them
    if (E.stream2) (void)hipStreamSynchronize(E.stream2);
    if (E.stream) (void)hipStreamSynchronize(E.stream);
    engine_release_buffers(E);
    if (E.h_counts && !E.static_mode) (void)hipHostFree(E.h_counts);
    E.h_counts = nullptr;
    if (E.ev0) (void)hipEventDestroy(E.ev0);
    if (E.ev1) (void)hipEventDestroy(E.ev1);
    if (E.stream && !E.static_mode) (void)hipStreamDestroy(E.stream);
    E.ev0 = E.ev1 = nullptr;
    E.stream = E.stream2 = nullptr;
}

// Side streams for the retry launches of cells crossing empty space: kSideStreams per device for
// the whole process, handed to engines in turn at their first such retry and never destroyed. Each
// engine owning a second stream would double the streams of a caller that holds thousands of
// engines (one per small tessellation), and every stream holds runtime queues and memory. Engines
// that share a side stream stay correct: each orders its own launches on it with its own events;
// only their far-cell retries queue behind each other.
constexpr int kSideStreams = 4;
std::mutex g_side_mu;
std::map<int, std::vector<hipStream_t>> g_side;
unsigned g_side_next = 0;

int engine_stream2(Engine &E) {
    if (E.stream2) return VD_OK;
    std::lock_guard<std::mutex> lock(g_side_mu);
    std::vector<hipStream_t> &pool = g_side[E.device];
    if (pool.size() < (size_t)kSideStreams) {
        hipStream_t st = nullptr;
        const hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) return fail(VD_EDEVICE, std::string("side stream: ") + hipGetErrorString(e));
        pool.push_back(st);
        E.stream2 = st;
    } else {
        E.stream2 = pool[g_side_next++ % pool.size()];
    }
    return VD_OK;
}

int scan(Engine &E, const uint32_t *in, uint32_t *out, uint32_t n) {
    const size_t words = vd::scan_tmp_words(n);
    VD_TRY(ensure(E.d_scan, words * 4));
    VD_HIP(vd::launch_scan_u32(in, out, n, E.d_scan.as<uint32_t>(), words, E.stream));
    return VD_OK;
}

// Builds the prefix with radius parameter k of one geometry on the device (algo_3d_grid.rs:47-95).
int bits_for(uint64_t v) {   // bits to hold 0 .. v - 1
    int b = 0;
    while (b < 32 && (1ull << b) < v) ++b;
    return b;
}

int build_table(Engine &E, const int64_t n[3], const double cs[3], uint32_t k, ShellTable &t) {
    const ShellBox b = shell_box(n, cs, k);
    const uint64_t w = (uint64_t)(2 * b.K[0] + 1) * (uint64_t)(2 * b.K[1] + 1) * (uint64_t)(2 * b.K[2] + 1);
    const int bx = bits_for(2 * b.K[0] + 1), by = bits_for(2 * b.K[1] + 1), bz = bits_for(2 * b.K[2] + 1);
    if (w >= 0x7fffffffull || bx + by + bz > 32)
        return fail(VD_ECAPACITY, "shell order prefix too large (grid too sparse for the walk radius)");
    vd::ShellPack pack{};
    for (int a = 0; a < 3; ++a) pack.k[a] = (int16_t)b.K[a];
    pack.sy = (uint8_t)bx;
    pack.sz = (uint8_t)(bx + by);
    const size_t sb = vd::shell_scratch_bytes((uint32_t)w);
    VD_TRY(ensure(E.d_shell_tmp, sb));
    VD_HIP(vd::shell_count(pack, cs, b.T, (uint32_t)w, E.d_shell_tmp.p, E.stream));
    uint32_t count = 0;
    VD_HIP(hipMemcpyAsync(&count, vd::shell_counter(E.d_shell_tmp.p, (uint32_t)w), 4, hipMemcpyDeviceToHost,
                          E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    release(t.buf);
    const size_t body = ((size_t)std::max<uint32_t>(count, 1) * sizeof(vd::ShellCode) + 15) & ~(size_t)15;
    VD_TRY(ensure(t.buf, body + 16));
    uint64_t *d_next = reinterpret_cast<uint64_t *>(t.buf.as<char>() + body);
    VD_HIP(vd::shell_finish(pack, cs, (uint32_t)w, count, E.d_shell_tmp.p, sb, t.buf.as<vd::ShellCode>(), d_next,
                            E.stream));
    uint64_t nb = 0;
    VD_HIP(hipMemcpyAsync(&nb, d_next, 8, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    double next;
    std::memcpy(&next, &nb, 8);
    t.pack = pack;
    t.k = k;
    t.count = count;
    t.next = std::min(next, b.outside);
    t.complete = std::isinf(t.next);
    if (k == kShellK0) { t.count0 = t.count; t.next0 = t.next; }
    return VD_OK;
}

// Groups the members by grid geometry and makes sure every geometry has a device prefix.
int engine_shells(Engine &E) {
    E.shapes.clear();
    E.epoch += 1;
    for (Member &m : E.mem) {
        int found = -1;
        for (size_t k = 0; k < E.shapes.size(); ++k) {
            const Shape &s = E.shapes[k];
            if (s.n[0] == m.n[0] && s.n[1] == m.n[1] && s.n[2] == m.n[2] &&
                std::memcmp(s.cs, m.grid.cs, sizeof(s.cs)) == 0) { found = (int)k; break; }
        }
        if (found < 0) {
            Shape s;
            for (int a = 0; a < 3; ++a) { s.n[a] = m.n[a]; s.cs[a] = m.grid.cs[a]; }
            E.shapes.push_back(s);
            found = (int)E.shapes.size() - 1;
        }
        m.shape = found;
    }
    for (Shape &s : E.shapes) {
        uint64_t c[3];
        std::memcpy(c, s.cs, sizeof(c));
        ShellTable &t = E.tables[GeomKey{s.n[0], s.n[1], s.n[2], c[0], c[1], c[2]}];
        t.epoch = E.epoch;
        if (!t.k) VD_TRY(build_table(E, s.n, s.cs, kShellK0, t));
        s.tab = &t;
    }
    // least recently used tables of other geometries go first when the cache grows too large
    size_t bytes = 0;
    for (auto &kv : E.tables) bytes += kv.second.buf.bytes;
    while (bytes > E.shell_cache_bytes) {
        auto victim = E.tables.end();
        for (auto it = E.tables.begin(); it != E.tables.end(); ++it)
            if (it->second.epoch != E.epoch && (victim == E.tables.end() || it->second.epoch < victim->second.epoch))
                victim = it;
        if (victim == E.tables.end()) break;
        bytes -= victim->second.buf.bytes;
        release(victim->second.buf);
        E.tables.erase(victim);
    }
    return VD_OK;
}

int upload_desc(Engine &E) {
    const size_t M = E.mem.size();
    if (E.shapes.empty() && M) VD_TRY(engine_shells(E));
    // polyhedron planes and tube samples of every member's walls, in one device array
    size_t words = 0;
    for (const Member &m : E.mem)
        for (const std::vector<double> &d : m.wdata) words += d.size();
    if (words) {
        std::vector<double> all;
        all.reserve(words);
        for (const Member &m : E.mem)
            for (const std::vector<double> &d : m.wdata) all.insert(all.end(), d.begin(), d.end());
        VD_TRY(ensure(E.d_wdata, words * 8));
        VD_HIP(hipMemcpyAsync(E.d_wdata.p, all.data(), words * 8, hipMemcpyHostToDevice, E.stream));
        VD_HIP(hipStreamSynchronize(E.stream));
    }
    size_t woff = 0;
    E.ntube = 0;
    for (const Member &m : E.mem) {
        uint32_t nt = 0;
        for (const vd::WallDev &w : m.walls) nt += w.kind == VD_WALL_TUBE;
        E.ntube = std::max(E.ntube, nt);
    }
    E.desc.assign(M, vd::TessDesc{});
    for (size_t t = 0; t < M; ++t) {
        const Member &m = E.mem[t];
        vd::TessDesc &d = E.desc[t];
        d.grid = m.grid;
        for (int a = 0; a < 3; ++a) { d.box_lo[a] = m.lo[a]; d.box_hi[a] = m.hi[a]; }
        d.raw_off = m.raw_off;
        d.gen_off = m.gen_off;
        d.bin_off = m.bin_off;
        d.nraw = (uint32_t)m.nraw;
        d.ngen = (uint32_t)m.ngen;
        Shape &sh = E.shapes[m.shape];
        const ShellTable &tab = *sh.tab;
        sh.trunc = E.trunc0 && far_geometry(tab) && tab.k > kShellK0;
        d.shell = tab.buf.as<vd::ShellCode>();
        d.pack = tab.pack;
        d.nshell = sh.trunc ? tab.count0 : tab.count;
        d.shell_next_d2 = sh.trunc ? tab.next0 : tab.next;
        d.nwalls = (int32_t)m.walls.size();
        for (size_t k = 0; k < m.walls.size(); ++k) {
            d.walls[k] = m.walls[k];
            d.walls[k].pts = m.wdata[k].empty() ? nullptr : E.d_wdata.as<double>() + woff;
            woff += m.wdata[k].size();
        }
    }
    VD_TRY(ensure(E.d_desc, std::max<size_t>(M, 1) * sizeof(vd::TessDesc)));
    if (M)
        VD_HIP(hipMemcpyAsync(E.d_desc.p, E.desc.data(), M * sizeof(vd::TessDesc), hipMemcpyHostToDevice,
                              E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    return VD_OK;
}

// Member layout of the raw points (raw_off from nraw) and of the bins (bin_off from the grids).
void engine_layout(Engine &E) {
    uint64_t r = 0, b = 0;
    for (Member &m : E.mem) {
        m.raw_off = r;
        r += m.nraw;
        m.bin_off = b;
        m.nbins = (uint64_t)m.n[0] * m.n[1] * m.n[2];
        b += m.nbins;
    }
    E.nraw = r;
    E.nbins = b;
}

// The cut planes of the caller's walls for every generator (WallGeometry::cut, src/wall/mod.rs:38-40),
// evaluated on the host: generator g's planes of its member's k-th such wall are
// d_host_planes[6 q], d_host_off[g * nhost + k] <= q < d_host_off[g * nhost + k + 1].
int host_planes(Engine &E) {
    const uint64_t ng = E.ngen, nh = E.nhost;
    std::vector<double> gen(3 * ng);
    if (ng) VD_HIP(hipMemcpyAsync(gen.data(), E.d_gen.p, ng * 24, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    std::vector<uint64_t> &off = E.h_host_off;
    std::vector<double> &planes = E.h_host_planes;
    off.assign(ng * nh + 1, 0);
    planes.clear();
    std::vector<double> buf(6 * 16);
    E.host_max.assign(nh, 0u);
    for (const Member &m : E.mem) {
        for (uint64_t g = m.gen_off; g < m.gen_off + m.ngen; ++g) {
            for (uint64_t k = 0; k < nh; ++k) {
                if (k < m.host.size()) {
                    const vd_wall_host &cb = m.host[k];
                    int c = cb.cut(cb.ctx, &gen[3 * g], buf.data(), (int)(buf.size() / 6));
                    if (c > (int)(buf.size() / 6)) {
                        if (c > VD_WALL_MAX_POINTS) return fail(VD_EINVAL, "a wall's cut emitted too many planes");
                        buf.resize(6 * (size_t)c);
                        c = cb.cut(cb.ctx, &gen[3 * g], buf.data(), c);
                    }
                    if (c < 0 || c > (int)(buf.size() / 6)) return fail(VD_EINVAL, "a wall's cut returned a bad count");
                    planes.insert(planes.end(), buf.begin(), buf.begin() + 6 * (size_t)c);
                    E.host_max[k] = std::max<uint32_t>(E.host_max[k], (uint32_t)c);
                }
                off[g * nh + k + 1] = planes.size() / 6;
            }
        }
    }
    VD_TRY(ensure(E.d_host_off, off.size() * 8));
    VD_TRY(ensure(E.d_host_planes, std::max<size_t>(planes.size(), 6) * 8));
    VD_HIP(hipMemcpyAsync(E.d_host_off.p, off.data(), off.size() * 8, hipMemcpyHostToDevice, E.stream));
    if (!planes.empty())
        VD_HIP(hipMemcpyAsync(E.d_host_planes.p, planes.data(), planes.size() * 8, hipMemcpyHostToDevice,
                              E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    return VD_OK;
}

// The caller's walls' cut planes for generator g of a single tessellation, now at q: written over
// its old planes when every wall emits as many as before (one small upload), else spliced into the
// CSR, which is uploaded whole.
int host_planes_one(Engine &E, uint64_t g, const double q[3]) {
    const uint64_t nh = E.nhost;
    const Member &m = E.mem[0];
    std::vector<uint64_t> &off = E.h_host_off;
    std::vector<double> &planes = E.h_host_planes;
    std::vector<double> mine;
    std::vector<uint64_t> cnt(nh, 0);
    std::vector<double> buf(6 * 16);
    for (uint64_t k = 0; k < nh && k < m.host.size(); ++k) {
        const vd_wall_host &cb = m.host[k];
        int c = cb.cut(cb.ctx, q, buf.data(), (int)(buf.size() / 6));
        if (c > (int)(buf.size() / 6)) {
            if (c > VD_WALL_MAX_POINTS) return fail(VD_EINVAL, "a wall's cut emitted too many planes");
            buf.resize(6 * (size_t)c);
            c = cb.cut(cb.ctx, q, buf.data(), c);
        }
        if (c < 0 || c > (int)(buf.size() / 6)) return fail(VD_EINVAL, "a wall's cut returned a bad count");
        mine.insert(mine.end(), buf.begin(), buf.begin() + 6 * (size_t)c);
        cnt[k] = (uint64_t)c;
        E.host_max[k] = std::max<uint32_t>(E.host_max[k], (uint32_t)c);
    }
    const uint64_t a = off[g * nh], b = off[g * nh + nh];
    bool same = true;
    for (uint64_t k = 0; k < nh; ++k) same &= off[g * nh + k + 1] - off[g * nh + k] == cnt[k];
    if (same) {
        if (b > a) {
            std::copy(mine.begin(), mine.end(), planes.begin() + 6 * a);
            VD_HIP(hipMemcpyAsync(E.d_host_planes.as<double>() + 6 * a, mine.data(), mine.size() * 8,
                                  hipMemcpyHostToDevice, E.stream));
            VD_HIP(hipStreamSynchronize(E.stream));
        }
        return VD_OK;
    }
    planes.erase(planes.begin() + 6 * a, planes.begin() + 6 * b);
    planes.insert(planes.begin() + 6 * a, mine.begin(), mine.end());
    const int64_t shift = (int64_t)(mine.size() / 6) - (int64_t)(b - a);
    for (uint64_t k = 0; k < nh; ++k) off[g * nh + k + 1] = off[g * nh + k] + cnt[k];
    for (uint64_t j = g * nh + nh + 1; j < off.size(); ++j) off[j] = (uint64_t)((int64_t)off[j] + shift);
    VD_TRY(ensure(E.d_host_planes, std::max<size_t>(planes.size(), 6) * 8));
    VD_HIP(hipMemcpyAsync(E.d_host_off.p, off.data(), off.size() * 8, hipMemcpyHostToDevice, E.stream));
    if (!planes.empty())
        VD_HIP(hipMemcpyAsync(E.d_host_planes.p, planes.data(), planes.size() * 8, hipMemcpyHostToDevice,
                              E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    return VD_OK;
}

// The wall filter of set_generators for every member from E.d_raw (tessellation.rs:43-67): the
// points every wall of their member contains, compacted in input order into E.d_gen.
int engine_filter(Engine &E) {
    E.prepared = E.calculated = false;
    E.custom_bins = E.bins_dirty = E.pos_dirty = false;
    const size_t M = E.mem.size();
    const uint64_t n = E.nraw;
    if (E.nbins + 1 >= 0xffffffffull || n >= 0xffffffffull) return fail(VD_EINVAL, "too many bins or points");
    if (E.shapes.empty()) VD_TRY(engine_shells(E));
    // member of every raw point
    std::vector<uint64_t> offs(M + 1);
    for (size_t t = 0; t < M; ++t) offs[t] = E.mem[t].raw_off;
    offs[M] = n;
    VD_TRY(ensure(E.d_offs, (M + 1) * 8));
    VD_HIP(hipMemcpyAsync(E.d_offs.p, offs.data(), (M + 1) * 8, hipMemcpyHostToDevice, E.stream));
    VD_TRY(ensure(E.d_raw_tess, (n + 1) * 4));
    VD_HIP(vd::launch_member_of(E.d_offs.as<uint64_t>(), (uint32_t)M, E.d_raw_tess.as<uint32_t>(), n, E.stream));
    bool walls = false;
    for (const Member &m : E.mem) walls |= !m.walls.empty();
    // the caller's walls: contains on the host for every raw point, in wall order
    E.nhost = 0;
    for (const Member &m : E.mem) E.nhost = std::max<uint32_t>(E.nhost, (uint32_t)m.host.size());
    const uint32_t *hflags = nullptr;
    if (E.nhost && n) {
        std::vector<double> raw(3 * n);
        VD_HIP(hipMemcpyAsync(raw.data(), E.d_raw.p, n * 24, hipMemcpyDeviceToHost, E.stream));
        VD_HIP(hipStreamSynchronize(E.stream));
        std::vector<uint32_t> hf(n + 1, 1u);
        for (const Member &m : E.mem) {
            if (m.host.empty()) continue;
            for (uint64_t g = m.raw_off; g < m.raw_off + m.nraw; ++g)
                for (const vd::WallDev &w : m.walls)
                    if (w.kind == 7) {
                        const vd_wall_host &cb = m.host[w.npts];
                        if (!cb.contains(cb.ctx, &raw[3 * g])) { hf[g] = 0u; break; }
                    }
        }
        VD_TRY(ensure(E.d_hflags, (n + 1) * 4));
        VD_HIP(hipMemcpyAsync(E.d_hflags.p, hf.data(), (n + 1) * 4, hipMemcpyHostToDevice, E.stream));
        hflags = E.d_hflags.as<uint32_t>();
    }
    // descriptors with the raw layout (the wall filter reads the walls)
    for (Member &m : E.mem) { m.gen_off = m.raw_off; m.ngen = m.nraw; }
    VD_TRY(upload_desc(E));
    VD_TRY(ensure(E.d_gen, (n + 1) * 24));
    VD_TRY(ensure(E.d_gen_tess, (n + 1) * 4));
    if (!walls) {
        if (n) {
            VD_HIP(hipMemcpyAsync(E.d_gen.p, E.d_raw.p, n * 24, hipMemcpyDeviceToDevice, E.stream));
            VD_HIP(hipMemcpyAsync(E.d_gen_tess.p, E.d_raw_tess.p, n * 4, hipMemcpyDeviceToDevice, E.stream));
        }
        E.ngen = n;
    } else {
        VD_TRY(ensure(E.d_flags, (n + 1) * 4));
        VD_TRY(ensure(E.d_pos, (n + 1) * 4));
        VD_HIP(vd::launch_fill_u32(E.d_flags.as<uint32_t>(), 0u, n + 1, E.stream));
        VD_HIP(vd::launch_wall_flags(E.d_raw.as<double>(), n, E.d_desc.as<vd::TessDesc>(),
                                     E.d_raw_tess.as<uint32_t>(), hflags, E.d_flags.as<uint32_t>(), E.stream));
        VD_TRY(scan(E, E.d_flags.as<uint32_t>(), E.d_pos.as<uint32_t>(), (uint32_t)(n + 1)));
        VD_HIP(vd::launch_compact(E.d_raw.as<double>(), E.d_raw_tess.as<uint32_t>(), E.d_flags.as<uint32_t>(),
                                  E.d_pos.as<uint32_t>(), n, E.d_gen.as<double>(), E.d_gen_tess.as<uint32_t>(),
                                  E.stream));
        // kept count before every member boundary
        VD_TRY(ensure(E.d_gat, (M + 1) * 4));
        VD_HIP(vd::launch_gather_u32(E.d_pos.as<uint32_t>(), E.d_offs.as<uint64_t>(), (uint32_t)(M + 1),
                                     E.d_gat.as<uint32_t>(), E.stream));
        std::vector<uint32_t> kept(M + 1);
        VD_HIP(hipMemcpyAsync(kept.data(), E.d_gat.p, (M + 1) * 4, hipMemcpyDeviceToHost, E.stream));
        VD_HIP(hipStreamSynchronize(E.stream));
        for (size_t t = 0; t < M; ++t) {
            E.mem[t].gen_off = kept[t];
            E.mem[t].ngen = kept[t + 1] - kept[t];
        }
        E.ngen = kept[M];
        VD_TRY(upload_desc(E));
    }
    return VD_OK;
}

// set_generators for every member from E.d_raw (tessellation.rs:43-67): wall filter in input order,
// compaction, then the stable counting-sort binning (algo_3d_grid.rs:111-130).
int engine_prepare(Engine &E) {
    VD_TRY(engine_filter(E));
    // the caller's walls: cut planes of every kept generator, as CSR over (generator, wall)
    if (E.nhost) VD_TRY(host_planes(E));
    // buffers of the binning (it runs at the start of calculate, see enqueue_binning)
    const uint64_t nb = E.nbins, ng = E.ngen;
    VD_TRY(ensure(E.d_counts, (nb + 1) * 4));
    VD_TRY(ensure(E.d_bin_start, (nb + 1) * 4));
    VD_TRY(ensure(E.d_fill, (nb + 1) * 4));
    VD_TRY(ensure(E.d_gen_bin, (ng + 1) * 4));
    VD_TRY(ensure(E.d_items, (ng + 1) * 4));
    VD_TRY(ensure(E.d_spos, (ng + 1) * 32));
    VD_TRY(ensure(E.d_scan, vd::scan_tmp_words((uint32_t)(nb + 1)) * 4));
    if (E.ntube) VD_TRY(ensure(E.d_tube, (ng + 1) * E.ntube * 7 * sizeof(double)));
    VD_HIP(hipStreamSynchronize(E.stream));
    E.binned = false;
    E.prepared = true;
    return VD_OK;
}

// The stable counting-sort binning of all members at once (algo_3d_grid.rs:111-130). Launches
// only: no allocation and no host synchronisation, so it can be captured into a graph.
int enqueue_binning(Engine &E) {
    const uint64_t nb = E.nbins, ng = E.ngen;
    VD_HIP(vd::launch_fill_u32(E.d_counts.as<uint32_t>(), 0u, nb + 1, E.stream));
    VD_HIP(vd::launch_fill_u32(E.d_fill.as<uint32_t>(), 0u, nb + 1, E.stream));
    VD_HIP(vd::launch_bin_count(E.d_gen.as<double>(), ng, E.d_desc.as<vd::TessDesc>(), E.d_gen_tess.as<uint32_t>(),
                                E.d_gen_bin.as<uint32_t>(), E.d_counts.as<uint32_t>(), E.stream));
    VD_TRY(scan(E, E.d_counts.as<uint32_t>(), E.d_bin_start.as<uint32_t>(), (uint32_t)(nb + 1)));
    VD_HIP(vd::launch_bin_scatter(E.d_gen_bin.as<uint32_t>(), ng, E.d_bin_start.as<uint32_t>(),
                                  E.d_fill.as<uint32_t>(), E.d_items.as<uint32_t>(), E.stream));
    VD_HIP(vd::launch_bin_sort(E.d_bin_start.as<uint32_t>(), nb, E.d_items.as<uint32_t>(), E.stream));
    VD_HIP(vd::launch_fill_spos(E.d_gen.as<double>(), E.d_items.as<uint32_t>(), ng, E.d_desc.as<vd::TessDesc>(),
                                E.d_gen_tess.as<uint32_t>(), E.d_spos.as<double4>(), E.stream));
    if (E.ntube)
        VD_HIP(vd::launch_tube_planes(E.d_gen.as<double>(), ng, E.d_desc.as<vd::TessDesc>(),
                                      E.d_gen_tess.as<uint32_t>(), E.ntube, E.d_tube.as<double>(), E.stream));
    return VD_OK;
}

// One cell launch over d_list on stream st (the engine's stream by default), drawing cells from work
// counter `slot` of d_queue (launches that may run at the same time use different counters), with
// optional per-cell shell lists (far_lists) for the entries of d_list.
int run_cells(Engine &E, int tier, const uint32_t *d_list, uint32_t nlist, float *ms, bool probe = false,
              int slot = 0, hipStream_t st = nullptr, const vd::ShellCode *lshell = nullptr,
              const uint64_t *lshell_off = nullptr, const uint32_t *lshell_n = nullptr,
              const double *lshell_next = nullptr) {
    if (!st) st = E.stream;
    vd::CellArgs a{};
    a.desc = E.d_desc.as<vd::TessDesc>();
    a.gen_tess = E.d_gen_tess.as<uint32_t>();
    a.gen = E.d_gen.as<double>();
    a.spos = E.d_spos.as<double4>();
    a.bin_start = E.d_bin_start.as<uint32_t>();
    a.tube_planes = E.d_tube.as<double>();
    a.ntube = E.ntube;
    a.host_off = E.nhost ? E.d_host_off.as<uint64_t>() : nullptr;
    a.host_planes = E.nhost ? E.d_host_planes.as<double>() : nullptr;
    a.nhost = E.nhost;
    a.seal_slots = E.seal_slots;
    if (E.seal) {
        a.seal_mask = E.d_smask.as<uint32_t>();
        a.seal_off = E.d_soff.as<uint32_t>();
        a.seal_nb = E.d_snb.as<uint32_t>();
        a.seal_log = E.d_slog.as<int32_t>();
        a.seal_logcnt = E.d_slogcnt.as<uint32_t>();
    }
    a.cells = d_list;
    a.ncells = nlist;
    a.queue = E.d_queue.as<uint32_t>() + slot;
    a.qgrain = E.qgrain;
    a.lshell = lshell;
    a.lshell_off = lshell_off;
    a.lshell_n = lshell_n;
    a.lshell_next = lshell_next;
    a.vol = E.d_vol.as<double>();
    a.cen = E.d_cen.as<double>();
    a.nface = E.d_nface.as<uint32_t>();
    a.nvert = E.d_nvert.as<uint32_t>();
    a.status = E.d_status.as<int32_t>();
    a.vtx = E.keep_cells ? E.d_vtx.as<double>() : nullptr;
    a.fcnt = E.d_fcnt.as<uint8_t>();
    a.fnb = (E.keep_cells || E.keep_nb) ? E.d_fnb.as<int32_t>() : nullptr;
    a.fidx = E.d_fidx.as<uint16_t>();
    a.arena_top = E.d_top.as<uint64_t>();
    a.cell_off = E.d_cell_off.as<uint64_t>();
    VD_HIP(vd::launch_fill_u32(E.d_queue.as<uint32_t>() + slot, 0u, 1, st));
    int waves = std::max(1, E.cus - E.reserve_cus) * E.occ[tier];
    if (tier == kGlobalTier) {
        waves = std::min(waves, E.global_slots);
        VD_TRY(ensure(E.d_gmem, (size_t)E.global_slots * vd::tier_global_bytes()));
        a.gmem = E.d_gmem.p;
    }
    if ((uint32_t)waves > nlist) waves = (int)nlist;
    // arena chunks per wave: up to kArenaChunk[k] entries, at most an eighth of the capacity over
    // all waves (the unused tails of the last chunks are holes in the arena)
    for (int k = 0; k < 3; ++k) {
        a.arena_cap[k] = E.cap[k];
        const uint64_t share = E.cap[k] / (8 * (uint64_t)std::max(waves, 1));
        a.arena_chunk[k] = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(kArenaChunk[k], share));
    }
    if (ms) VD_HIP(hipEventRecord(E.ev0, st));
    VD_HIP(vd::launch_cells(tier, a, waves, st, E.seal != 0, probe, E.t0_waves));
    if (ms) {
        VD_HIP(hipEventRecord(E.ev1, st));
        VD_HIP(hipEventSynchronize(E.ev1));
        VD_HIP(hipEventElapsedTime(ms, E.ev0, E.ev1));
    }
    return VD_OK;
}

// Launches of the first round of a calculate: the binning (if the generators changed), the cell
// launch over every cell in the first tier (tier 0 unless VD_OPT_FIRST_TIER says otherwise), the
// collect of its retry lists and the readback of the outcome counts into pinned memory. Launches
// and copies only, so the sequence can be captured.
int enqueue_round0(Engine &E, uint32_t n, float *ms) {
    if (!E.binned) VD_TRY(enqueue_binning(E));
    VD_HIP(vd::launch_fill_u32(reinterpret_cast<uint32_t *>(E.d_top.p), 0u, 12, E.stream));
    VD_TRY(run_cells(E, E.first_tier, E.d_items.as<uint32_t>(), n, ms, E.first_probe != 0));
    VD_HIP(vd::launch_fill_u32(E.d_counts4.as<uint32_t>(), 0u, 4 * kTiers, E.stream));
    VD_HIP(vd::launch_collect(E.d_status.as<int32_t>(), E.d_items.as<uint32_t>(), n, E.d_out0.as<uint32_t>(),
                              E.d_out1.as<uint32_t>(), E.d_counts4.as<uint32_t>() + 4 * E.first_tier, E.stream));
    VD_HIP(hipMemcpyAsync(E.h_counts, E.d_counts4.p, 4 * kTiers * sizeof(uint32_t), hipMemcpyDeviceToHost,
                          E.stream));
    return VD_OK;
}

// Per-cell shell lists (k_far_lists) for the far cells of d_list[0, n) whose reruns walk about
// E.far_list_min shell entries or more, each over the prefix entries its rerun can visit (key <=
// 4 r^2 of its stop, with a margin for rounding). Built on E.stream and finished on return. use:
// whether any cell got one; lsh holds the codes, and per entry of d_list lof the offset, lnn the
// length (kNoList: none) and lnx the key after the last entry of its list. The caller releases
// the four buffers after the launch.
int far_lists(Engine &E, const uint32_t *d_list, uint32_t n, DevBuf &lsh, DevBuf &lof, DevBuf &lnn, DevBuf &lnx,
              bool &use) {
    use = false;
    if (E.far_list_min <= 0.0 || n == 0) return VD_OK;
    DevBuf tmp[9];
    const ScopedBufs keep_tmp{tmp, 9};   // released on every return (after the last sync)
    DevBuf &flags = tmp[0], &lim = tmp[1], &sel = tmp[2], &slim = tmp[3], &cpd = tmp[4], &counts = tmp[5],
           &offs = tmp[6], &gidx = tmp[7], &at = tmp[8];
    VD_TRY(ensure(flags, (size_t)n * 4));
    VD_TRY(ensure(lim, (size_t)n * 4));
    VD_TRY(ensure(lnx, (size_t)n * 8));
    const vd::TessDesc *desc = E.d_desc.as<vd::TessDesc>();
    const uint32_t *gt = E.d_gen_tess.as<uint32_t>();
    VD_HIP(vd::launch_far_pick(d_list, n, E.d_status.as<int32_t>(), E.d_vol.as<double>(), gt, desc,
                               E.far_list_min, flags.as<uint32_t>(), lim.as<uint32_t>(), lnx.as<double>(),
                               E.stream));
    std::vector<uint32_t> hf(n), hl(n), hlim(n);
    VD_HIP(hipMemcpyAsync(hf.data(), flags.p, (size_t)n * 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipMemcpyAsync(hlim.data(), lim.p, (size_t)n * 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipMemcpyAsync(hl.data(), d_list, (size_t)n * 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    // selected cells and their chunks: cp[s] = first block of cell s
    std::vector<uint32_t> sq, hsel, hslim, cp(1, 0u);
    uint64_t blocks = 0;
    // the cells with the most entries to walk first, at most kFarListCells of them
    std::vector<uint32_t> cand;
    for (uint32_t q = 0; q < n; ++q)
        if (hf[q] && hlim[q]) cand.push_back(q);
    std::stable_sort(cand.begin(), cand.end(), [&](uint32_t a, uint32_t b) { return hlim[a] > hlim[b]; });
    if (cand.size() > kFarListCells) cand.resize(kFarListCells);
    uint64_t entries = 0;   // bounds the kept entries, whose offsets are 32-bit
    for (uint32_t q : cand) {
        const uint32_t lq = hlim[q];
        const uint64_t c = (lq + vd::kListChunk - 1) / vd::kListChunk;
        if (blocks + c >= (uint64_t(1) << 31) || entries + lq > 0xffffffffull) break;
        entries += lq;
        sq.push_back(q);
        hsel.push_back(hl[q]);
        hslim.push_back(lq);
        blocks += c;
        cp.push_back((uint32_t)blocks);
    }
    if (sq.empty()) return VD_OK;
    const uint32_t nsel = (uint32_t)sq.size(), nb = (uint32_t)blocks;
    VD_TRY(ensure(sel, (size_t)nsel * 4));
    VD_TRY(ensure(slim, (size_t)nsel * 4));
    VD_TRY(ensure(cpd, (size_t)(nsel + 1) * 4));
    VD_TRY(ensure(counts, (size_t)nb * 4));
    VD_TRY(ensure(offs, (size_t)nb * 4));
    VD_HIP(hipMemcpyAsync(sel.p, hsel.data(), (size_t)nsel * 4, hipMemcpyHostToDevice, E.stream));
    VD_HIP(hipMemcpyAsync(slim.p, hslim.data(), (size_t)nsel * 4, hipMemcpyHostToDevice, E.stream));
    VD_HIP(hipMemcpyAsync(cpd.p, cp.data(), (size_t)(nsel + 1) * 4, hipMemcpyHostToDevice, E.stream));
    VD_HIP(vd::launch_far_lists(sel.as<uint32_t>(), slim.as<uint32_t>(), cpd.as<uint32_t>(), nsel, nb, desc, gt,
                                E.d_gen.as<double>(), E.d_bin_start.as<uint32_t>(), counts.as<uint32_t>(),
                                nullptr, nullptr, E.stream));
    VD_TRY(scan(E, counts.as<uint32_t>(), offs.as<uint32_t>(), nb));
    // first list entry of each selected cell (the scanned count of its first block) and the total
    std::vector<uint64_t> hidx(nsel);
    for (uint32_t k = 0; k < nsel; ++k) hidx[k] = cp[k];
    VD_TRY(ensure(gidx, (size_t)nsel * 8));
    VD_TRY(ensure(at, (size_t)nsel * 4));
    VD_HIP(hipMemcpyAsync(gidx.p, hidx.data(), (size_t)nsel * 8, hipMemcpyHostToDevice, E.stream));
    VD_HIP(vd::launch_gather_u32(offs.as<uint32_t>(), gidx.as<uint64_t>(), nsel, at.as<uint32_t>(), E.stream));
    std::vector<uint32_t> start(nsel);
    uint32_t tail[2] = {0, 0};
    VD_HIP(hipMemcpyAsync(start.data(), at.p, (size_t)nsel * 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipMemcpyAsync(&tail[0], offs.as<uint32_t>() + nb - 1, 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipMemcpyAsync(&tail[1], counts.as<uint32_t>() + nb - 1, 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    const uint64_t all = (uint64_t)tail[0] + tail[1];
    // the cells (in the order above) whose lists fit in kFarListEntries: a prefix of the blocks
    uint32_t keep = nsel;
    while (keep > 0 && (keep < nsel ? (uint64_t)start[keep] : all) > kFarListEntries) --keep;
    if (keep == 0) return VD_OK;
    const uint64_t total = keep < nsel ? (uint64_t)start[keep] : all;
    // one spare entry: a walk's bulk skip reads entry 0 of its list even when the list is empty
    VD_TRY(ensure(lsh, (size_t)(total + 1) * sizeof(vd::ShellCode)));
    VD_HIP(vd::launch_far_lists(sel.as<uint32_t>(), slim.as<uint32_t>(), cpd.as<uint32_t>(), keep, cp[keep], desc,
                                gt, E.d_gen.as<double>(), E.d_bin_start.as<uint32_t>(), nullptr,
                                offs.as<uint32_t>(), lsh.as<vd::ShellCode>(), E.stream));
    std::vector<uint64_t> hoff(n, 0);
    std::vector<uint32_t> hn(n, vd::kNoList);
    for (uint32_t k = 0; k < keep; ++k) {
        const uint64_t end = k + 1 < keep ? start[k + 1] : total;
        hoff[sq[k]] = start[k];
        hn[sq[k]] = (uint32_t)(end - start[k]);
    }
    VD_TRY(ensure(lof, (size_t)n * 8));
    VD_TRY(ensure(lnn, (size_t)n * 4));
    VD_HIP(hipMemcpyAsync(lof.p, hoff.data(), (size_t)n * 8, hipMemcpyHostToDevice, E.stream));
    VD_HIP(hipMemcpyAsync(lnn.p, hn.data(), (size_t)n * 4, hipMemcpyHostToDevice, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    use = true;
    return VD_OK;
}

// Everything the first-round sequence depends on besides the generator coordinates: sizes,
// options, buffer addresses and the member descriptors. Equal keys give identical launches.
std::vector<uint64_t> round0_key(const Engine &E, uint32_t n) {
    std::vector<uint64_t> k = {n, E.nbins, E.nraw, (uint64_t)E.binned, (uint64_t)E.keep_cells, E.cap[0], E.cap[1],
                               E.cap[2], (uint64_t)E.occ[0], (uint64_t)E.cus, (uint64_t)E.reserve_cus,
                               (uint64_t)E.mem.size(), E.ntube, (uint64_t)E.seal, E.nhost, E.seal_slots,
                               (uint64_t)E.t0_waves, E.qgrain, (uint64_t)E.first_tier,
                               (uint64_t)E.first_probe};
    const DevBuf *bufs[] = {&E.d_gen, &E.d_gen_tess, &E.d_gen_bin, &E.d_counts, &E.d_bin_start, &E.d_fill,
                            &E.d_items, &E.d_spos, &E.d_scan, &E.d_desc, &E.d_vol, &E.d_cen, &E.d_nface,
                            &E.d_nvert, &E.d_status, &E.d_cell_off, &E.d_vtx, &E.d_fcnt, &E.d_fnb, &E.d_fidx,
                            &E.d_top, &E.d_queue, &E.d_counts4, &E.d_out0, &E.d_out1, &E.d_tube,
                            &E.d_wdata, &E.d_smask, &E.d_soff, &E.d_snb, &E.d_slog, &E.d_slogcnt,
                            &E.d_host_off, &E.d_host_planes};
    for (const DevBuf *b : bufs) k.push_back(reinterpret_cast<uint64_t>(b->p));
    k.push_back(reinterpret_cast<uint64_t>(E.h_counts));
    const size_t words = E.desc.size() * sizeof(vd::TessDesc) / 8;
    const size_t at = k.size();
    k.resize(at + words);
    if (words) std::memcpy(k.data() + at, E.desc.data(), words * 8);
    return k;
}

// Runs the first round, from a graph when the same sequence ran just before, and waits for it.
// *ms: time of the tier-0 launch (eager) or of the whole replayed sequence (graph).
int run_round0(Engine &E, uint32_t n, float *ms) {
    std::vector<uint64_t> key = round0_key(E, n);
    const bool repeat = E.graphs && key == E.last_key;
    E.last_key = key;
    if (repeat) {
        if (!E.gexec || key != E.gkey) {
            if (E.gexec) (void)hipGraphExecDestroy(E.gexec);
            E.gexec = nullptr;
            E.gkey.clear();
            VD_HIP(hipStreamBeginCapture(E.stream, hipStreamCaptureModeThreadLocal));
            const int rc = enqueue_round0(E, n, nullptr);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(E.stream, &g);
            if (rc != VD_OK) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (ec != hipSuccess) return fail(VD_EDEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
            const hipError_t ei = hipGraphInstantiate(&E.gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ei != hipSuccess) {
                E.gexec = nullptr;
                return fail(VD_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(ei));
            }
            E.gkey = key;
            E.graph_captures += 1;
        }
        VD_HIP(hipEventRecord(E.ev0, E.stream));
        VD_HIP(hipGraphLaunch(E.gexec, E.stream));
        VD_HIP(hipEventRecord(E.ev1, E.stream));
        VD_HIP(hipEventSynchronize(E.ev1));
        VD_HIP(hipEventElapsedTime(ms, E.ev0, E.ev1));
        E.graph_launches += 1;
    } else {
        VD_TRY(enqueue_round0(E, n, E.static_mode ? nullptr : ms));
        VD_HIP(hipStreamSynchronize(E.stream));
    }
    E.binned = true;
    return VD_OK;
}

// get_bin_index (algo_3d_grid.rs:98-107) on the host, the expression of k_bin_count.
uint32_t host_axis_bin(double x, double lo, double s, double lim) {
    double v = (x - lo) * s;
    if (v < 0.0) v = 0.0;
    if (v > lim) v = lim;
    if (!(v == v)) return 0;
    return (uint32_t)v;
}

uint32_t host_bin_of(const vd::GridGeom &G, const double q[3]) {
    const uint32_t ix = host_axis_bin(q[0], G.lo[0], G.s[0], G.lim[0]);
    const uint32_t iy = host_axis_bin(q[1], G.lo[1], G.s[1], G.lim[1]);
    const uint32_t iz = host_axis_bin(q[2], G.lo[2], G.s[2], G.lim[2]);
    return ix + iy * (uint32_t)G.n[0] + iz * (uint32_t)G.n[0] * (uint32_t)G.n[1];
}

// The bins of a single tessellation's current generators on the host, as the binning builds them
// (algo_3d_grid.rs:111-130: every bin lists its generators in index order), and their positions.
int host_bins(Engine &E) {
    const Member &m = E.mem[0];
    const uint64_t n = E.ngen, nb = m.nbins;
    std::vector<double> &g = E.h_gen;
    g.resize(3 * n);
    if (n) VD_HIP(hipMemcpyAsync(g.data(), E.d_gen.p, n * 24, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    E.h_gen_bin.resize(n);
    E.h_bin_start.assign(nb + 1, 0u);
    for (uint64_t i = 0; i < n; ++i) {
        E.h_gen_bin[i] = host_bin_of(m.grid, &g[3 * i]);
        E.h_bin_start[E.h_gen_bin[i] + 1]++;
    }
    for (uint64_t b = 0; b < nb; ++b) E.h_bin_start[b + 1] += E.h_bin_start[b];
    std::vector<uint32_t> fill(E.h_bin_start.begin(), E.h_bin_start.end() - 1);
    E.h_items.resize(n);
    for (uint64_t i = 0; i < n; ++i) E.h_items[fill[E.h_gen_bin[i]]++] = (uint32_t)i;
    E.h_over.clear();
    E.pos_dirty = false;
    return VD_OK;
}

// The list of bin b as moves left it: its CSR slice until a move first touches it.
std::vector<uint32_t> &over_list(Engine &E, uint32_t b) {
    auto ins = E.h_over.try_emplace(b);
    if (ins.second) ins.first->second.assign(E.h_items.begin() + E.h_bin_start[b], E.h_items.begin() + E.h_bin_start[b + 1]);
    return ins.first->second;
}

// update_generator (algo_3d_grid.rs:132-143): generator i leaves bin `from` by swap_remove (its
// slot takes the bin's last entry) and is pushed at the end of bin `to`.
int host_move(Engine &E, uint32_t i, uint32_t from, uint32_t to) {
    std::vector<uint32_t> &src = over_list(E, from);
    const auto it = std::find(src.begin(), src.end(), i);
    if (it == src.end()) return fail(VD_ESTATE, "generator missing from its bin list");
    *it = src.back();
    src.pop_back();
    over_list(E, to).push_back(i);
    E.h_gen_bin[i] = to;
    return VD_OK;
}

// The moved bins' lists merged into the CSR: bins between two moved ones keep their entries (one
// block copy) and shift by the count change so far.
void merge_over(Engine &E) {
    if (E.h_over.empty()) return;
    const std::vector<uint32_t> &bs = E.h_bin_start, &it = E.h_items;
    const uint32_t nb = (uint32_t)(bs.size() - 1);
    std::vector<uint32_t> keys;
    keys.reserve(E.h_over.size());
    for (const auto &kv : E.h_over) keys.push_back(kv.first);
    std::sort(keys.begin(), keys.end());
    std::vector<uint32_t> ns(bs.size()), ni(it.size());
    int64_t delta = 0;
    uint32_t next = 0;   // first bin not written yet
    auto copy_run = [&](uint32_t b0, uint32_t b1) {   // unmoved bins [b0, b1)
        for (uint32_t b = b0; b < b1; ++b) ns[b] = (uint32_t)((int64_t)bs[b] + delta);
        if (bs[b1] > bs[b0])
            std::memcpy(&ni[(size_t)((int64_t)bs[b0] + delta)], &it[bs[b0]], (size_t)(bs[b1] - bs[b0]) * 4);
    };
    for (const uint32_t k : keys) {
        copy_run(next, k);
        const std::vector<uint32_t> &l = E.h_over[k];
        ns[k] = (uint32_t)((int64_t)bs[k] + delta);
        if (!l.empty()) std::memcpy(&ni[ns[k]], l.data(), l.size() * 4);
        delta += (int64_t)l.size() - (int64_t)(bs[k + 1] - bs[k]);
        next = k + 1;
    }
    copy_run(next, nb);
    ns[nb] = (uint32_t)((int64_t)bs[nb] + delta);   // delta is 0: moves keep the count
    E.h_bin_start.swap(ns);
    E.h_items.swap(ni);
    E.h_over.clear();
}

// Positions written by set_generator to the device, in one copy.
int flush_positions(Engine &E) {
    if (!E.custom_bins || !E.pos_dirty) return VD_OK;
    if (E.ngen) VD_HIP(hipMemcpyAsync(E.d_gen.p, E.h_gen.data(), E.ngen * 24, hipMemcpyHostToDevice, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    E.pos_dirty = false;
    return VD_OK;
}

// Bins and positions kept by set_generator to the device: the CSR, the positions, the bin-ordered
// positions and the tube planes.
int upload_custom_bins(Engine &E) {
    const uint64_t n = E.ngen, nb = E.mem[0].nbins;
    merge_over(E);
    VD_TRY(flush_positions(E));
    VD_HIP(hipMemcpyAsync(E.d_bin_start.p, E.h_bin_start.data(), (nb + 1) * 4, hipMemcpyHostToDevice, E.stream));
    if (n) VD_HIP(hipMemcpyAsync(E.d_items.p, E.h_items.data(), n * 4, hipMemcpyHostToDevice, E.stream));
    VD_HIP(vd::launch_fill_spos(E.d_gen.as<double>(), E.d_items.as<uint32_t>(), n, E.d_desc.as<vd::TessDesc>(),
                                E.d_gen_tess.as<uint32_t>(), E.d_spos.as<double4>(), E.stream));
    if (E.ntube)
        VD_HIP(vd::launch_tube_planes(E.d_gen.as<double>(), n, E.d_desc.as<vd::TessDesc>(),
                                      E.d_gen_tess.as<uint32_t>(), E.ntube, E.d_tube.as<double>(), E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    E.bins_dirty = false;
    return VD_OK;
}

// calculate() of every member (tessellation.rs:213-230).
int engine_calculate(Engine &E) {
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    if (!E.prepared) return fail(VD_ESTATE, "generators are not set");
    const uint64_t n = E.ngen;
    for (double &s : E.stats) s = 0.0;
    VD_TRY(ensure(E.d_vol, n * 8));
    VD_TRY(ensure(E.d_cen, n * 24));
    VD_TRY(ensure(E.d_nface, n * 4));
    VD_TRY(ensure(E.d_nvert, n * 4));
    VD_TRY(ensure(E.d_status, n * 4));
    if (E.keep_cells || E.keep_nb) VD_TRY(ensure(E.d_cell_off, n * 24));
    VD_TRY(ensure(E.d_top, 64));
    VD_TRY(ensure(E.d_queue, 64));
    VD_TRY(ensure(E.d_counts4, 8 * kTiers * sizeof(uint32_t)));
    // arena sizes before the first calculate or after more cells: arena_hint vertices, 3/5 of
    // that many faces and three times as many face indices per cell (a uniform cell has about 27,
    // 16 and 82); a launch that fills an arena doubles all three and runs again
    const uint64_t h = E.arena_hint;
    if (E.cap_cells < n) {
        const uint64_t want[3] = {n * h + 4096, n * h * 3 / 5 + 4096, n * h * 3 + 4096};
        for (int k = 0; k < 3; ++k) E.cap[k] = E.cap_cells ? std::max(E.cap[k], want[k]) : want[k];
        E.cap_cells = n;
    }
    E.calculated = false;
    if (n == 0) { E.calculated = true; return VD_OK; }
    // geometries with cells crossing empty regions: the first launch sees their first prefix only
    bool far = false;
    for (const Shape &s : E.shapes) far |= s.tab && far_geometry(*s.tab) && s.tab->k > kShellK0;
    if (far || E.trunc0) {
        E.trunc0 = far;
        VD_TRY(upload_desc(E));
    }
    if (E.custom_bins && E.bins_dirty) VD_TRY(upload_custom_bins(E));
    // every cell computed within the attempts and rounds; otherwise calculate() fails and
    // leaves no results (cells could be left unfinished, with arena offsets never written)
    bool done = false;
    for (int attempt = 0; attempt < E.max_attempts && !done; ++attempt) {
        if (E.keep_cells) {
            VD_TRY(ensure(E.d_vtx, E.cap[0] * 24));
            VD_TRY(ensure(E.d_fcnt, E.cap[1]));
            VD_TRY(ensure(E.d_fidx, E.cap[2] * 2));
        }
        if (E.keep_cells || E.keep_nb) VD_TRY(ensure(E.d_fnb, E.cap[1] * 4));
        VD_TRY(ensure(E.d_out0, n * 4));
        VD_TRY(ensure(E.d_out1, n * 4));

        // Tier 0 (register fast path, small LDS; or the first tier set) over every cell in bin order. A cell that
        // outgrows its tier moves to the next one; a cell whose walk runs past the shell prefix
        // reruns in its tier after the prefixes are extended. Reruns start from scratch and give
        // the same result. Cells that once ran past their prefix reach across empty space: they
        // rerun in the tiers' skip-ahead instantiation from then on. Each tier's list holds
        // those cells first (group 1, nsh of them), then the others (group 0).
        constexpr int NT = kTiers;
        DevBuf L[NT];
        const ScopedBufs keep_l{L, NT};
        const uint32_t *list[NT] = {E.d_items.as<uint32_t>()};
        uint32_t nlist[NT] = {(uint32_t)n};
        uint32_t nsh[NT] = {};
        bool arena_full = false;
        auto pending = [&]() { for (int k = 0; k < NT; ++k) if (nlist[k]) return true; return false; };
        for (int round = 0; round < 64 && pending(); ++round) {
            // outcome per group g and tier: counts[4 * (NT * g + tier) + {0 capacity, 1 shell,
            // 2 arena, 3 error}], and the lists of capacity and shell failures per group and tier
            uint32_t counts[8 * NT] = {};
            DevBuf out[4 * NT];
            DevBuf flist[4 * NT];   // per-cell shell lists of the far cells of each tier (far_lists)
            const ScopedBufs keep_out{out, 4 * NT}, keep_flist{flist, 4 * NT};
            bool lon[NT] = {};
            const uint32_t *ocap[2][NT] = {}, *oshell[2][NT] = {};
            auto cnt = [&](int g, int tier, int what) { return counts[4 * (NT * g + tier) + what]; };
            if (round == 0 && attempt == 0) {
                float ms = 0.f;
                VD_TRY(run_round0(E, (uint32_t)n, &ms));
                E.stats[4] = 1000.0 * ms;
                E.stats[0] = (double)n;
                std::memcpy(counts, E.h_counts, 4 * NT * sizeof(uint32_t));
                ocap[0][E.first_tier] = E.d_out0.as<uint32_t>();
                oshell[0][E.first_tier] = E.d_out1.as<uint32_t>();
            } else {
                if (round == 0) VD_HIP(vd::launch_fill_u32(reinterpret_cast<uint32_t *>(E.d_top.p), 0u, 12, E.stream));
                // group 1 (cells crossing empty space: few, long walks) runs on a side stream, so
                // the rest of the round fills the CUs its long-running waves leave idle; ev0 and ev1
                // (timing events of the first round's launch, free in later rounds) order it
                // lists first: the side stream's launches are ordered after everything before ev0
                for (int tier = 0; tier < NT; ++tier)
                    if (nsh[tier])
                        VD_TRY(far_lists(E, list[tier], nsh[tier], flist[4 * tier], flist[4 * tier + 1],
                                         flist[4 * tier + 2], flist[4 * tier + 3], lon[tier]));
                bool side = false;
                if (E.side_streams)
                    for (int tier = 0; tier < NT; ++tier) side = side || nsh[tier] > 0;
                if (side) {
                    VD_TRY(engine_stream2(E));
                    VD_HIP(hipEventRecord(E.ev0, E.stream));
                    VD_HIP(hipStreamWaitEvent(E.stream2, E.ev0, 0));
                }
                for (int tier = 0; tier < NT; ++tier) {
                    if (!nlist[tier]) continue;
                    if (nsh[tier])
                        VD_TRY(run_cells(E, tier, list[tier], nsh[tier], nullptr, true, 1 + tier,
                                         side ? E.stream2 : nullptr,
                                         lon[tier] ? flist[4 * tier].as<vd::ShellCode>() : nullptr,
                                         lon[tier] ? flist[4 * tier + 1].as<uint64_t>() : nullptr,
                                         lon[tier] ? flist[4 * tier + 2].as<uint32_t>() : nullptr,
                                         lon[tier] ? flist[4 * tier + 3].as<double>() : nullptr));
                    if (nlist[tier] > nsh[tier])
                        VD_TRY(run_cells(E, tier, list[tier] + nsh[tier], nlist[tier] - nsh[tier], nullptr, false,
                                         1 + NT + tier));
                    if (tier > 0) E.stats[1] += nlist[tier];
                }
                if (side) {
                    VD_HIP(hipEventRecord(E.ev1, E.stream2));
                    VD_HIP(hipStreamWaitEvent(E.stream, E.ev1, 0));
                }
                VD_HIP(vd::launch_fill_u32(E.d_counts4.as<uint32_t>(), 0u, 8 * NT, E.stream));
                for (int tier = 0; tier < NT; ++tier) {
                    for (int g = 0; g < 2; ++g) {
                        const uint32_t *lst = g ? list[tier] : list[tier] + nsh[tier];
                        const uint32_t m = g ? nsh[tier] : nlist[tier] - nsh[tier];
                        if (!m) continue;
                        DevBuf &oc = out[2 * (NT * g + tier)], &os = out[2 * (NT * g + tier) + 1];
                        VD_TRY(ensure(oc, (size_t)m * 4));
                        VD_TRY(ensure(os, (size_t)m * 4));
                        ocap[g][tier] = oc.as<uint32_t>();
                        oshell[g][tier] = os.as<uint32_t>();
                        VD_HIP(vd::launch_collect(E.d_status.as<int32_t>(), lst, m, oc.as<uint32_t>(),
                                                  os.as<uint32_t>(), E.d_counts4.as<uint32_t>() + 4 * (NT * g + tier),
                                                  E.stream));
                    }
                }
                VD_HIP(hipMemcpyAsync(E.h_counts, E.d_counts4.p, sizeof(counts), hipMemcpyDeviceToHost, E.stream));
                VD_HIP(hipStreamSynchronize(E.stream));
                std::memcpy(counts, E.h_counts, sizeof(counts));
            }
            bool shell = false;
            for (int g = 0; g < 2; ++g)
                for (int tier = 0; tier < NT; ++tier) {
                    if (cnt(g, tier, 3)) return fail(VD_ECAPACITY, "a cell has a face with 256 or more vertices");
                    if (cnt(g, tier, 2)) arena_full = true;
                    shell |= cnt(g, tier, 1) != 0;
                }
            if (arena_full) break;
            if (cnt(0, NT - 1, 0) || cnt(1, NT - 1, 0))
                return fail(VD_ECAPACITY, "a cell exceeds the largest polytope tier");
            if (shell) {
                // The geometries of the cells that walked past their prefix get a longer one: at
                // least twice as long, and long enough for the largest radius r^2 those cells had
                // when they stopped (the walk needs every key up to 4 r^2, and r^2 only shrinks as
                // the walk goes on), so one extension usually covers cells reaching across empty space.
                const size_t M = E.mem.size();
                VD_TRY(ensure(E.d_mflags, (M + 1) * 4));
                VD_TRY(ensure(E.d_mrmax, M * 8));
                VD_HIP(vd::launch_fill_u32(E.d_mflags.as<uint32_t>(), 0u, M + 1, E.stream));
                VD_HIP(vd::launch_fill_u32(E.d_mrmax.as<uint32_t>(), 0u, 2 * M, E.stream));
                uint64_t *d_rmax = E.d_mrmax.as<uint64_t>();
                for (int g = 0; g < 2; ++g)
                    for (int tier = 0; tier < NT; ++tier)
                        VD_HIP(vd::launch_mark_members(oshell[g][tier], cnt(g, tier, 1), E.d_gen_tess.as<uint32_t>(),
                                                       E.d_vol.as<double>(), E.d_mflags.as<uint32_t>(), d_rmax,
                                                       E.stream));
                std::vector<uint32_t> flags(M);
                std::vector<double> rmax(M);
                VD_HIP(hipMemcpyAsync(flags.data(), E.d_mflags.p, M * 4, hipMemcpyDeviceToHost, E.stream));
                VD_HIP(hipMemcpyAsync(rmax.data(), d_rmax, M * 8, hipMemcpyDeviceToHost, E.stream));
                VD_HIP(hipStreamSynchronize(E.stream));
                std::vector<char> grow(E.shapes.size(), 0);
                std::vector<double> need(E.shapes.size(), 0.0);
                for (size_t t = 0; t < M; ++t)
                    if (flags[t]) {
                        grow[E.mem[t].shape] = 1;
                        need[E.mem[t].shape] = std::max(need[E.mem[t].shape], rmax[t]);
                    }
                uint64_t nfar = 0;
                for (int g = 0; g < 2; ++g)
                    for (int tier = 0; tier < NT; ++tier) nfar += cnt(g, tier, 1);
                bool rebuilt = false;
                for (size_t k = 0; k < E.shapes.size(); ++k) {
                    if (!grow[k]) continue;
                    Shape &s = E.shapes[k];
                    if (s.trunc) continue;   // its whole prefix is there: the reruns see it
                    if (s.tab->complete) return fail(VD_EDEVICE, "shell walk past a complete table");
                    const double frac = std::max(s.tab->far_frac, (double)nfar / (double)n);
                    VD_TRY(build_table(E, s.n, s.cs, shell_k_for(s, need[k]), *s.tab));
                    s.tab->far_frac = frac;
                    rebuilt = true;
                }
                E.trunc0 = false;
                VD_TRY(upload_desc(E));
                if (rebuilt) E.stats[2] += 1;
            }
            // next list of tier k: group 1 = the shell failures that rerun in tier k (both groups;
            // those of tier 0 rerun in tier 1: cells reaching across empty space mostly outgrow
            // tier 0, and any tier gives the same cell) + the capacity failures of group 1 of tier
            // k-1; group 0 = the capacity failures of group 0 of tier k-1
            auto far_tier = [](int tier) { return tier < kFarTier ? kFarTier : tier; };
            DevBuf next[NT];
            const ScopedBufs keep_next{next, NT};
            uint32_t nnext[NT] = {}, nnsh[NT] = {};
            for (int tier = 0; tier < NT; ++tier) {
                uint32_t parts[2 * NT + 2] = {};
                const uint32_t *src[2 * NT + 2] = {};
                int np = 0;
                for (int t = 0; t < NT; ++t)
                    if (far_tier(t) == tier)
                        for (int g = 0; g < 2; ++g) { parts[np] = cnt(g, t, 1); src[np++] = oshell[g][t]; }
                if (tier > 0) { parts[np] = cnt(1, tier - 1, 0); src[np++] = ocap[1][tier - 1]; }
                for (int q = 0; q < np; ++q) nnsh[tier] += parts[q];
                if (tier > 0) { parts[np] = cnt(0, tier - 1, 0); src[np++] = ocap[0][tier - 1]; }
                for (int q = 0; q < np; ++q) nnext[tier] += parts[q];
                if (!nnext[tier]) continue;
                VD_TRY(ensure(next[tier], (size_t)nnext[tier] * 4));
                uint32_t at = 0;
                for (int q = 0; q < np; ++q) {
                    if (parts[q])
                        VD_HIP(hipMemcpyAsync(next[tier].as<uint32_t>() + at, src[q], parts[q] * 4,
                                              hipMemcpyDeviceToDevice, E.stream));
                    at += parts[q];
                }
            }
            VD_HIP(hipStreamSynchronize(E.stream));
            for (int tier = 0; tier < NT; ++tier) {
                release(L[tier]);
                L[tier] = next[tier];
                next[tier] = DevBuf();
                list[tier] = L[tier].as<uint32_t>();
                nlist[tier] = nnext[tier];
                nsh[tier] = nnsh[tier];
            }
            for (DevBuf &b : out) release(b);
            for (DevBuf &b : flist) release(b);
        }
        for (DevBuf &b : L) release(b);
        if (arena_full) {
            for (int k = 0; k < 3; ++k) E.cap[k] *= 2;
            continue;
        }
        if (pending()) return fail(VD_ECAPACITY, "cells still need reruns after 64 rounds");
        done = true;
    }
    if (!done)
        return fail(VD_ECAPACITY, "the cell arenas filled up on every attempt (" + std::to_string(E.max_attempts) +
                                      " launches, each with twice the arena of the one before)");
    uint64_t top[6] = {0, 0, 0, 0, 0, 0};
    VD_HIP(hipMemcpyAsync(top, E.d_top.p, sizeof(top), hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    E.calculated = true;
    E.stats[6] = E.keep_cells ? (double)top[3] : 0.0;
    E.stats[7] = E.keep_cells ? (double)top[5] : 0.0;
    size_t entries = 0;
    for (const Shape &s : E.shapes) entries += s.tab->count;
    E.stats[3] = (double)entries;
    E.stats[5] = std::chrono::duration<double, std::micro>(clock::now() - t0).count();
    return VD_OK;
}

// Topology of the last calculate for the post-passes (tessellation.rs:244-264): per cell the mask
// of its curved walls and, for cells with one, the neighbour list in face order.
int engine_cell_topology(Engine &E, bool lists) {
    const uint64_t n = E.ngen;
    VD_TRY(ensure(E.d_smask, (n + 1) * 4));
    VD_TRY(ensure(E.d_scnt, (n + 1) * 4));
    VD_TRY(ensure(E.d_soff, (n + 1) * 4));
    VD_HIP(vd::launch_fill_u32(E.d_scnt.as<uint32_t>(), 0u, n + 1, E.stream));
    VD_HIP(vd::launch_fill_u32(E.d_smask.as<uint32_t>(), 0u, n + 1, E.stream));
    VD_HIP(vd::launch_cell_topology(E.d_desc.as<vd::TessDesc>(), E.d_gen_tess.as<uint32_t>(), n,
                                    E.d_nface.as<uint32_t>(), E.d_cell_off.as<uint64_t>(), E.d_fnb.as<int32_t>(),
                                    E.d_smask.as<uint32_t>(), E.d_scnt.as<uint32_t>(), E.stream));
    if (!lists) return VD_OK;
    VD_TRY(scan(E, E.d_scnt.as<uint32_t>(), E.d_soff.as<uint32_t>(), (uint32_t)(n + 1)));
    uint32_t total = 0;
    VD_HIP(hipMemcpyAsync(&total, E.d_soff.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    VD_TRY(ensure(E.d_snb, ((size_t)total + 1) * 4));
    // log room per neighbour: one entry per curved-wall plane the neighbour's generator emits
    uint32_t slots = 0;
    for (const Member &m : E.mem)
        for (const vd::WallDev &w : m.walls)
            if (!vd::wall_planar(w))
                slots += w.kind == 7 ? (w.npts < E.host_max.size() ? E.host_max[w.npts] : 1u) : 1u;
    E.seal_slots = std::max<uint32_t>((uint32_t)vd::kMaxWalls, slots);
    VD_TRY(ensure(E.d_slog, ((size_t)total + 1) * 3 * E.seal_slots * 4));
    VD_TRY(ensure(E.d_slogcnt, (n + 1) * 4));
    VD_HIP(vd::launch_cell_neighbors(E.d_desc.as<vd::TessDesc>(), E.d_gen_tess.as<uint32_t>(), n,
                                     E.d_nface.as<uint32_t>(), E.d_cell_off.as<uint64_t>(), E.d_fnb.as<int32_t>(),
                                     E.d_smask.as<uint32_t>(), E.d_soff.as<uint32_t>(), E.d_snb.as<uint32_t>(),
                                     E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    return VD_OK;
}

// calculate_sealed (tessellation.rs:234-318): calculate, the topology of its cells, then a second
// calculate whose cells take the sealing clips after their walk. log: the entries in cell order.
int engine_calculate_sealed(Engine &E, std::vector<int32_t> &log) {
    log.clear();
    VD_TRY(engine_calculate(E));
    const uint64_t n = E.ngen;
    if (n == 0) return VD_OK;
    VD_TRY(engine_cell_topology(E, true));
    E.seal = 1;
    const int rc = engine_calculate(E);
    E.seal = 0;
    VD_TRY(rc);
    std::vector<uint32_t> cnt(n), off(n + 1);
    VD_HIP(hipMemcpyAsync(cnt.data(), E.d_slogcnt.p, n * 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipMemcpyAsync(off.data(), E.d_soff.p, (n + 1) * 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    std::vector<int32_t> all((size_t)off[n] * 3 * E.seal_slots);
    if (!all.empty())
        VD_HIP(hipMemcpyAsync(all.data(), E.d_slog.p, all.size() * 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    for (uint64_t i = 0; i < n; ++i)
        for (uint32_t k = 0; k < cnt[i]; ++k)
            for (int c = 0; c < 3; ++c) log.push_back(all[(size_t)off[i] * 3 * E.seal_slots + 3 * k + c]);
    return VD_OK;
}

// prune_boundaries (tessellation.rs:324-385) over the cells of the last calculate; needs full cells.
int engine_prune(Engine &E, std::vector<int32_t> &log, std::vector<double> &pos) {
    if (!E.calculated) return fail(VD_ESTATE, "prune_boundaries() needs calculate() first");
    if (!E.keep_cells) return fail(VD_ESTATE, "prune_boundaries() needs full cells (VD_OPT_KEEP_CELLS)");
    log.clear();
    pos.clear();
    const uint64_t n = E.ngen;
    if (n == 0) return VD_OK;
    VD_TRY(engine_cell_topology(E, false));
    VD_TRY(ensure(E.d_pcnt, (n + 1) * 8));
    VD_TRY(ensure(E.d_poff, (n + 1) * 8));
    uint32_t *icnt = E.d_pcnt.as<uint32_t>(), *pcnt = icnt + (n + 1);
    uint32_t *ioff = E.d_poff.as<uint32_t>(), *poff = ioff + (n + 1);
    VD_HIP(vd::launch_fill_u32(icnt, 0u, 2 * (n + 1), E.stream));
    const auto prune = [&](int32_t *oi, double *op) {
        return vd::launch_prune(E.d_desc.as<vd::TessDesc>(), E.d_gen_tess.as<uint32_t>(), n,
                                E.d_nface.as<uint32_t>(), E.d_cell_off.as<uint64_t>(), E.d_fnb.as<int32_t>(),
                                E.d_fcnt.as<uint8_t>(), E.d_fidx.as<uint16_t>(), E.d_vtx.as<double>(),
                                E.d_smask.as<uint32_t>(), icnt, pcnt, ioff, poff, oi, op, E.stream);
    };
    VD_HIP(prune(nullptr, nullptr));
    VD_TRY(scan(E, icnt, ioff, (uint32_t)(n + 1)));
    VD_TRY(scan(E, pcnt, poff, (uint32_t)(n + 1)));
    uint32_t tot[2] = {0, 0};
    VD_HIP(hipMemcpyAsync(&tot[0], ioff + n, 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipMemcpyAsync(&tot[1], poff + n, 4, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    VD_TRY(ensure(E.d_pi, ((size_t)tot[0] + 1) * 4));
    VD_TRY(ensure(E.d_pp, ((size_t)tot[1] + 1) * 8));
    VD_HIP(prune(E.d_pi.as<int32_t>(), E.d_pp.as<double>()));
    log.resize(tot[0]);
    pos.resize(tot[1]);
    if (tot[0]) VD_HIP(hipMemcpyAsync(log.data(), E.d_pi.p, (size_t)tot[0] * 4, hipMemcpyDeviceToHost, E.stream));
    if (tot[1]) VD_HIP(hipMemcpyAsync(pos.data(), E.d_pp.p, (size_t)tot[1] * 8, hipMemcpyDeviceToHost, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    return VD_OK;
}

// relax() of every member (tessellation.rs:482-495): new input points, then set_generators.
int engine_relax(Engine &E) {
    if (!E.calculated) return fail(VD_ESTATE, "relax() needs calculate() first");
    const uint64_t n = E.ngen;
    VD_TRY(ensure(E.d_raw, (n + 1) * 24));
    VD_HIP(vd::launch_relax(E.d_gen.as<double>(), E.d_cen.as<double>(), E.d_nvert.as<uint32_t>(), n,
                            E.d_raw.as<double>(), E.stream));
    for (Member &m : E.mem) m.nraw = m.ngen;
    engine_layout(E);
    return engine_prepare(E);
}

const Member *member(const Engine &E, size_t t) {
    if (t >= E.mem.size()) {
        fail(VD_ERANGE, "member index out of range");
        return nullptr;
    }
    return &E.mem[t];
}

int need_results(const Engine &E) {
    if (!E.calculated) return fail(VD_ESTATE, "calculate() has not run since the generators changed");
    return VD_OK;
}

int copy_slice(const DevBuf &src, size_t elem, const Member &m, void *out, size_t n) {
    if (n < m.ngen) return fail(VD_EINVAL, "output buffer smaller than the cell count");
    if (m.ngen)
        VD_HIP(hipMemcpy(out, src.as<char>() + m.gen_off * elem, m.ngen * elem, hipMemcpyDeviceToHost));
    return VD_OK;
}

int get_neighbors(const Engine &E, const Member &m, uint64_t *offsets, size_t n, int32_t *ids, size_t cap) {
    VD_TRY(need_results(E));
    if (!offsets) return fail(VD_EINVAL, "NULL offsets");
    if (n < m.ngen) return fail(VD_EINVAL, "offsets buffer smaller than the cell count");
    const size_t nc = m.ngen;
    std::vector<uint32_t> nf(nc);
    if (nc) VD_HIP(hipMemcpy(nf.data(), E.d_nface.as<uint32_t>() + m.gen_off, nc * 4, hipMemcpyDeviceToHost));
    offsets[0] = 0;
    for (size_t i = 0; i < nc; ++i) offsets[i + 1] = offsets[i] + nf[i];
    if (!ids) return VD_OK;
    if (cap < offsets[nc]) return fail(VD_EINVAL, "ids buffer too small");
    const uint64_t total = offsets[nc];
    if (total == 0) return VD_OK;
    // the member's ids are packed on the device (a batch member's cells are spread over the
    // shared face arena), then copied once
    DevBuf tmp[2];
    const ScopedBufs tmp_g{tmp, 2};
    VD_TRY(ensure(tmp[0], (nc + 1) * 8));
    VD_TRY(ensure(tmp[1], total * 4));
    VD_HIP(hipMemcpy(tmp[0].p, offsets, (nc + 1) * 8, hipMemcpyHostToDevice));
    VD_HIP(vd::launch_gather_nb(E.d_fnb.as<int32_t>(), E.d_cell_off.as<uint64_t>() + 3 * m.gen_off,
                                tmp[0].as<uint64_t>(), nc, tmp[1].as<int32_t>(), nullptr));
    VD_HIP(hipMemcpy(ids, tmp[1].p, total * 4, hipMemcpyDeviceToHost));
    return VD_OK;
}

int get_cell(const Engine &E, const Member &m, size_t i, uint64_t sizes[3], double *vertices,
             uint32_t *face_counts, uint32_t *face_indices, int32_t *face_neighbors) {
    if (!sizes) return fail(VD_EINVAL, "NULL sizes");
    VD_TRY(need_results(E));
    if (i >= m.ngen) return fail(VD_ERANGE, "cell index out of range");
    if (!E.keep_cells) return fail(VD_ESTATE, "cells are not kept (VD_OPT_KEEP_CELLS = 0)");
    const size_t g = m.gen_off + i;
    uint32_t nv = 0, nf = 0;
    uint64_t off[3];
    VD_HIP(hipMemcpy(&nv, E.d_nvert.as<uint32_t>() + g, 4, hipMemcpyDeviceToHost));
    VD_HIP(hipMemcpy(&nf, E.d_nface.as<uint32_t>() + g, 4, hipMemcpyDeviceToHost));
    VD_HIP(hipMemcpy(off, E.d_cell_off.as<uint64_t>() + 3 * g, 24, hipMemcpyDeviceToHost));
    std::vector<uint8_t> fc(nf);
    if (nf) VD_HIP(hipMemcpy(fc.data(), E.d_fcnt.as<uint8_t>() + off[1], nf, hipMemcpyDeviceToHost));
    uint64_t ns = 0;
    for (uint8_t c : fc) ns += c;
    sizes[0] = nv;
    sizes[1] = nf;
    sizes[2] = ns;
    if (vertices && nv) VD_HIP(hipMemcpy(vertices, E.d_vtx.as<double>() + 3 * off[0], nv * 24, hipMemcpyDeviceToHost));
    if (face_counts) for (uint32_t k = 0; k < nf; ++k) face_counts[k] = fc[k];
    if (face_neighbors && nf)
        VD_HIP(hipMemcpy(face_neighbors, E.d_fnb.as<int32_t>() + off[1], nf * 4, hipMemcpyDeviceToHost));
    if (face_indices && ns) {
        std::vector<uint16_t> fi(ns);
        VD_HIP(hipMemcpy(fi.data(), E.d_fidx.as<uint16_t>() + off[2], ns * 2, hipMemcpyDeviceToHost));
        for (uint64_t k = 0; k < ns; ++k) face_indices[k] = fi[k];
    }
    return VD_OK;
}

int check_grid(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny, uint32_t nz) {
    if (!bmin || !bmax || nx == 0 || ny == 0 || nz == 0) return fail(VD_EINVAL, "invalid grid");
    if (nx > 16384 || ny > 16384 || nz > 16384) return fail(VD_EINVAL, "grid resolution above 16384");
    if ((uint64_t)nx * ny * nz >= 0x7fffffffull) return fail(VD_EINVAL, "too many bins");
    return VD_OK;
}

}  // namespace

struct vd_tess {
    Engine E;
    std::vector<int32_t> seal_log;     // (cell, neighbour, wall id) per sealing clip
    std::vector<int32_t> prune_log;    // (cell, neighbour, wall id, vertex count) per entry
    std::vector<double> prune_pos;     // the vertices of the prune entries, xyz
};

struct vd_batch {
    Engine E;
    std::vector<double> staged;   // host points of every member, in member order
    bool uploaded = false;
    double per_bin = kPerBinDefault;
};


// ---- serving: continuous batching of tessellation and Lloyd jobs --------------------------
namespace {

struct Job {
    uint64_t id = 0;
    Member m;                      // geometry, grid and walls
    uint32_t iters = 1;            // calculate() runs; a relax() between consecutive ones
    uint32_t done = 0;
    std::vector<double> host;      // input points until admission
    DevBuf pts;                    // current input points on the device (after admission)
    uint64_t npts = 0;
    size_t bytes = 0;              // HBM charged to the budget while admitted
    int state = 0;                 // 0 pending, 1 active, 2 finished, 3 rejected
    uint64_t cells = 0;            // cells of the last calculate
    uint64_t total_cells = 0;      // cells over all its calculate() runs
    std::vector<double> vol, cen, gen;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int lane = 0;                  // 0 large jobs, 1 small jobs
    bool uploaded = false;         // pts holds the input (the lane uploads `host` at its next step)
};

// Device bytes a job can hold while it runs, an upper bound: its point buffer; its share of a lane
// engine's buffers in one step (the input and compacted generators with their member, flag and
// position words, bin and item ids, the bin-ordered copy, the reductions and status, the retry
// lists of a round, the relaxed copy, tube-wall planes); its grid's bins and scan scratch; its
// walls' data and descriptor; and the shell-prefix table of its grid at its largest (every offset
// of the grid's enumeration box) with the scratch that table's construction needs. Server lanes
// keep no neighbour ids or full cells (no cell arenas) and no per-cell far lists; their
// global-memory tier slots are the server's fixed reserve.
size_t job_bytes(const Member &m, uint64_t n) {
    uint64_t ntube = 0;
    for (const vd::WallDev &w : m.walls) ntube += w.kind == VD_WALL_TUBE;
    const uint64_t per_point = 24 + 24 + 4 + 4 + 4 + 24 + 4 + 4 + 4 + 32 + 8 + 24 + 4 + 4 + 4 + 8 + 16 + 24 + 56 * ntube;
    const uint64_t bins = (uint64_t)m.n[0] * m.n[1] * m.n[2];
    size_t b = (n + 1) * per_point + (bins + 1) * 12 + vd::scan_tmp_words((uint32_t)std::max(bins, n) + 1) * 4;
    for (const std::vector<double> &d : m.wdata) b += d.size() * 8;
    b += 4096;
    const int64_t nmax = std::max(m.n[0], std::max(m.n[1], m.n[2]));
    const int64_t kfull = (int64_t)std::ceil(std::sqrt(3.0) * (double)nmax) + 3;
    uint64_t box = 1;
    for (int a = 0; a < 3; ++a) box *= (uint64_t)(2 * std::min(kfull, m.n[a]) + 1);
    box = std::min<uint64_t>(box, 0x7fffffffull);
    b += box * sizeof(vd::ShellCode) + 32 + vd::shell_scratch_bytes((uint32_t)box);
    return b;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ------------------------------------------------------------------------------------------
// vd_calculate_static: one tessellation computed in a file-scope device array instead of
// hipMalloc'd buffers. The array's name is a host-side stand-in; its device address on each GPU is
// taken once with hipGetSymbolAddress (the device must be current) and used wherever host code
// hands the memory to the device.
// ------------------------------------------------------------------------------------------
static_assert(VD_STATIC_SCRATCH_BYTES % 256 == 0, "whole 256-byte blocks");
static __device__ __align__(256) unsigned char scratch[VD_STATIC_SCRATCH_BYTES];

namespace {

std::mutex g_scratch_mu;
std::map<int, char *> g_scratch_ptr;   // device -> address of that device's copy of `scratch`

int scratch_address(int device, char **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it == g_scratch_ptr.end()) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(VD_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        it = g_scratch_ptr.emplace(device, static_cast<char *>(ptr)).first;
    }
    *out = it->second;
    return VD_OK;
}

// Sets g_static for the scope of one vd_calculate_static call.
struct StaticScope {
    StaticArena arena;
    explicit StaticScope(char *base) : arena{base, VD_STATIC_SCRATCH_BYTES, 0} { g_static = &arena; }
    ~StaticScope() { g_static = nullptr; }
};

}  // namespace

// One lane of the server: an engine (with its own stream) and the jobs it runs. The lane of large
// jobs runs its steps on a worker thread, so that small jobs, stepped by the calling thread on the
// other lane, do not wait for them; its cell launches leave kReservedCUs to the small lane. While
// a lane's step runs, only that step touches the lane's engine and jobs; the calling thread applies
// the outcome (job states, budget) when it collects the step.
struct Lane {
    Engine E;
    AllocAccount *acct = nullptr;    // the server's: every device buffer of the lane is charged to it
    std::vector<Job *> active;       // jobs of the running or next step
    std::vector<uint64_t> cells;     // per active job: cells of the step's calculate
    std::thread worker;
    std::atomic<bool> running{false};
    std::atomic<bool> ready{false};
    int rc = VD_OK;
    std::string err;
    uint64_t step_cells = 0;
};

constexpr uint64_t kSmallJobDefault = 65536;   // jobs with at most this many points use the small lane
#ifndef VD_RESERVED_CUS
#define VD_RESERVED_CUS 16
#endif
constexpr int kReservedCUs = VD_RESERVED_CUS;
constexpr size_t kSpareBuffers = 64;

// Makes an account current on this thread for a scope.
struct AcctScope {
    AllocAccount *prev;
    explicit AcctScope(AllocAccount *a) : prev(g_acct) { g_acct = a; }
    ~AcctScope() { g_acct = prev; }
};

constexpr int kServerGlobalSlots = 2;   // global-memory tier slots of each server lane

struct vd_server {
    Lane lanes[2];                 // 0: large jobs (worker thread), 1: small jobs (calling thread)
    AllocAccount acct;             // every device byte the server holds (lanes, jobs, spare buffers)
    size_t reserve = 0;            // fixed part of the budget: the lanes' global-memory tier slots
    size_t budget = 0, used = 0;   // used: the admitted jobs' job_bytes
    int max_slots = 64;
    uint64_t next_id = 1;
    std::map<uint64_t, Job *> jobs;
    std::deque<Job *> pending;
    std::vector<Job *> admitted0;  // admitted large jobs waiting for the large lane's next step
    std::vector<DevBuf> spare;     // point buffers of finished jobs, reused by later ones: freeing
                                   // device memory waits for the whole device, running lanes included
    std::vector<DevBuf> deferred;  // memory the small lane released while a large step ran
    size_t nslots = 0;             // admitted, unfinished jobs
    double per_bin = kPerBinDefault;
    uint64_t small_max = kSmallJobDefault;
    uint64_t steps = 0, cells = 0;
};

// Launch shape of the cell kernels (VD_OPT_T0_WAVES, VD_OPT_QUEUE_GRAIN, VD_OPT_FIRST_TIER,
// VD_OPT_FIRST_PROBE).
enum LaunchKnob : int { kKnobT0Waves = 0, kKnobGrain = 1, kKnobFirstTier = 2, kKnobFirstProbe = 3 };
constexpr int kLaunchKnobs = 4;

int engine_set_launch(Engine &E, int knob, int64_t value) {
    if (knob == kKnobT0Waves) {
        if (value < 5 || value > 8 || !vd::tier0_waves_valid((int)value))
            return fail(VD_EINVAL, "tier-0 waves per SIMD must be 5, 6, 7 or 8");
        int occ = 1;
        VD_HIP(vd::cells_occupancy(0, &occ, (int)value));
        E.t0_waves = (int)value;
        E.occ[0] = std::max(1, occ);
        return VD_OK;
    }
    if (knob == kKnobGrain) {
        if (value < 1 || value > 64) return fail(VD_EINVAL, "queue grain must be in [1, 64]");
        E.qgrain = (uint32_t)value;
        return VD_OK;
    }
    if (knob == kKnobFirstTier) {
        if (value < 0 || value > 1) return fail(VD_EINVAL, "first tier must be 0 or 1");
        E.first_tier = (int)value;
        return VD_OK;
    }
    if (value < 0 || value > 1) return fail(VD_EINVAL, "first-launch skip-ahead must be 0 or 1");
    E.first_probe = (int)value;
    return VD_OK;
}

int64_t engine_launch_value(const Engine &E, int knob) {
    return knob == kKnobT0Waves ? E.t0_waves : knob == kKnobGrain ? (int64_t)E.qgrain
         : knob == kKnobFirstTier ? E.first_tier : E.first_probe;
}

// Coordinate descent over the launch shape: from the current one, each knob in turn (skip-ahead,
// waves, grain, first tier) takes the candidate under which run() (one calculate, returning its
// time in microseconds) was fastest, the other knobs held. Every shape gives the same results, so
// whichever calculate ran last left valid ones. picks: tier-0 waves per SIMD, queue grain, first
// tier, first-launch skip-ahead.
template <class Run>
int engine_tune_launch(Engine &E, Run run, int picks[kLaunchKnobs]) {
    static const std::vector<int64_t> cand[kLaunchKnobs] = {{5, 6, 7, 8}, {1, 2, 4, 8}, {0, 1}, {0, 1}};
    static const int order[kLaunchKnobs] = {kKnobFirstProbe, kKnobT0Waves, kKnobGrain, kKnobFirstTier};
    double cur = 0.0;
    VD_TRY(run(&cur));   // warm: shell tables, buffers
    VD_TRY(run(&cur));
    for (int k : order) {
        const int64_t keep = engine_launch_value(E, k);
        int64_t bv = keep;
        double bt = cur;
        for (int64_t v : cand[k]) {
            if (v == keep) continue;
            VD_TRY(engine_set_launch(E, k, v));
            double us = 0.0;
            VD_TRY(run(&us));
            if (us < bt) { bt = us; bv = v; }
        }
        VD_TRY(engine_set_launch(E, k, bv));
        cur = bt;
    }
    if (picks)
        for (int k = 0; k < kLaunchKnobs; ++k) picks[k] = (int)engine_launch_value(E, k);
    return VD_OK;
}

extern "C" {

const char *vd_last_error(void) { return g_err.c_str(); }

int vd_shell_prefix(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny, uint32_t nz,
                    uint32_t kmax, int32_t *dxyz, double *keys, size_t cap, size_t *count,
                    double *next_key) {
    VD_TRY(check_grid(bmin, bmax, nx, ny, nz));
    const int64_t n[3] = {nx, ny, nz};
    const vd::GridGeom g = make_grid(bmin, bmax, n);
    ShellPrefix sp;
    build_shell(n, g.cs, kmax, sp);
    const size_t m = std::min(cap, sp.e.size());
    for (size_t k = 0; k < m; ++k) {
        if (dxyz) {
            dxyz[3 * k] = sp.e[k].dx;
            dxyz[3 * k + 1] = sp.e[k].dy;
            dxyz[3 * k + 2] = sp.e[k].dz;
        }
        if (keys) keys[k] = sp.e[k].d2;
    }
    if (count) *count = sp.e.size();
    if (next_key) *next_key = sp.next;
    return VD_OK;
}

int vd_shell_prefix_device(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny, uint32_t nz,
                           uint32_t kmax, int32_t *dxyz, double *keys, size_t cap, size_t *count,
                           double *next_key) {
    VD_TRY(check_grid(bmin, bmax, nx, ny, nz));
    const int64_t n[3] = {nx, ny, nz};
    const vd::GridGeom g = make_grid(bmin, bmax, n);
    Engine E;   // only its stream (the null stream) and scratch buffer are used
    ShellTable t;
    int rc = build_table(E, n, g.cs, kmax, t);
    std::vector<vd::ShellCode> e(t.count);
    if (rc == VD_OK && t.count) {
        hipError_t h = hipMemcpy(e.data(), t.buf.p, (size_t)t.count * sizeof(vd::ShellCode), hipMemcpyDeviceToHost);
        if (h != hipSuccess) rc = fail(VD_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(h));
    }
    release(t.buf);
    release(E.d_shell_tmp);
    if (rc != VD_OK) return rc;
    const size_t m = std::min(cap, e.size());
    for (size_t k = 0; k < m; ++k) {
        int dx, dy, dz;
        vd::shell_unpack(e[k], t.pack, dx, dy, dz);
        if (dxyz) {
            dxyz[3 * k] = dx;
            dxyz[3 * k + 1] = dy;
            dxyz[3 * k + 2] = dz;
        }
        if (keys) keys[k] = vd::shell_key(dx, dy, dz, g.cs);
    }
    if (count) *count = e.size();
    if (next_key) *next_key = t.next;
    return VD_OK;
}

int vd_grid_for_points(const double bmin[3], const double bmax[3], const double *xyz, size_t n, double per_bin,
                       uint32_t out[3]) {
    if (!bmin || !bmax || !out || (!xyz && n) || !(per_bin > 0.0)) return fail(VD_EINVAL, "invalid argument");
    int64_t g[3];
    density_grid(bmin, bmax, xyz, n, per_bin, false, g);
    for (int a = 0; a < 3; ++a) out[a] = (uint32_t)g[a];
    return VD_OK;
}

int vd_calculate_static(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny, uint32_t nz,
                        const vd_wall *walls, int nwalls, const double *d_xyz, size_t n, double *d_vol,
                        double *d_cen, uint32_t *d_nface, size_t *ncells, void *stream) {
    // every check that needs no device comes first
    VD_TRY(check_grid(bmin, bmax, nx, ny, nz));
    if (!ncells || (n && (!d_xyz || !d_vol || !d_cen || !d_nface)) || nwalls < 0 || (nwalls && !walls))
        return fail(VD_EINVAL, "NULL argument");
    if (n > VD_STATIC_MAX_POINTS)
        return fail(VD_EINVAL, "vd_calculate_static takes at most " + std::to_string(VD_STATIC_MAX_POINTS) +
                                   " points, got " + std::to_string(n));
    if ((uint64_t)nx * ny * nz > VD_STATIC_MAX_BINS)
        return fail(VD_EINVAL, "vd_calculate_static takes at most " + std::to_string(VD_STATIC_MAX_BINS) +
                                   " bins, got " + std::to_string((uint64_t)nx * ny * nz));
    if (nwalls > vd::kMaxWalls) return fail(VD_EINVAL, "too many walls (max 8)");
    Member m;
    for (int a = 0; a < 3; ++a) { m.lo[a] = bmin[a]; m.hi[a] = bmax[a]; }
    m.n[0] = nx; m.n[1] = ny; m.n[2] = nz;
    m.grid = make_grid(m.lo, m.hi, m.n);
    for (int k = 0; k < nwalls; ++k) {
        vd::WallDev d;
        std::vector<double> data;
        VD_TRY(make_wall(walls[k], d, data));
        m.walls.push_back(d);
        m.wdata.push_back(std::move(data));
    }
    *ncells = 0;
    Engine E;
    E.static_mode = true;
    E.stream = static_cast<hipStream_t>(stream);
    E.h_counts = E.h_local;
    E.graphs = 0;
    E.side_streams = false;
    E.far_list_min = 0.0;
    E.keep_cells = 0;
    E.keep_nb = false;
    VD_HIP(hipGetDevice(&E.device));
    VD_HIP(hipDeviceGetAttribute(&E.cus, hipDeviceAttributeMultiprocessorCount, E.device));
    for (int k = 0; k < kTiers; ++k) {
        VD_HIP(vd::cells_occupancy(k, &E.occ[k], E.t0_waves));
        E.occ[k] = std::max(1, E.occ[k]);
    }
    char *base = nullptr;
    VD_TRY(scratch_address(E.device, &base));
    const StaticScope scope(base);
    struct Free {   // the engine's carved buffers go back before the scope ends
        Engine &E;
        ~Free() { engine_free(E); }
    } free_engine{E};
    E.mem.push_back(m);
    VD_TRY(ensure(E.d_raw, (n + 1) * 24));
    if (n) VD_HIP(hipMemcpyAsync(E.d_raw.p, d_xyz, n * 24, hipMemcpyDeviceToDevice, E.stream));
    E.mem[0].nraw = n;
    engine_layout(E);
    VD_TRY(engine_prepare(E));
    VD_TRY(engine_calculate(E));
    const uint64_t kept = E.ngen;
    if (kept) {
        VD_HIP(hipMemcpyAsync(d_vol, E.d_vol.p, kept * 8, hipMemcpyDeviceToDevice, E.stream));
        VD_HIP(hipMemcpyAsync(d_cen, E.d_cen.p, kept * 24, hipMemcpyDeviceToDevice, E.stream));
        VD_HIP(hipMemcpyAsync(d_nface, E.d_nface.p, kept * 4, hipMemcpyDeviceToDevice, E.stream));
    }
    VD_HIP(hipStreamSynchronize(E.stream));
    *ncells = kept;
    return VD_OK;
}

int vd_grid_for(const double bmin[3], const double bmax[3], size_t n, double per_bin, uint32_t out[3]) {
    if (!bmin || !bmax || !out || !(per_bin > 0.0)) return fail(VD_EINVAL, "invalid argument");
    int64_t g[3];
    grid_for(bmin, bmax, n, per_bin, g);
    for (int a = 0; a < 3; ++a) out[a] = (uint32_t)g[a];
    return VD_OK;
}

// ---- vd_tess: one Tessellation -----------------------------------------------------------
int vd_tess_create(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny, uint32_t nz,
                   vd_tess **out) {
    if (!out) return fail(VD_EINVAL, "out is NULL");
    *out = nullptr;
    VD_TRY(check_grid(bmin, bmax, nx, ny, nz));
    vd_tess *t = new vd_tess();
    Member m;
    for (int a = 0; a < 3; ++a) { m.lo[a] = bmin[a]; m.hi[a] = bmax[a]; }
    m.n[0] = nx; m.n[1] = ny; m.n[2] = nz;
    m.grid = make_grid(m.lo, m.hi, m.n);
    t->E.mem.push_back(m);
    int rc = engine_init(t->E);
    if (!rc) rc = ensure(t->E.d_raw, 24);
    if (!rc) { engine_layout(t->E); rc = engine_prepare(t->E); }
    if (rc) { vd_tess_destroy(t); return rc; }
    *out = t;
    return VD_OK;
}

void vd_tess_destroy(vd_tess *t) {
    if (!t) return;
    engine_free(t->E);
    delete t;
}

// add_wall (tessellation.rs:175-200) of a wall in device form; host: the caller's callbacks for
// a wall of kind 7 (nullptr otherwise).
static int tess_add_wall_dev(vd_tess *t, vd::WallDev d, std::vector<double> data, const vd_wall_host *host) {
    Engine &E = t->E;
    Member &m = E.mem[0];
    if ((int)m.walls.size() >= vd::kMaxWalls) return fail(VD_EINVAL, "too many walls (max 8)");
    // prune_outside_generators (tessellation.rs:175-194): the current generators are filtered
    // again by every wall. They become the raw input first (filtering them by the walls they
    // already passed keeps them all, and re-binning an unchanged set gives the same bins), so a
    // failure below leaves the handle as it was, without the new wall.
    const uint64_t n = E.ngen;
    VD_TRY(flush_positions(E));
    if (n) {
        DevBuf keep;
        const ScopedBufs keep_g{&keep, 1};
        VD_TRY(ensure(keep, n * 24));
        VD_HIP(hipMemcpyAsync(keep.p, E.d_gen.p, n * 24, hipMemcpyDeviceToDevice, E.stream));
        VD_TRY(ensure(E.d_raw, (n + 1) * 24));
        VD_HIP(hipMemcpyAsync(E.d_raw.p, keep.p, n * 24, hipMemcpyDeviceToDevice, E.stream));
        VD_HIP(hipStreamSynchronize(E.stream));
        m.nraw = n;
        engine_layout(E);
    }
    if (host) {
        d.npts = (uint32_t)m.host.size();
        m.host.push_back(*host);
    }
    m.walls.push_back(d);
    m.wdata.push_back(std::move(data));
    const bool custom = E.custom_bins;
    const int rc = n ? engine_prepare(E) : upload_desc(E);
    if (rc) {
        const std::string msg = g_err;
        m.walls.pop_back();
        m.wdata.pop_back();
        if (host) m.host.pop_back();
        (void)(n ? engine_prepare(E) : upload_desc(E));
        if (custom && E.ngen == n) E.custom_bins = E.bins_dirty = E.binned = true;
        return fail(rc, msg);
    }
    // prune_outside_generators (tessellation.rs:178-194) bins afresh only when the wall removed a
    // generator; otherwise the bins set_generator left stay as they are
    if (custom && E.ngen == n) E.custom_bins = E.bins_dirty = E.binned = true;
    return VD_OK;
}

static int tess_add_wall(vd_tess *t, const vd_wall &w) {
    vd::WallDev d;
    std::vector<double> data;
    VD_TRY(make_wall(w, d, data));
    return tess_add_wall_dev(t, d, std::move(data), nullptr);
}

int vd_tess_add_wall_planes_per_generator(vd_tess *t, int32_t id, const vd_wall_host *wall) {
    if (!t || !wall || !wall->contains || !wall->cut) return fail(VD_EINVAL, "NULL argument");
    if (id > VD_WALL_ID_MAX) return fail(VD_EINVAL, "Wall ID must be <= -1000");   // wall/mod.rs:25-27
    vd::WallDev d{};
    d.id = id;
    d.kind = vd::kWallHost;
    d.closed = wall->planar ? 1 : 0;
    return tess_add_wall_dev(t, d, {}, wall);
}

int vd_tess_add_wall_plane(vd_tess *t, int32_t id, const double point[3], const double normal[3]) {
    if (!t || !point || !normal) return fail(VD_EINVAL, "NULL argument");
    vd_wall w{};
    w.id = id;
    w.kind = 0;
    for (int k = 0; k < 3; ++k) { w.a[k] = point[k]; w.b[k] = normal[k]; }
    return tess_add_wall(t, w);
}

int vd_tess_add_wall_sphere(vd_tess *t, int32_t id, const double center[3], double radius) {
    if (!t || !center) return fail(VD_EINVAL, "NULL argument");
    vd_wall w{};
    w.id = id;
    w.kind = 1;
    w.r = radius;
    for (int k = 0; k < 3; ++k) w.a[k] = center[k];
    return tess_add_wall(t, w);
}

int vd_tess_add_wall(vd_tess *t, const vd_wall *wall) {
    if (!t || !wall) return fail(VD_EINVAL, "NULL argument");
    return tess_add_wall(t, *wall);
}

int vd_tess_clear_walls(vd_tess *t) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    t->E.mem[0].walls.clear();
    t->E.mem[0].wdata.clear();
    t->E.mem[0].host.clear();
    t->E.calculated = false;
    return upload_desc(t->E);
}

int vd_tess_set_generators(vd_tess *t, const double *xyz, size_t n) {
    if (!t || (!xyz && n)) return fail(VD_EINVAL, "NULL argument");
    if (n >= 0x7fffffffull) return fail(VD_EINVAL, "too many generators");
    Engine &E = t->E;
    VD_TRY(ensure(E.d_raw, (n + 1) * 24));
    if (n) VD_HIP(hipMemcpyAsync(E.d_raw.p, xyz, n * 24, hipMemcpyHostToDevice, E.stream));
    E.mem[0].nraw = n;
    engine_layout(E);
    return engine_prepare(E);
}

int vd_tess_set_generators_device(vd_tess *t, const double *d_xyz, size_t n) {
    if (!t || (!d_xyz && n)) return fail(VD_EINVAL, "NULL argument");
    if (n >= 0x7fffffffull) return fail(VD_EINVAL, "too many generators");
    Engine &E = t->E;
    VD_TRY(ensure(E.d_raw, (n + 1) * 24));
    if (n) VD_HIP(hipMemcpyAsync(E.d_raw.p, d_xyz, n * 24, hipMemcpyDeviceToDevice, E.stream));
    E.mem[0].nraw = n;
    engine_layout(E);
    return engine_prepare(E);
}

// random_generators (tessellation.rs:96-122): candidates in batches on the device, each batch
// through the wall filter of set_generators; the first `count` kept candidates in stream order
// among the first count * 1000 become the generators.
int vd_tess_random_generators(vd_tess *t, size_t count, uint64_t seed) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    if (count >= 0x7fffffffull) return fail(VD_EINVAL, "too many generators");
    Engine &E = t->E;
    Member &m = E.mem[0];
    const uint64_t limit = (uint64_t)count * 1000;   // max_attempts (:101)
    constexpr uint64_t kMaxBatch = uint64_t(1) << 24;
    struct Acc {
        DevBuf b;
        ~Acc() { release(b); }
    } acc;
    VD_TRY(ensure(acc.b, ((uint64_t)count + 1) * 24));
    uint64_t found = 0, tried = 0;
    uint64_t batch = std::min<uint64_t>(std::max<uint64_t>(count + count / 4, 4096), kMaxBatch);
    while (found < count && tried < limit) {
        const uint64_t b = std::min(batch, limit - tried);
        VD_TRY(ensure(E.d_raw, (b + 1) * 24));
        VD_HIP(vd::launch_random_points(seed, tried, b, m.lo, m.hi, E.d_raw.as<double>(), E.stream));
        m.nraw = b;
        engine_layout(E);
        VD_TRY(engine_filter(E));
        const uint64_t take = std::min<uint64_t>(E.ngen, count - found);
        if (take)
            VD_HIP(hipMemcpyAsync(acc.b.as<double>() + 3 * found, E.d_gen.p, take * 24, hipMemcpyDeviceToDevice,
                                  E.stream));
        found += take;
        tried += b;
        // the next batch from the share kept so far, with a quarter more for its spread
        const double rate = tried ? (double)std::max<uint64_t>(found, 1) / (double)tried : 1.0;
        const double want = (double)(count - found) / rate * 1.25 + 4096.0;
        batch = (uint64_t)std::min<double>(want, (double)kMaxBatch);
    }
    VD_HIP(hipStreamSynchronize(E.stream));
    // the kept points become the generators; every wall contains them, so the filter keeps all
    VD_TRY(ensure(E.d_raw, (found + 1) * 24));
    if (found)
        VD_HIP(hipMemcpyAsync(E.d_raw.p, acc.b.p, found * 24, hipMemcpyDeviceToDevice, E.stream));
    m.nraw = found;
    engine_layout(E);
    return engine_prepare(E);
}

int vd_tess_set_generator(vd_tess *t, size_t index, const double xyz[3]) {
    if (!t || !xyz) return fail(VD_EINVAL, "NULL argument");
    Engine &E = t->E;
    if (index >= E.ngen) return fail(VD_ERANGE, "generator index out of range (the reference ignores the call)");
    Member &m = E.mem[0];
    // every wall must contain the new position, or nothing changes (tessellation.rs:76-80)
    for (size_t k = 0; k < m.walls.size(); ++k) {
        const vd::WallDev &w = m.walls[k];
        bool in;
        if (w.kind == vd::kWallHost) {
            const vd_wall_host &cb = m.host[w.npts];
            in = cb.contains(cb.ctx, xyz) != 0;
        } else {
            vd::WallDev hw = w;
            hw.pts = m.wdata[k].empty() ? nullptr : m.wdata[k].data();
            in = vd::wall_contains(hw, xyz[0], xyz[1], xyz[2]);
        }
        if (!in) return VD_OK;
    }
    if (!E.custom_bins) VD_TRY(host_bins(E));
    const uint32_t to = host_bin_of(m.grid, xyz), from = E.h_gen_bin[index];
    if (to != from) VD_TRY(host_move(E, (uint32_t)index, from, to));   // algo_3d_grid.rs:136-142
    std::memcpy(&E.h_gen[3 * index], xyz, 24);                          // tessellation.rs:89-91
    E.custom_bins = E.bins_dirty = E.pos_dirty = E.binned = true;
    E.calculated = false;
    if (E.nhost) VD_TRY(host_planes_one(E, index, xyz));   // the caller's walls cut the moved generator anew
    return VD_OK;
}

int vd_tess_calculate(vd_tess *t) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    t->seal_log.clear();   // tessellation.rs:214-216
    t->prune_log.clear();
    t->prune_pos.clear();
    return engine_calculate(t->E);
}

int vd_tess_autotune_launch(vd_tess *t, int picks[4]) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    if (!t->E.prepared) return fail(VD_ESTATE, "generators are not set");
    return engine_tune_launch(t->E, [&](double *us) {
        VD_TRY(vd_tess_calculate(t));
        *us = t->E.stats[5];
        return VD_OK;
    }, picks);
}

int vd_tess_calculate_sealed(vd_tess *t) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    t->prune_log.clear();
    t->prune_pos.clear();
    return engine_calculate_sealed(t->E, t->seal_log);
}

int vd_tess_prune_boundaries(vd_tess *t) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    return engine_prune(t->E, t->prune_log, t->prune_pos);
}

int vd_tess_get_seal_log(const vd_tess *t, int32_t *out, size_t cap, size_t *count) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    if (out) std::memcpy(out, t->seal_log.data(), std::min(cap, t->seal_log.size()) * 4);
    if (count) *count = t->seal_log.size();
    return VD_OK;
}

int vd_tess_get_prune_log(const vd_tess *t, int32_t *ints, size_t icap, double *pos, size_t pcap,
                          size_t counts[2]) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    if (ints) std::memcpy(ints, t->prune_log.data(), std::min(icap, t->prune_log.size()) * 4);
    if (pos) std::memcpy(pos, t->prune_pos.data(), std::min(pcap, t->prune_pos.size()) * 8);
    if (counts) { counts[0] = t->prune_log.size(); counts[1] = t->prune_pos.size(); }
    return VD_OK;
}

int vd_tess_relax(vd_tess *t) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    return engine_relax(t->E);
}

size_t vd_tess_count_generators(const vd_tess *t) { return t ? t->E.ngen : 0; }
size_t vd_tess_count_cells(const vd_tess *t) { return (t && t->E.calculated) ? t->E.ngen : 0; }

int vd_tess_get_generators(const vd_tess *t, double *out, size_t n) {
    if (!t || (!out && t->E.ngen)) return fail(VD_EINVAL, "NULL argument");
    const Engine &E = t->E;
    if (E.custom_bins && E.pos_dirty) {   // moves not uploaded yet: the host copy is current
        if (n < E.ngen) return fail(VD_EINVAL, "output buffer smaller than the cell count");
        if (E.ngen) std::memcpy(out, E.h_gen.data(), E.ngen * 24);
        return VD_OK;
    }
    return copy_slice(E.d_gen, 24, E.mem[0], out, n);
}

int vd_tess_get_volumes(const vd_tess *t, double *out, size_t n) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    VD_TRY(need_results(t->E));
    return copy_slice(t->E.d_vol, 8, t->E.mem[0], out, n);
}

int vd_tess_get_centroids(const vd_tess *t, double *out3, size_t n) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    VD_TRY(need_results(t->E));
    return copy_slice(t->E.d_cen, 24, t->E.mem[0], out3, n);
}

int vd_tess_get_face_counts(const vd_tess *t, uint32_t *out, size_t n) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    VD_TRY(need_results(t->E));
    return copy_slice(t->E.d_nface, 4, t->E.mem[0], out, n);
}

int vd_tess_get_neighbors(const vd_tess *t, uint64_t *offsets, size_t n, int32_t *ids, size_t cap) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    return get_neighbors(t->E, t->E.mem[0], offsets, n, ids, cap);
}

int vd_tess_get_cell(const vd_tess *t, size_t i, uint64_t sizes[3], double *vertices, uint32_t *face_counts,
                     uint32_t *face_indices, int32_t *face_neighbors) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    return get_cell(t->E, t->E.mem[0], i, sizes, vertices, face_counts, face_indices, face_neighbors);
}

int vd_tess_set_option(vd_tess *t, int key, int64_t value) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    if (key == VD_OPT_KEEP_CELLS) {
        t->E.keep_cells = value ? 1 : 0;
        t->E.calculated = false;
        return VD_OK;
    }
    if (key == VD_OPT_GRAPHS) {
        t->E.graphs = value ? 1 : 0;
        return VD_OK;
    }
    if (key == VD_OPT_ARENA) {
        if (value < 1) return fail(VD_EINVAL, "at least one arena vertex per cell");
        t->E.arena_hint = (uint64_t)value;
        t->E.cap_cells = 0;   // resize at the next calculate
        return VD_OK;
    }
    if (key == VD_OPT_FAR_LIST) {
        if (value < 0) return fail(VD_EINVAL, "walk length must not be negative");
        t->E.far_list_min = (double)value;
        return VD_OK;
    }
    if (key == VD_OPT_ATTEMPTS) {
        if (value < 1 || value > 62) return fail(VD_EINVAL, "attempts must be in [1, 62]");
        t->E.max_attempts = (int)value;
        return VD_OK;
    }
    if (key == VD_OPT_T0_WAVES) return engine_set_launch(t->E, kKnobT0Waves, value);
    if (key == VD_OPT_QUEUE_GRAIN) return engine_set_launch(t->E, kKnobGrain, value);
    if (key == VD_OPT_FIRST_TIER) return engine_set_launch(t->E, kKnobFirstTier, value);
    if (key == VD_OPT_FIRST_PROBE) return engine_set_launch(t->E, kKnobFirstProbe, value);
    return fail(VD_EINVAL, "unknown option");
}

void engine_stats(const Engine &E, double out[10]) {
    for (int k = 0; k < 8; ++k) out[k] = E.stats[k];
    out[8] = (double)E.graph_launches;
    out[9] = (double)E.graph_captures;
}

int vd_tess_get_stats(const vd_tess *t, double out[10]) {
    if (!t || !out) return fail(VD_EINVAL, "NULL argument");
    engine_stats(t->E, out);
    return VD_OK;
}

// ---- vd_batch: many tessellations in one launch -------------------------------------------
int vd_batch_create(vd_batch **out) {
    if (!out) return fail(VD_EINVAL, "out is NULL");
    vd_batch *b = new vd_batch();
    b->E.keep_cells = 0;
    int rc = engine_init(b->E);
    if (rc) { engine_free(b->E); delete b; return rc; }
    *out = b;
    return VD_OK;
}

void vd_batch_destroy(vd_batch *b) {
    if (!b) return;
    engine_free(b->E);
    delete b;
}

int vd_batch_clear(vd_batch *b) {
    if (!b) return fail(VD_EINVAL, "NULL handle");
    b->E.mem.clear();
    b->E.shapes.clear();
    b->E.prepared = b->E.calculated = false;
    b->staged.clear();
    b->uploaded = false;
    return VD_OK;
}

int vd_batch_add(vd_batch *b, const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny, uint32_t nz,
                 const vd_wall *walls, int nwalls, const double *xyz, size_t n, uint32_t *index) {
    if (!b || !bmin || !bmax || (!xyz && n) || (!walls && nwalls > 0)) return fail(VD_EINVAL, "NULL argument");
    if (nwalls < 0 || nwalls > vd::kMaxWalls) return fail(VD_EINVAL, "0 to 8 walls per member");
    Member m;
    for (int a = 0; a < 3; ++a) { m.lo[a] = bmin[a]; m.hi[a] = bmax[a]; }
    if (nx == 0 && ny == 0 && nz == 0) {
        grid_for(bmin, bmax, n, b->per_bin, m.n);
        m.auto_grid = true;
    } else {
        VD_TRY(check_grid(bmin, bmax, nx, ny, nz));
        m.n[0] = nx; m.n[1] = ny; m.n[2] = nz;
    }
    m.grid = make_grid(m.lo, m.hi, m.n);
    for (int k = 0; k < nwalls; ++k) {
        vd::WallDev d;
        std::vector<double> data;
        VD_TRY(make_wall(walls[k], d, data));
        m.walls.push_back(d);
        m.wdata.push_back(std::move(data));
    }
    m.nraw = n;
    b->E.mem.push_back(m);
    b->staged.insert(b->staged.end(), xyz, xyz + 3 * n);
    b->uploaded = false;
    b->E.shapes.clear();
    b->E.prepared = b->E.calculated = false;
    if (index) *index = (uint32_t)(b->E.mem.size() - 1);
    return VD_OK;
}

int vd_batch_upload(vd_batch *b) {
    if (!b) return fail(VD_EINVAL, "NULL handle");
    Engine &E = b->E;
    for (size_t t = 0; t < E.mem.size(); ++t) E.mem[t].nraw = E.mem[t].nraw;
    engine_layout(E);
    if (E.nraw * 3 != b->staged.size()) return fail(VD_ESTATE, "staged points and members disagree");
    VD_TRY(ensure(E.d_raw, (E.nraw + 1) * 24));
    if (E.nraw)
        VD_HIP(hipMemcpyAsync(E.d_raw.p, b->staged.data(), E.nraw * 24, hipMemcpyHostToDevice, E.stream));
    VD_HIP(hipStreamSynchronize(E.stream));
    b->uploaded = true;
    return VD_OK;
}

int vd_batch_calculate(vd_batch *b) {
    if (!b) return fail(VD_EINVAL, "NULL handle");
    if (!b->uploaded) VD_TRY(vd_batch_upload(b));
    engine_layout(b->E);
    VD_TRY(engine_prepare(b->E));
    return engine_calculate(b->E);
}

int vd_batch_relax(vd_batch *b) {
    if (!b) return fail(VD_EINVAL, "NULL handle");
    VD_TRY(engine_relax(b->E));
    // the relaxed points are now the device input of every member; the host copy is stale
    b->staged.clear();
    b->uploaded = true;
    return VD_OK;
}

size_t vd_batch_count(const vd_batch *b) { return b ? b->E.mem.size() : 0; }

int vd_batch_member_info(const vd_batch *b, uint32_t t, uint64_t info[5]) {
    if (!b || !info) return fail(VD_EINVAL, "NULL argument");
    const Member *m = member(b->E, t);
    if (!m) return VD_ERANGE;
    info[0] = b->E.prepared ? m->ngen : m->nraw;
    info[1] = (uint64_t)m->n[0];
    info[2] = (uint64_t)m->n[1];
    info[3] = (uint64_t)m->n[2];
    info[4] = m->gen_off;
    return VD_OK;
}

int vd_batch_get_volumes(const vd_batch *b, uint32_t t, double *out, size_t n) {
    if (!b || !out) return fail(VD_EINVAL, "NULL argument");
    VD_TRY(need_results(b->E));
    const Member *m = member(b->E, t);
    if (!m) return VD_ERANGE;
    return copy_slice(b->E.d_vol, 8, *m, out, n);
}

int vd_batch_get_centroids(const vd_batch *b, uint32_t t, double *out3, size_t n) {
    if (!b || !out3) return fail(VD_EINVAL, "NULL argument");
    VD_TRY(need_results(b->E));
    const Member *m = member(b->E, t);
    if (!m) return VD_ERANGE;
    return copy_slice(b->E.d_cen, 24, *m, out3, n);
}

int vd_batch_get_face_counts(const vd_batch *b, uint32_t t, uint32_t *out, size_t n) {
    if (!b || !out) return fail(VD_EINVAL, "NULL argument");
    VD_TRY(need_results(b->E));
    const Member *m = member(b->E, t);
    if (!m) return VD_ERANGE;
    return copy_slice(b->E.d_nface, 4, *m, out, n);
}

int vd_batch_get_generators(const vd_batch *b, uint32_t t, double *out, size_t n) {
    if (!b || !out) return fail(VD_EINVAL, "NULL argument");
    if (!b->E.prepared) return fail(VD_ESTATE, "the batch has not been calculated");
    const Member *m = member(b->E, t);
    if (!m) return VD_ERANGE;
    return copy_slice(b->E.d_gen, 24, *m, out, n);
}

int vd_batch_get_neighbors(const vd_batch *b, uint32_t t, uint64_t *offsets, size_t n, int32_t *ids, size_t cap) {
    if (!b) return fail(VD_EINVAL, "NULL handle");
    const Member *m = member(b->E, t);
    if (!m) return VD_ERANGE;
    return get_neighbors(b->E, *m, offsets, n, ids, cap);
}

int vd_batch_get_cell(const vd_batch *b, uint32_t t, size_t i, uint64_t sizes[3], double *vertices,
                      uint32_t *face_counts, uint32_t *face_indices, int32_t *face_neighbors) {
    if (!b) return fail(VD_EINVAL, "NULL handle");
    const Member *m = member(b->E, t);
    if (!m) return VD_ERANGE;
    return get_cell(b->E, *m, i, sizes, vertices, face_counts, face_indices, face_neighbors);
}

int vd_batch_set_option(vd_batch *b, int key, int64_t value) {
    if (!b) return fail(VD_EINVAL, "NULL handle");
    if (key == VD_BOPT_KEEP_CELLS) {
        b->E.keep_cells = value ? 1 : 0;
        b->E.calculated = false;
        return VD_OK;
    }
    if (key == VD_BOPT_PER_BIN) {
        if (value <= 0) return fail(VD_EINVAL, "generators per bin must be positive");
        b->per_bin = (double)value / 1000.0;
        return VD_OK;
    }
    if (key == VD_BOPT_GRAPHS) {
        b->E.graphs = value ? 1 : 0;
        return VD_OK;
    }
    if (key == VD_BOPT_FAR_LIST) {
        if (value < 0) return fail(VD_EINVAL, "walk length must not be negative");
        b->E.far_list_min = (double)value;
        return VD_OK;
    }
    if (key == VD_BOPT_T0_WAVES) return engine_set_launch(b->E, kKnobT0Waves, value);
    if (key == VD_BOPT_QUEUE_GRAIN) return engine_set_launch(b->E, kKnobGrain, value);
    if (key == VD_BOPT_FIRST_TIER) return engine_set_launch(b->E, kKnobFirstTier, value);
    if (key == VD_BOPT_FIRST_PROBE) return engine_set_launch(b->E, kKnobFirstProbe, value);
    return fail(VD_EINVAL, "unknown option");
}

int vd_batch_get_stats(const vd_batch *b, double out[10]) {
    if (!b || !out) return fail(VD_EINVAL, "NULL argument");
    engine_stats(b->E, out);
    return VD_OK;
}

int vd_batch_autotune(vd_batch *b, const double *per_bin, int ncand, double *best) {
    if (!b || !per_bin || ncand <= 0) return fail(VD_EINVAL, "invalid argument");
    Engine &E = b->E;
    if (!b->uploaded) VD_TRY(vd_batch_upload(b));
    double best_t = std::numeric_limits<double>::infinity(), best_r = b->per_bin;
    for (int pass = 0; pass <= ncand; ++pass) {
        const double r = pass < ncand ? per_bin[pass] : best_r;   // the last pass applies the winner
        if (!(r > 0.0)) return fail(VD_EINVAL, "generators per bin must be positive");
        for (Member &m : E.mem)
            if (m.auto_grid) {
                grid_for(m.lo, m.hi, m.nraw, r, m.n);
                m.grid = make_grid(m.lo, m.hi, m.n);
            }
        E.shapes.clear();
        if (pass == ncand) break;
        VD_TRY(vd_batch_calculate(b));
        const double us = E.stats[5];
        if (us < best_t) { best_t = us; best_r = r; }
    }
    b->per_bin = best_r;
    E.prepared = E.calculated = false;
    if (best) *best = best_r;
    return VD_OK;
}

int vd_batch_autotune_launch(vd_batch *b, int picks[4]) {
    if (!b) return fail(VD_EINVAL, "NULL handle");
    if (b->E.mem.empty()) return fail(VD_ESTATE, "the batch has no members");
    return engine_tune_launch(b->E, [&](double *us) {
        VD_TRY(vd_batch_calculate(b));
        *us = b->E.stats[5];
        return VD_OK;
    }, picks);
}

// ---- vd_server -----------------------------------------------------------------------------
int vd_server_create(size_t hbm_budget, int max_slots, vd_server **out) {
    if (!out || hbm_budget == 0 || max_slots <= 0) return fail(VD_EINVAL, "invalid argument");
    vd_server *S = new vd_server();
    S->budget = hbm_budget;
    S->max_slots = max_slots;
    int rc = VD_OK;
    for (Lane &L : S->lanes) {
        L.acct = &S->acct;
        L.E.keep_cells = 0;
        L.E.keep_nb = false;      // results are volumes, centroids and generators
        L.E.side_streams = false;
        L.E.far_list_min = 0.0;   // measured: the served mix ran faster without the lists
        L.E.shell_cache_bytes = 0;   // tables of geometries no job of the step uses are freed
        L.E.global_slots = kServerGlobalSlots;
        if (!rc) rc = engine_init(L.E);
    }
    S->reserve = 2 * (size_t)kServerGlobalSlots * vd::tier_global_bytes();
    if (!rc && hbm_budget <= S->reserve)
        rc = fail(VD_EINVAL, "HBM budget below the server's fixed reserve of " + std::to_string(S->reserve) + " bytes");
    if (rc) {
        for (Lane &L : S->lanes) engine_free(L.E);
        delete S;
        return rc;
    }
    *out = S;
    return VD_OK;
}

void vd_server_destroy(vd_server *S) {
    if (!S) return;
    for (Lane &L : S->lanes)
        if (L.worker.joinable()) L.worker.join();
    for (auto &kv : S->jobs) { release(kv.second->pts); delete kv.second; }
    for (DevBuf &b : S->spare) release(b);
    for (DevBuf &b : S->deferred) release(b);
    for (Lane &L : S->lanes) engine_free(L.E);
    delete S;
}

int vd_server_job_grid(const vd_server *S, uint64_t job, uint32_t out[3]) {
    if (!S || !out) return fail(VD_EINVAL, "NULL argument");
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return fail(VD_ERANGE, "unknown job");
    for (int a = 0; a < 3; ++a) out[a] = (uint32_t)it->second->m.n[a];
    return VD_OK;
}

int vd_server_memory(const vd_server *S, uint64_t out[4]) {
    if (!S || !out) return fail(VD_EINVAL, "NULL argument");
    out[0] = (uint64_t)std::max<int64_t>(0, S->acct.bytes.load());
    out[1] = (uint64_t)std::max<int64_t>(0, S->acct.peak.load());
    out[2] = S->reserve;
    out[3] = S->budget;
    return VD_OK;
}

int vd_server_submit(vd_server *S, const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny,
                     uint32_t nz, const vd_wall *walls, int nwalls, const double *xyz, size_t n,
                     uint32_t iterations, uint64_t *job) {
    if (!S || !bmin || !bmax || (!xyz && n) || (!walls && nwalls > 0) || !job) return fail(VD_EINVAL, "NULL argument");
    if (nwalls < 0 || nwalls > vd::kMaxWalls) return fail(VD_EINVAL, "0 to 8 walls per job");
    if (iterations == 0) return fail(VD_EINVAL, "a job runs at least one calculate()");
    Job *J = new Job();
    for (int a = 0; a < 3; ++a) { J->m.lo[a] = bmin[a]; J->m.hi[a] = bmax[a]; }
    int rc = VD_OK;
    if (nx == 0 && ny == 0 && nz == 0) {
        density_grid(bmin, bmax, xyz, n, S->per_bin, iterations > 1, J->m.n);
        J->m.auto_grid = true;
    } else {
        rc = check_grid(bmin, bmax, nx, ny, nz);
        J->m.n[0] = nx; J->m.n[1] = ny; J->m.n[2] = nz;
    }
    for (int k = 0; k < nwalls && rc == VD_OK; ++k) {
        vd::WallDev d;
        std::vector<double> data;
        rc = make_wall(walls[k], d, data);
        J->m.walls.push_back(d);
        J->m.wdata.push_back(std::move(data));
    }
    if (rc) { delete J; return rc; }
    J->m.grid = make_grid(J->m.lo, J->m.hi, J->m.n);
    J->iters = iterations;
    J->host.assign(xyz, xyz + 3 * n);
    J->npts = n;
    J->lane = n <= S->small_max ? 1 : 0;
    J->bytes = job_bytes(J->m, n);
    J->t_submit = now_s();
    J->id = S->next_id++;
    if (J->bytes + S->reserve > S->budget) {
        J->state = 3;   // can never be admitted: refuse it now instead of blocking the queue
        S->jobs[J->id] = J;
        *job = J->id;
        return fail(VD_ENOMEM, "job needs more device memory than the server's HBM budget");
    }
    S->jobs[J->id] = J;
    S->pending.push_back(J);
    *job = J->id;
    return VD_OK;
}

// One step of a lane: upload the points of newly admitted jobs, one batched calculate over the
// lane's jobs, the results of the jobs whose last iteration this was copied to their host buffers,
// the others' generators relaxed in place (tessellation.rs:482-495; their next input never leaves
// the device). Touches only the lane's engine and jobs, so it can run on a worker thread.
int lane_run(Lane &L) {
    const AcctScope scope(L.acct);
    Engine &E = L.E;
    for (Job *J : L.active)
        if (!J->uploaded && J->npts)
            VD_HIP(hipMemcpyAsync(J->pts.p, J->host.data(), J->npts * 24, hipMemcpyHostToDevice, E.stream));
    E.mem.clear();
    for (Job *J : L.active) {
        Member m = J->m;
        m.nraw = J->npts;
        E.mem.push_back(m);
    }
    E.shapes.clear();
    engine_layout(E);
    VD_TRY(ensure(E.d_raw, (E.nraw + 1) * 24));
    for (size_t t = 0; t < L.active.size(); ++t) {
        const Job *J = L.active[t];
        if (J->npts)
            VD_HIP(hipMemcpyAsync(E.d_raw.as<double>() + 3 * E.mem[t].raw_off, J->pts.p, J->npts * 24,
                                  hipMemcpyDeviceToDevice, E.stream));
    }
    VD_TRY(engine_prepare(E));
    VD_TRY(engine_calculate(E));
    bool any_more = false;
    for (const Job *J : L.active) any_more |= J->done + 1 < J->iters;
    if (any_more) {
        VD_TRY(ensure(E.d_next, (E.ngen + 1) * 24));
        VD_HIP(vd::launch_relax(E.d_gen.as<double>(), E.d_cen.as<double>(), E.d_nvert.as<uint32_t>(), E.ngen,
                                E.d_next.as<double>(), E.stream));
    }
    L.cells.assign(L.active.size(), 0);
    for (size_t k = 0; k < L.active.size(); ++k) {
        Job *J = L.active[k];
        const Member &m = E.mem[k];
        L.cells[k] = m.ngen;
        if (J->done + 1 >= J->iters) {
            J->vol.resize(m.ngen);
            J->cen.resize(3 * m.ngen);
            J->gen.resize(3 * m.ngen);
            if (m.ngen) {
                VD_HIP(hipMemcpyAsync(J->vol.data(), E.d_vol.as<double>() + m.gen_off, m.ngen * 8,
                                      hipMemcpyDeviceToHost, E.stream));
                VD_HIP(hipMemcpyAsync(J->cen.data(), E.d_cen.as<double>() + 3 * m.gen_off, m.ngen * 24,
                                      hipMemcpyDeviceToHost, E.stream));
                VD_HIP(hipMemcpyAsync(J->gen.data(), E.d_gen.as<double>() + 3 * m.gen_off, m.ngen * 24,
                                      hipMemcpyDeviceToHost, E.stream));
            }
        } else if (m.ngen) {
            VD_HIP(hipMemcpyAsync(J->pts.p, E.d_next.as<double>() + 3 * m.gen_off, m.ngen * 24,
                                  hipMemcpyDeviceToDevice, E.stream));
        }
    }
    VD_HIP(hipStreamSynchronize(E.stream));
    for (size_t k = 0; k < L.active.size(); ++k) {
        Job *J = L.active[k];
        if (!J->uploaded) { J->host.clear(); J->host.shrink_to_fit(); J->uploaded = true; }
        J->npts = L.cells[k];
    }
    L.step_cells = E.ngen;
    return VD_OK;
}

// Applies a finished lane step on the calling thread: job states and counters, finished jobs leave
// the lane and the budget, their point buffers go to the spare list.
void lane_collect(vd_server *S, Lane &L, uint32_t *finished) {
    const double t = now_s();
    std::vector<Job *> keep;
    for (size_t k = 0; k < L.active.size(); ++k) {
        Job *J = L.active[k];
        J->done += 1;
        J->cells = L.cells[k];
        J->total_cells += L.cells[k];
        if (J->done >= J->iters) {
            J->state = 2;
            J->t_done = t;
            S->spare.push_back(J->pts);
            J->pts = DevBuf();
            S->used -= J->bytes;
            S->nslots -= 1;
            if (finished) *finished += 1;
        } else {
            keep.push_back(J);
        }
    }
    L.active.swap(keep);
    S->steps += 1;
    S->cells += L.step_cells;
}

// A device buffer of at least `bytes` for a job's points: the smallest spare one that fits, else a
// new allocation.
int take_buffer(vd_server *S, size_t bytes, DevBuf &out) {
    const AcctScope scope(&S->acct);
    size_t best = S->spare.size();
    for (size_t k = 0; k < S->spare.size(); ++k)
        if (S->spare[k].bytes >= bytes && (best == S->spare.size() || S->spare[k].bytes < S->spare[best].bytes))
            best = k;
    if (best < S->spare.size()) {
        out = S->spare[best];
        S->spare.erase(S->spare.begin() + (long)best);
        return VD_OK;
    }
    return ensure(out, bytes);
}

// The large lane's step finished on its worker: join it and apply it.
int collect_large(vd_server *S, uint32_t *finished) {
    Lane &L = S->lanes[0];
    if (!L.running.load() || !L.ready.load()) return VD_OK;
    L.worker.join();
    L.running.store(false);
    L.ready.store(false);
    if (L.rc) return fail(L.rc, L.err);
    lane_collect(S, L, finished);
    return VD_OK;
}

// One server step: admit pending jobs (FIFO, while slots and the HBM budget allow) to the lane of
// their size, collect the large lane's step if it finished and start its next one on the worker
// (new large jobs join there), then run one step of the small lane on the calling thread. With no
// small work it waits up to a few milliseconds for the large lane, so a polling caller does not spin.
int vd_server_step(vd_server *S, uint32_t *finished) {
    if (!S) return fail(VD_EINVAL, "NULL handle");
    if (finished) *finished = 0;
    // admission: the admitted jobs' bounds plus the reserve, or what the server holds now if that
    // is more (buffers an idle lane kept from earlier steps, spare point buffers), must leave room
    // for the next job's bound. When they do not, an idle lane's buffers and the spare buffers are
    // freed first (only while no large step runs: hipFree waits for the whole device).
    auto committed = [S]() {
        return std::max<int64_t>((int64_t)(S->used + S->reserve), S->acct.bytes.load());
    };
    if (!S->pending.empty() && committed() + (int64_t)S->pending.front()->bytes > (int64_t)S->budget &&
        !S->lanes[0].running.load()) {
        for (Lane &L : S->lanes)
            if (L.active.empty() && (&L != &S->lanes[0] || S->admitted0.empty())) engine_release_buffers(L.E);
        for (DevBuf &b : S->spare) release(b);
        S->spare.clear();
        for (DevBuf &b : S->deferred) release(b);
        S->deferred.clear();
    }
    while (!S->pending.empty() && (int)S->nslots < S->max_slots &&
           committed() + (int64_t)S->pending.front()->bytes <= (int64_t)S->budget) {
        Job *J = S->pending.front();
        S->pending.pop_front();
        VD_TRY(take_buffer(S, (J->npts + 1) * 24, J->pts));
        J->state = 1;
        J->t_admit = now_s();
        S->used += J->bytes;
        S->nslots += 1;
        if (J->lane == 1) S->lanes[1].active.push_back(J);
        else S->admitted0.push_back(J);
    }
    VD_TRY(collect_large(S, finished));
    Lane &L0 = S->lanes[0];
    if (!L0.running.load()) {
        // no step in flight: deferred memory and spare buffers beyond kSpareBuffers can be freed
        for (DevBuf &b : S->deferred) release(b);
        S->deferred.clear();
        while (S->spare.size() > kSpareBuffers) { release(S->spare.front()); S->spare.erase(S->spare.begin()); }
    }
    if (!L0.running.load() && (!L0.active.empty() || !S->admitted0.empty())) {
        L0.active.insert(L0.active.end(), S->admitted0.begin(), S->admitted0.end());
        S->admitted0.clear();
        L0.E.reserve_cus = S->small_max ? kReservedCUs : 0;
        L0.running.store(true);
        L0.ready.store(false);
        L0.worker = std::thread([&L0]() {
            L0.rc = lane_run(L0);
            if (L0.rc) L0.err = vd_last_error();
            L0.ready.store(true);
        });
    }
    Lane &L1 = S->lanes[1];
    if (!L1.active.empty()) {
        g_defer = &S->deferred;        // a running large step must not hold up the small lane
        const int rc = lane_run(L1);
        g_defer = nullptr;
        VD_TRY(rc);
        lane_collect(S, L1, finished);
    } else if (L0.running.load()) {
        for (int k = 0; k < 20 && !L0.ready.load(); ++k) std::this_thread::sleep_for(std::chrono::microseconds(250));
    }
    VD_TRY(collect_large(S, finished));
    return VD_OK;
}

int vd_server_pending(const vd_server *S, uint64_t counts[3]) {
    if (!S || !counts) return fail(VD_EINVAL, "NULL argument");
    counts[0] = S->pending.size();
    counts[1] = S->nslots;
    counts[2] = S->used;
    return VD_OK;
}

int vd_server_job_info(const vd_server *S, uint64_t job, int32_t *state, uint64_t info[3], double times[3]) {
    if (!S) return fail(VD_EINVAL, "NULL handle");
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return fail(VD_ERANGE, "unknown job");
    const Job *J = it->second;
    if (state) *state = J->state;
    if (info) { info[0] = J->cells; info[1] = J->done; info[2] = J->total_cells; }
    if (times) { times[0] = J->t_submit; times[1] = J->t_admit; times[2] = J->t_done; }
    return VD_OK;
}

int vd_server_result(const vd_server *S, uint64_t job, double *volumes, double *centroids, double *generators,
                     size_t n) {
    if (!S) return fail(VD_EINVAL, "NULL handle");
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return fail(VD_ERANGE, "unknown job");
    const Job *J = it->second;
    if (J->state != 2) return fail(VD_ESTATE, "job not finished");
    if (n < J->cells) return fail(VD_EINVAL, "output buffer smaller than the cell count");
    if (volumes) std::memcpy(volumes, J->vol.data(), J->cells * 8);
    if (centroids) std::memcpy(centroids, J->cen.data(), J->cells * 24);
    if (generators) std::memcpy(generators, J->gen.data(), J->cells * 24);
    return VD_OK;
}

int vd_server_release(vd_server *S, uint64_t job) {
    if (!S) return fail(VD_EINVAL, "NULL handle");
    auto it = S->jobs.find(job);
    if (it == S->jobs.end()) return fail(VD_ERANGE, "unknown job");
    Job *J = it->second;
    if (J->state == 0 || J->state == 1) return fail(VD_ESTATE, "job still queued or running");
    release(J->pts);
    delete J;
    S->jobs.erase(it);
    return VD_OK;
}

int vd_server_set_option(vd_server *S, int key, int64_t value) {
    if (!S) return fail(VD_EINVAL, "NULL handle");
    if (key == VD_SOPT_PER_BIN) {
        if (value <= 0) return fail(VD_EINVAL, "per_bin must be positive");
        S->per_bin = (double)value / 1000.0;
        return VD_OK;
    }
    if (key == VD_SOPT_SMALL_JOB) {
        if (value < 0) return fail(VD_EINVAL, "the small-job limit cannot be negative");
        S->small_max = (uint64_t)value;
        return VD_OK;
    }
    return fail(VD_EINVAL, "unknown option");
}

int vd_server_stats(const vd_server *S, double out[4]) {
    if (!S || !out) return fail(VD_EINVAL, "NULL argument");
    out[0] = (double)S->steps;
    out[1] = (double)S->cells;
    out[2] = (double)S->used;
    out[3] = (double)S->budget;
    return VD_OK;
}

}  // extern "C"
