// C ABI of libvoronoid_hip.so (declared in include/vd.h): handle state, device memory, the bin
// shell-order prefix, the generator pipeline (wall filter -> compaction -> binning) and the cell
// launches with their retry tiers.
#include "vd.h"
#include "vd_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

namespace {

constexpr int kTiers = 3;   // polytope capacity tiers of the cell kernel (vd_cells.hip)

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

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    template <class T> T *as() const { return static_cast<T *>(p); }
};

int ensure(DevBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return VD_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(VD_ENOMEM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
    }
    b.bytes = bytes;
    return VD_OK;
}

void release(DevBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

// ------------------------------------------------------------------------------------------
// Shell order prefix (Algorithm3DGrid::new, algo_3d_grid.rs:47-95, with a stable sort).
// ------------------------------------------------------------------------------------------
struct ShellPrefix {
    std::vector<vd::ShellEntry> e;
    double next = std::numeric_limits<double>::infinity();
    bool complete = false;
};

inline double axis_term(int64_t d, double cs) {  // algo_3d_grid.rs:61-66
    if (d > 0) return (double)(d - 1) * cs;
    if (d < 0) return (double)(-d - 1) * cs;
    return 0.0;
}

// All entries with key <= T, T = ((kmax - 1) * max cs)^2, in stable order; `next` is the smallest
// key outside the prefix.
void build_shell(const int64_t n[3], const double cs[3], uint32_t kmax, ShellPrefix &out) {
    const double csm = std::max(cs[0], std::max(cs[1], cs[2]));
    const double tm = (double)(kmax > 0 ? kmax - 1 : 0) * csm;
    const double T = tm * tm;
    int64_t K[3];
    for (int a = 0; a < 3; ++a) {
        int64_t k = 1;
        while (k < n[a]) {
            const double m = axis_term(k + 1, cs[a]);
            if (m * m > T) break;
            ++k;
        }
        K[a] = std::min<int64_t>(k, n[a]);
    }
    struct Tmp { double d2; uint64_t gen; int16_t dx, dy, dz; };
    std::vector<Tmp> all;
    all.reserve((size_t)(2 * K[0] + 1) * (2 * K[1] + 1) * (2 * K[2] + 1));
    double next = std::numeric_limits<double>::infinity();
    const uint64_t wx = 2 * (uint64_t)n[0] + 1, wy = 2 * (uint64_t)n[1] + 1;
    for (int64_t z = -K[2]; z <= K[2]; ++z)
        for (int64_t y = -K[1]; y <= K[1]; ++y)
            for (int64_t x = -K[0]; x <= K[0]; ++x) {
                const double mx = axis_term(x, cs[0]), my = axis_term(y, cs[1]), mz = axis_term(z, cs[2]);
                const double d2 = mx * mx + my * my + mz * mz;
                if (d2 <= T) {
                    const uint64_t g = ((uint64_t)(z + n[2]) * wy + (uint64_t)(y + n[1])) * wx + (uint64_t)(x + n[0]);
                    all.push_back({d2, g, (int16_t)x, (int16_t)y, (int16_t)z});
                } else {
                    next = std::min(next, d2);
                }
            }
    // entries beyond the enumerated box: the smallest has one axis at K+1 and the others at 0
    for (int a = 0; a < 3; ++a)
        if (K[a] < n[a]) {
            const double m = axis_term(K[a] + 1, cs[a]);
            next = std::min(next, m * m);
        }
    std::sort(all.begin(), all.end(), [](const Tmp &p, const Tmp &q) {
        return p.d2 < q.d2 || (p.d2 == q.d2 && p.gen < q.gen);
    });
    out.e.resize(all.size());
    for (size_t k = 0; k < all.size(); ++k) out.e[k] = vd::ShellEntry{all[k].d2, all[k].dx, all[k].dy, all[k].dz, 0};
    out.next = next;
    out.complete = std::isinf(next);
}

}  // namespace

struct vd_tess {
    double lo[3], hi[3];
    int64_t n[3];
    uint64_t nbins = 0;
    vd::GridGeom grid{};
    std::vector<vd::WallDev> walls;
    int device = 0;
    int cus = 1;
    int occ[kTiers] = {1, 1, 1};
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int keep_cells = 1;

    // shell prefix
    uint32_t shell_k = 0;
    ShellPrefix shell;
    DevBuf d_shell;

    // generators and bins
    size_t ngen = 0;
    DevBuf d_gen, d_raw, d_flags, d_pos, d_scan;
    DevBuf d_gen_bin, d_counts, d_bin_start, d_fill, d_items, d_spos;

    // results
    bool calculated = false;
    size_t ncells = 0;
    DevBuf d_vol, d_cen, d_nface, d_nvert, d_status, d_cell_off;
    DevBuf d_vtx, d_fcnt, d_fnb, d_fidx, d_top, d_queue, d_counts4;
    uint64_t cap[3] = {0, 0, 0};
    double stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

namespace {

int scan(vd_tess *t, const uint32_t *in, uint32_t *out, uint32_t n) {
    size_t words = vd::scan_tmp_words(n);
    int rc = ensure(t->d_scan, words * 4);
    if (rc) return rc;
    VD_HIP(vd::launch_scan_u32(in, out, n, t->d_scan.as<uint32_t>(), words, t->stream));
    return VD_OK;
}

// Wall filter of n points in d_src (input order kept), result in t->d_gen; t->ngen updated.
int filter_into_gen(vd_tess *t, const double *d_src, size_t n) {
    int rc;
    if (t->walls.empty()) {
        if ((rc = ensure(t->d_gen, n * 24))) return rc;
        if (n) VD_HIP(hipMemcpyAsync(t->d_gen.p, d_src, n * 24, hipMemcpyDeviceToDevice, t->stream));
        t->ngen = n;
        return VD_OK;
    }
    if ((rc = ensure(t->d_flags, (n + 1) * 4)) || (rc = ensure(t->d_pos, (n + 1) * 4))) return rc;
    VD_HIP(vd::launch_fill_u32(t->d_flags.as<uint32_t>(), 0u, n + 1, t->stream));
    VD_HIP(vd::launch_wall_flags(d_src, (uint32_t)n, t->walls.data(), (int)t->walls.size(),
                                 t->d_flags.as<uint32_t>(), t->stream));
    if ((rc = scan(t, t->d_flags.as<uint32_t>(), t->d_pos.as<uint32_t>(), (uint32_t)(n + 1)))) return rc;
    uint32_t kept = 0;
    VD_HIP(hipMemcpyAsync(&kept, t->d_pos.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, t->stream));
    VD_HIP(hipStreamSynchronize(t->stream));
    DevBuf out;
    if (d_src == t->d_gen.p) {
        if ((rc = ensure(out, (size_t)kept * 24))) return rc;
    } else {
        if ((rc = ensure(t->d_gen, (size_t)kept * 24))) return rc;
    }
    double *dst = (d_src == t->d_gen.p) ? out.as<double>() : t->d_gen.as<double>();
    VD_HIP(vd::launch_compact3(d_src, t->d_flags.as<uint32_t>(), t->d_pos.as<uint32_t>(), (uint32_t)n,
                               dst, t->stream));
    if (d_src == t->d_gen.p) {
        VD_HIP(hipStreamSynchronize(t->stream));
        release(t->d_gen);
        t->d_gen = out;
    }
    t->ngen = kept;
    return VD_OK;
}

// Algorithm3DGrid::set_generators (algo_3d_grid.rs:111-130) on the device.
int rebin(vd_tess *t) {
    int rc;
    const size_t n = t->ngen;
    const uint64_t nb = t->nbins;
    if ((rc = ensure(t->d_counts, (nb + 1) * 4)) || (rc = ensure(t->d_bin_start, (nb + 1) * 4)) ||
        (rc = ensure(t->d_fill, (nb + 1) * 4)) || (rc = ensure(t->d_gen_bin, (n + 1) * 4)) ||
        (rc = ensure(t->d_items, (n + 1) * 4)) || (rc = ensure(t->d_spos, (n + 1) * 32)))
        return rc;
    VD_HIP(vd::launch_fill_u32(t->d_counts.as<uint32_t>(), 0u, nb + 1, t->stream));
    VD_HIP(vd::launch_fill_u32(t->d_fill.as<uint32_t>(), 0u, nb + 1, t->stream));
    VD_HIP(vd::launch_bin_count(t->d_gen.as<double>(), (uint32_t)n, t->grid, t->d_gen_bin.as<uint32_t>(),
                                t->d_counts.as<uint32_t>(), t->stream));
    if ((rc = scan(t, t->d_counts.as<uint32_t>(), t->d_bin_start.as<uint32_t>(), (uint32_t)(nb + 1)))) return rc;
    VD_HIP(vd::launch_bin_scatter(t->d_gen_bin.as<uint32_t>(), (uint32_t)n, t->d_bin_start.as<uint32_t>(),
                                  t->d_fill.as<uint32_t>(), t->d_items.as<uint32_t>(), t->stream));
    VD_HIP(vd::launch_bin_sort(t->d_bin_start.as<uint32_t>(), (uint32_t)nb, t->d_items.as<uint32_t>(), t->stream));
    VD_HIP(vd::launch_fill_spos(t->d_gen.as<double>(), t->d_items.as<uint32_t>(), (uint32_t)n,
                                t->d_spos.as<double4>(), t->stream));
    t->calculated = false;
    t->ncells = 0;
    return VD_OK;
}

int upload_shell(vd_tess *t, uint32_t k) {
    build_shell(t->n, t->grid.cs, k, t->shell);
    t->shell_k = k;
    int rc = ensure(t->d_shell, t->shell.e.size() * sizeof(vd::ShellEntry));
    if (rc) return rc;
    VD_HIP(hipMemcpyAsync(t->d_shell.p, t->shell.e.data(), t->shell.e.size() * sizeof(vd::ShellEntry),
                          hipMemcpyHostToDevice, t->stream));
    VD_HIP(hipStreamSynchronize(t->stream));
    return VD_OK;
}

int add_wall(vd_tess *t, const vd::WallDev &w) {
    if (w.id > VD_WALL_ID_MAX) return fail(VD_EINVAL, "Wall ID must be <= -1000");
    if ((int)t->walls.size() >= vd::kMaxWalls) return fail(VD_EINVAL, "too many walls (max 8)");
    t->walls.push_back(w);
    // prune_outside_generators (tessellation.rs:175-194): re-filter, re-bin if the count changed
    if (t->ngen == 0) return VD_OK;
    const size_t before = t->ngen;
    int rc = filter_into_gen(t, t->d_gen.as<double>(), before);
    if (rc) return rc;
    if (t->ngen != before) return rebin(t);
    return VD_OK;
}

int run_cells(vd_tess *t, int tier, const uint32_t *d_list, uint32_t nlist, float *ms) {
    vd::TessArgs a{};
    a.grid = t->grid;
    a.gen = t->d_gen.as<double>();
    a.spos = t->d_spos.as<double4>();
    a.bin_start = t->d_bin_start.as<uint32_t>();
    a.shell = t->d_shell.as<vd::ShellEntry>();
    a.nshell = (uint32_t)t->shell.e.size();
    a.shell_next_d2 = t->shell.next;
    a.nwalls = (int32_t)t->walls.size();
    for (size_t k = 0; k < t->walls.size(); ++k) a.walls[k] = t->walls[k];
    for (int k = 0; k < 3; ++k) { a.box_lo[k] = t->lo[k]; a.box_hi[k] = t->hi[k]; }
    a.ngen = (uint32_t)t->ngen;
    a.cells = d_list;
    a.ncells = nlist;
    a.queue = t->d_queue.as<uint32_t>();
    a.vol = t->d_vol.as<double>();
    a.cen = t->d_cen.as<double>();
    a.nface = t->d_nface.as<uint32_t>();
    a.nvert = t->d_nvert.as<uint32_t>();
    a.status = t->d_status.as<int32_t>();
    a.vtx = t->keep_cells ? t->d_vtx.as<double>() : nullptr;
    a.fcnt = t->d_fcnt.as<uint8_t>();
    a.fnb = t->d_fnb.as<int32_t>();
    a.fidx = t->d_fidx.as<uint16_t>();
    a.arena_top = t->d_top.as<uint64_t>();
    for (int k = 0; k < 3; ++k) a.arena_cap[k] = t->cap[k];
    a.cell_off = t->d_cell_off.as<uint64_t>();
    a.fnb_only = 1;
    VD_HIP(vd::launch_fill_u32(t->d_queue.as<uint32_t>(), 0u, 1, t->stream));
    int waves = t->cus * t->occ[tier];
    if ((uint32_t)waves > nlist) waves = (int)nlist;
    if (ms) VD_HIP(hipEventRecord(t->ev0, t->stream));
    VD_HIP(vd::launch_cells(tier, a, waves, t->stream));
    if (ms) {
        VD_HIP(hipEventRecord(t->ev1, t->stream));
        VD_HIP(hipEventSynchronize(t->ev1));
        VD_HIP(hipEventElapsedTime(ms, t->ev0, t->ev1));
    }
    return VD_OK;
}

int calculate(vd_tess *t) {
    using clock = std::chrono::steady_clock;
    const auto t0 = clock::now();
    int rc;
    const size_t n = t->ngen;
    for (double &s : t->stats) s = 0.0;
    if ((rc = ensure(t->d_vol, n * 8)) || (rc = ensure(t->d_cen, n * 24)) || (rc = ensure(t->d_nface, n * 4)) ||
        (rc = ensure(t->d_nvert, n * 4)) || (rc = ensure(t->d_status, n * 4)) ||
        (rc = ensure(t->d_cell_off, n * 24)) || (rc = ensure(t->d_top, 64)) || (rc = ensure(t->d_queue, 64)) ||
        (rc = ensure(t->d_counts4, 64)))
        return rc;
    if (t->cap[1] == 0 || t->cap[1] < n * 16) {
        t->cap[0] = n * 40 + 4096;
        t->cap[1] = n * 24 + 4096;
        t->cap[2] = n * 120 + 4096;
    }
    t->calculated = false;
    t->ncells = n;
    if (n == 0) { t->calculated = true; return VD_OK; }
    if (t->shell.e.empty()) {
        if ((rc = upload_shell(t, 7))) return rc;
    }
    for (int attempt = 0; attempt < 8; ++attempt) {
        if (t->keep_cells) {
            if ((rc = ensure(t->d_vtx, t->cap[0] * 24)) || (rc = ensure(t->d_fcnt, t->cap[1])) ||
                (rc = ensure(t->d_fidx, t->cap[2] * 2)))
                return rc;
        }
        if ((rc = ensure(t->d_fnb, t->cap[1] * 4))) return rc;
        VD_HIP(vd::launch_fill_u32(reinterpret_cast<uint32_t *>(t->d_top.p), 0u, 6, t->stream));

        // Tier 0 (register fast path, small LDS) over every cell in bin order. A cell that
        // outgrows its tier moves to the next one; a cell whose walk runs past the shell prefix
        // reruns in its tier after the prefix is extended. Reruns start from scratch and give
        // the same result.
        constexpr int NT = kTiers;
        DevBuf L[NT];                              // owned work lists per tier
        const uint32_t *list[NT] = {t->d_items.as<uint32_t>()};
        uint32_t nlist[NT] = {(uint32_t)n};
        bool arena_full = false;
        bool first = true;
        auto pending = [&]() { for (int k = 0; k < NT; ++k) if (nlist[k]) return true; return false; };
        for (int round = 0; round < 64 && pending(); ++round) {
            for (int tier = 0; tier < NT; ++tier) {
                if (!nlist[tier]) continue;
                float ms = 0.f;
                if ((rc = run_cells(t, tier, list[tier], nlist[tier], first ? &ms : nullptr))) return rc;
                if (first) { t->stats[4] = 1000.0 * ms; t->stats[0] = (double)nlist[0]; first = false; }
                else if (tier > 0) t->stats[1] += nlist[tier];
            }
            // outcome per tier: counts[4*tier + {0 capacity, 1 shell, 2 arena, 3 error}]
            uint32_t counts[4 * NT] = {};
            DevBuf out[2 * NT];   // per tier: capacity list, shell list
            VD_HIP(vd::launch_fill_u32(t->d_counts4.as<uint32_t>(), 0u, 4 * NT, t->stream));
            for (int tier = 0; tier < NT; ++tier) {
                if (!nlist[tier]) continue;
                if ((rc = ensure(out[2 * tier], (size_t)nlist[tier] * 4)) ||
                    (rc = ensure(out[2 * tier + 1], (size_t)nlist[tier] * 4)))
                    return rc;
                VD_HIP(vd::launch_collect(t->d_status.as<int32_t>(), list[tier], nlist[tier],
                                          out[2 * tier].as<uint32_t>(), out[2 * tier + 1].as<uint32_t>(),
                                          t->d_counts4.as<uint32_t>() + 4 * tier, t->stream));
            }
            VD_HIP(hipMemcpyAsync(counts, t->d_counts4.p, sizeof(counts), hipMemcpyDeviceToHost, t->stream));
            VD_HIP(hipStreamSynchronize(t->stream));
            bool shell = false;
            for (int tier = 0; tier < NT; ++tier) {
                if (counts[4 * tier + 3]) return fail(VD_ECAPACITY, "a cell has a face with 256 or more vertices");
                if (counts[4 * tier + 2]) arena_full = true;
                shell |= counts[4 * tier + 1] != 0;
            }
            if (arena_full) break;
            if (counts[4 * (NT - 1)]) return fail(VD_ECAPACITY, "a cell exceeds the largest polytope tier");
            if (shell) {
                if (t->shell.complete) return fail(VD_EDEVICE, "shell walk past a complete table");
                if ((rc = upload_shell(t, t->shell_k * 2))) return rc;
                t->stats[2] += 1;
            }
            // next list of tier k: its own shell failures + the capacity failures of tier k-1
            DevBuf next[NT];
            uint32_t nnext[NT] = {};
            for (int tier = 0; tier < NT; ++tier) {
                const uint32_t ns = counts[4 * tier + 1];
                const uint32_t nc = tier > 0 ? counts[4 * (tier - 1)] : 0u;
                nnext[tier] = ns + nc;
                if (!nnext[tier]) continue;
                if ((rc = ensure(next[tier], (size_t)nnext[tier] * 4))) return rc;
                if (ns)
                    VD_HIP(hipMemcpyAsync(next[tier].p, out[2 * tier + 1].p, ns * 4, hipMemcpyDeviceToDevice,
                                          t->stream));
                if (nc)
                    VD_HIP(hipMemcpyAsync(next[tier].as<uint32_t>() + ns, out[2 * (tier - 1)].p, nc * 4,
                                          hipMemcpyDeviceToDevice, t->stream));
            }
            VD_HIP(hipStreamSynchronize(t->stream));
            for (int tier = 0; tier < NT; ++tier) {
                release(L[tier]);
                L[tier] = next[tier];
                next[tier] = DevBuf();
                list[tier] = L[tier].as<uint32_t>();
                nlist[tier] = nnext[tier];
            }
            for (DevBuf &b : out) release(b);
        }
        for (DevBuf &b : L) release(b);
        if (!arena_full) break;
        for (int k = 0; k < 3; ++k) t->cap[k] *= 2;
    }
    VD_HIP(hipStreamSynchronize(t->stream));
    t->calculated = true;
    t->stats[3] = (double)t->shell.e.size();
    t->stats[5] = std::chrono::duration<double, std::micro>(clock::now() - t0).count();
    return VD_OK;
}

int need_results(const vd_tess *t, size_t n) {
    if (!t->calculated) return fail(VD_ESTATE, "calculate() has not run since the generators changed");
    if (n < t->ncells) return fail(VD_EINVAL, "output buffer smaller than count_cells()");
    return VD_OK;
}

}  // namespace

extern "C" {

const char *vd_last_error(void) { return g_err.c_str(); }

int vd_shell_prefix(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny, uint32_t nz,
                    uint32_t kmax, int32_t *dxyz, double *keys, size_t cap, size_t *count,
                    double *next_key) {
    if (!bmin || !bmax || nx == 0 || ny == 0 || nz == 0) return fail(VD_EINVAL, "invalid grid");
    const int64_t n[3] = {nx, ny, nz};
    double cs[3];
    for (int a = 0; a < 3; ++a) {
        const double s = (double)n[a] / (bmax[a] - bmin[a]);
        cs[a] = 1.0 / s;
    }
    ShellPrefix sp;
    build_shell(n, cs, kmax, sp);
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

int vd_tess_create(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny, uint32_t nz,
                   vd_tess **out) {
    if (!out) return fail(VD_EINVAL, "out is NULL");
    *out = nullptr;
    if (!bmin || !bmax || nx == 0 || ny == 0 || nz == 0) return fail(VD_EINVAL, "invalid grid");
    if (nx > 16384 || ny > 16384 || nz > 16384) return fail(VD_EINVAL, "grid resolution above 16384");
    const uint64_t nb = (uint64_t)nx * ny * nz;
    if (nb >= 0xffffffffull) return fail(VD_EINVAL, "too many bins");
    vd_tess *t = new vd_tess();
    for (int a = 0; a < 3; ++a) { t->lo[a] = bmin[a]; t->hi[a] = bmax[a]; }
    t->n[0] = nx; t->n[1] = ny; t->n[2] = nz;
    t->nbins = nb;
    for (int a = 0; a < 3; ++a) {
        t->grid.lo[a] = bmin[a];
        t->grid.s[a] = (double)t->n[a] / (bmax[a] - bmin[a]);   // algo_3d_grid.rs:48-50
        t->grid.cs[a] = 1.0 / t->grid.s[a];                     // :52-54, :161-163
        t->grid.lim[a] = (double)t->n[a] - 1e-5;                // :85-87
        t->grid.n[a] = (int32_t)t->n[a];
    }
    hipError_t e = hipGetDevice(&t->device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&t->cus, hipDeviceAttributeMultiprocessorCount, t->device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&t->ev0);
    if (e == hipSuccess) e = hipEventCreate(&t->ev1);
    for (int k = 0; k < kTiers && e == hipSuccess; ++k) e = vd::cells_occupancy(k, &t->occ[k]);
    if (e != hipSuccess) {
        vd_tess_destroy(t);
        return fail(VD_EDEVICE, std::string("device setup: ") + hipGetErrorString(e));
    }
    for (int k = 0; k < kTiers; ++k) t->occ[k] = std::max(1, t->occ[k]);
    int rc = ensure(t->d_gen, 16);
    if (!rc) rc = rebin(t);
    if (!rc) rc = hipStreamSynchronize(t->stream) == hipSuccess ? VD_OK : fail(VD_EDEVICE, "sync");
    if (rc) { vd_tess_destroy(t); return rc; }
    *out = t;
    return VD_OK;
}

void vd_tess_destroy(vd_tess *t) {
    if (!t) return;
    if (t->stream) (void)hipStreamSynchronize(t->stream);
    DevBuf *bufs[] = {&t->d_shell, &t->d_gen, &t->d_raw, &t->d_flags, &t->d_pos, &t->d_scan, &t->d_gen_bin,
                      &t->d_counts, &t->d_bin_start, &t->d_fill, &t->d_items, &t->d_spos, &t->d_vol,
                      &t->d_cen, &t->d_nface, &t->d_nvert, &t->d_status, &t->d_cell_off, &t->d_vtx,
                      &t->d_fcnt, &t->d_fnb, &t->d_fidx, &t->d_top, &t->d_queue, &t->d_counts4};
    for (DevBuf *b : bufs) release(*b);
    if (t->ev0) (void)hipEventDestroy(t->ev0);
    if (t->ev1) (void)hipEventDestroy(t->ev1);
    if (t->stream) (void)hipStreamDestroy(t->stream);
    delete t;
}

int vd_tess_add_wall_plane(vd_tess *t, int32_t id, const double point[3], const double normal[3]) {
    if (!t || !point || !normal) return fail(VD_EINVAL, "NULL argument");
    vd::WallDev w{};
    w.id = id;
    w.kind = 0;
    // PlaneGeometry::new, wall_3d.rs:23-27
    const double len = std::sqrt(normal[0] * normal[0] + normal[1] * normal[1] + normal[2] * normal[2]);
    if (len == 0.0) { w.b[0] = 0.0; w.b[1] = 0.0; w.b[2] = 1.0; }
    else { w.b[0] = normal[0] / len; w.b[1] = normal[1] / len; w.b[2] = normal[2] / len; }
    for (int k = 0; k < 3; ++k) w.a[k] = point[k];
    return add_wall(t, w);
}

int vd_tess_add_wall_sphere(vd_tess *t, int32_t id, const double center[3], double radius) {
    if (!t || !center) return fail(VD_EINVAL, "NULL argument");
    vd::WallDev w{};
    w.id = id;
    w.kind = 1;
    w.r = radius;
    for (int k = 0; k < 3; ++k) w.a[k] = center[k];
    return add_wall(t, w);
}

int vd_tess_clear_walls(vd_tess *t) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    t->walls.clear();
    t->calculated = false;
    return VD_OK;
}

int vd_tess_set_generators(vd_tess *t, const double *xyz, size_t n) {
    if (!t || (!xyz && n)) return fail(VD_EINVAL, "NULL argument");
    if (n >= 0x7fffffffull) return fail(VD_EINVAL, "too many generators");
    int rc = ensure(t->d_raw, n * 24);
    if (rc) return rc;
    if (n) VD_HIP(hipMemcpyAsync(t->d_raw.p, xyz, n * 24, hipMemcpyHostToDevice, t->stream));
    if ((rc = filter_into_gen(t, t->d_raw.as<double>(), n))) return rc;
    if ((rc = rebin(t))) return rc;
    VD_HIP(hipStreamSynchronize(t->stream));
    return VD_OK;
}

int vd_tess_set_generators_device(vd_tess *t, const double *d_xyz, size_t n) {
    if (!t || (!d_xyz && n)) return fail(VD_EINVAL, "NULL argument");
    if (n >= 0x7fffffffull) return fail(VD_EINVAL, "too many generators");
    int rc;
    if ((rc = filter_into_gen(t, d_xyz, n))) return rc;
    if ((rc = rebin(t))) return rc;
    VD_HIP(hipStreamSynchronize(t->stream));
    return VD_OK;
}

int vd_tess_calculate(vd_tess *t) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    return calculate(t);
}

int vd_tess_relax(vd_tess *t) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    if (!t->calculated) return fail(VD_ESTATE, "relax() needs calculate() first");
    const size_t n = t->ngen;
    int rc = ensure(t->d_raw, n * 24);
    if (rc) return rc;
    VD_HIP(vd::launch_relax(t->d_gen.as<double>(), t->d_cen.as<double>(), t->d_nvert.as<uint32_t>(),
                            (uint32_t)n, t->d_raw.as<double>(), t->stream));
    if ((rc = filter_into_gen(t, t->d_raw.as<double>(), n))) return rc;
    if ((rc = rebin(t))) return rc;
    VD_HIP(hipStreamSynchronize(t->stream));
    return VD_OK;
}

size_t vd_tess_count_generators(const vd_tess *t) { return t ? t->ngen : 0; }
size_t vd_tess_count_cells(const vd_tess *t) { return (t && t->calculated) ? t->ncells : 0; }

int vd_tess_get_generators(const vd_tess *t, double *out, size_t n) {
    if (!t || (!out && t->ngen)) return fail(VD_EINVAL, "NULL argument");
    if (n < t->ngen) return fail(VD_EINVAL, "output buffer smaller than count_generators()");
    if (t->ngen) VD_HIP(hipMemcpy(out, t->d_gen.p, t->ngen * 24, hipMemcpyDeviceToHost));
    return VD_OK;
}

int vd_tess_get_volumes(const vd_tess *t, double *out, size_t n) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    int rc = need_results(t, n);
    if (rc) return rc;
    if (t->ncells) VD_HIP(hipMemcpy(out, t->d_vol.p, t->ncells * 8, hipMemcpyDeviceToHost));
    return VD_OK;
}

int vd_tess_get_centroids(const vd_tess *t, double *out3, size_t n) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    int rc = need_results(t, n);
    if (rc) return rc;
    if (t->ncells) VD_HIP(hipMemcpy(out3, t->d_cen.p, t->ncells * 24, hipMemcpyDeviceToHost));
    return VD_OK;
}

int vd_tess_get_face_counts(const vd_tess *t, uint32_t *out, size_t n) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    int rc = need_results(t, n);
    if (rc) return rc;
    if (t->ncells) VD_HIP(hipMemcpy(out, t->d_nface.p, t->ncells * 4, hipMemcpyDeviceToHost));
    return VD_OK;
}

int vd_tess_get_neighbors(const vd_tess *t, uint64_t *offsets, size_t n, int32_t *ids, size_t cap) {
    if (!t || !offsets) return fail(VD_EINVAL, "NULL argument");
    int rc = need_results(t, n);
    if (rc) return rc;
    const size_t nc = t->ncells;
    std::vector<uint32_t> nf(nc);
    if (nc) VD_HIP(hipMemcpy(nf.data(), t->d_nface.p, nc * 4, hipMemcpyDeviceToHost));
    offsets[0] = 0;
    for (size_t i = 0; i < nc; ++i) offsets[i + 1] = offsets[i] + nf[i];
    if (!ids) return VD_OK;
    if (cap < offsets[nc]) return fail(VD_EINVAL, "ids buffer too small");
    std::vector<uint64_t> off(3 * nc);
    if (nc) VD_HIP(hipMemcpy(off.data(), t->d_cell_off.p, nc * 24, hipMemcpyDeviceToHost));
    uint64_t top = 0;
    VD_HIP(hipMemcpy(&top, t->d_top.as<uint64_t>() + 1, 8, hipMemcpyDeviceToHost));
    std::vector<int32_t> arena(top);
    if (top) VD_HIP(hipMemcpy(arena.data(), t->d_fnb.p, top * 4, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < nc; ++i)
        if (nf[i]) std::memcpy(ids + offsets[i], arena.data() + off[3 * i + 1], nf[i] * 4);
    return VD_OK;
}

int vd_tess_get_cell(const vd_tess *t, size_t i, uint64_t sizes[3], double *vertices, uint32_t *face_counts,
                     uint32_t *face_indices, int32_t *face_neighbors) {
    if (!t || !sizes) return fail(VD_EINVAL, "NULL argument");
    if (!t->calculated) return fail(VD_ESTATE, "calculate() has not run since the generators changed");
    if (i >= t->ncells) return fail(VD_ERANGE, "cell index out of range");
    if (!t->keep_cells) return fail(VD_ESTATE, "cells are not kept (VD_OPT_KEEP_CELLS = 0)");
    uint32_t nv = 0, nf = 0;
    uint64_t off[3];
    VD_HIP(hipMemcpy(&nv, t->d_nvert.as<uint32_t>() + i, 4, hipMemcpyDeviceToHost));
    VD_HIP(hipMemcpy(&nf, t->d_nface.as<uint32_t>() + i, 4, hipMemcpyDeviceToHost));
    VD_HIP(hipMemcpy(off, t->d_cell_off.as<uint64_t>() + 3 * i, 24, hipMemcpyDeviceToHost));
    std::vector<uint8_t> fc(nf);
    if (nf) VD_HIP(hipMemcpy(fc.data(), t->d_fcnt.as<uint8_t>() + off[1], nf, hipMemcpyDeviceToHost));
    uint64_t ns = 0;
    for (uint8_t c : fc) ns += c;
    sizes[0] = nv;
    sizes[1] = nf;
    sizes[2] = ns;
    if (vertices && nv) VD_HIP(hipMemcpy(vertices, t->d_vtx.as<double>() + 3 * off[0], nv * 24, hipMemcpyDeviceToHost));
    if (face_counts) for (uint32_t k = 0; k < nf; ++k) face_counts[k] = fc[k];
    if (face_neighbors && nf)
        VD_HIP(hipMemcpy(face_neighbors, t->d_fnb.as<int32_t>() + off[1], nf * 4, hipMemcpyDeviceToHost));
    if (face_indices && ns) {
        std::vector<uint16_t> fi(ns);
        VD_HIP(hipMemcpy(fi.data(), t->d_fidx.as<uint16_t>() + off[2], ns * 2, hipMemcpyDeviceToHost));
        for (uint64_t k = 0; k < ns; ++k) face_indices[k] = fi[k];
    }
    return VD_OK;
}

int vd_tess_set_option(vd_tess *t, int key, int64_t value) {
    if (!t) return fail(VD_EINVAL, "NULL handle");
    if (key == VD_OPT_KEEP_CELLS) {
        t->keep_cells = value ? 1 : 0;
        t->calculated = false;
        return VD_OK;
    }
    return fail(VD_EINVAL, "unknown option");
}

int vd_tess_get_stats(const vd_tess *t, double out[8]) {
    if (!t || !out) return fail(VD_EINVAL, "NULL argument");
    for (int k = 0; k < 8; ++k) out[k] = t->stats[k];
    return VD_OK;
}

}  // extern "C"
