// Internal interface between the C-ABI host code (vd_api.cpp) and the gfx950 kernels (*.hip).
// Not installed; the public interface is include/vd.h.
//
// Every launch works on a set of tessellations ("members"). A single Tessellation is a set of one.
// Generators, bins and outputs of all members are concatenated; a TessDesc per member holds its
// geometry, walls and offsets, and gen_tess maps each generator to its member.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace vd {

// One entry of the bin shell order of Algorithm3DGrid (src/algorithm/algo_3d_grid.rs:56-76): the
// offset (dx, dy, dz) packed in 32 bits as (dx + K0) | (dy + K1) << sy | (dz + K2) << sz with the
// prefix's enumeration box [-K, K]. Its key (the squared minimum bin distance) is recomputed from
// the offset with the expression the table was sorted by.
using ShellCode = uint32_t;

struct ShellPack {
    int16_t k[3];     // K per axis
    uint8_t sy, sz;   // bit positions of dy and dz
};

__host__ __device__ inline void shell_unpack(ShellCode c, const ShellPack &p, int &dx, int &dy, int &dz) {
    dx = (int)(c & ((1u << p.sy) - 1u)) - p.k[0];
    dy = (int)((c >> p.sy) & ((1u << (p.sz - p.sy)) - 1u)) - p.k[1];
    dz = (int)(c >> p.sz) - p.k[2];
}

// algo_3d_grid.rs:61-66 (one axis of get_min_dist_sq) and the key of an offset
__host__ __device__ inline double shell_axis(int64_t d, double cs) {
    if (d > 0) return (double)(d - 1) * cs;
    if (d < 0) return (double)(-d - 1) * cs;
    return 0.0;
}

__host__ __device__ inline double shell_key(int dx, int dy, int dz, const double cs[3]) {
    const double mx = shell_axis(dx, cs[0]), my = shell_axis(dy, cs[1]), mz = shell_axis(dz, cs[2]);
    return mx * mx + my * my + mz * mz;
}

// A wall as the kernels see it (src/wall/wall_3d.rs). kind: 0 plane, 1 sphere, 2 cylinder, 3 cone,
// 4 torus, 5 convex polyhedron, 6 tube around a sampled curve (the vd_wall kinds of vd.h), 7 a
// geometry of the caller's (vd_tess_add_wall_planes_per_generator): its contains verdicts and cut
// planes are evaluated on the host per generator; npts is its index among the member's such walls
// and closed its is_planar flag.
struct WallDev {
    int32_t id;
    int32_t kind;
    double a[3];          // plane point / sphere, cylinder, torus centre / cone tip
    double b[3];          // plane unit normal into the valid region / unit axis
    double r;             // sphere, cylinder, tube radius / torus major radius / cone half-angle
    double r2;            // torus minor radius
    double cs, sn, tn;    // cone: cos, sin and tan of the half-angle, evaluated on the host
    const double *pts;    // polyhedron planes (point, outward normal; 6 each) / tube samples (xyz)
    uint32_t npts;
    int32_t closed;       // tube: the last sample joins the first
};

constexpr int kMaxWalls = 8;

// Grid geometry, computed on the host exactly as Algorithm3DGrid::new (algo_3d_grid.rs:47-95).
struct GridGeom {
    double lo[3];
    double s[3];     // bins per unit length
    double cs[3];    // 1 / s
    double lim[3];   // n - 1e-5
    int32_t n[3];
};

// One tessellation of a launch.
struct TessDesc {
    GridGeom grid;
    double box_lo[3], box_hi[3];
    uint64_t raw_off;       // first input point of this member (before the wall filter)
    uint64_t gen_off;       // first generator (after the wall filter) = first output slot
    uint64_t bin_off;       // first bin of this member in the concatenated bin arrays
    uint32_t nraw, ngen;
    const ShellCode *shell;   // shell order prefix of this member's grid geometry
    uint32_t nshell;
    int32_t nwalls;
    double shell_next_d2;   // key of the first entry not in the prefix (+inf if complete)
    ShellPack pack;
    uint8_t pad[8];
    WallDev walls[kMaxWalls];
};

// Everything one cell launch reads and writes. Device pointers.
struct CellArgs {
    const TessDesc *desc;
    const uint32_t *gen_tess;     // member of each generator
    const double *gen;            // generators, xyz AoS, members concatenated
    const double4 *spos;          // generators in bin order: x, y, z, member-local index (u64 bits)
    const uint32_t *bin_start;    // concatenated bins + 1 (positions in spos)
    const double *tube_planes;    // per generator and tube wall: plane point, normal, emitted flag
    uint32_t ntube;               // tube-wall slots per generator (0: no member has one)
    // planes of the caller's walls (kind 7): generator g's planes of its k-th such wall are
    // host_planes[6 q] for host_off[g * nhost + k] <= q < host_off[g * nhost + k + 1]
    const uint64_t *host_off;
    const double *host_planes;
    uint32_t nhost;

    // work list: generator indices (global) to compute
    const uint32_t *cells;
    uint32_t ncells;
    uint32_t *queue;              // atomic work counter (zeroed before launch)
    uint32_t qgrain;              // list entries a wave takes per queue atomic (>= 1)
    // per-cell shell lists (optional, nullptr: none): work-list entry q walks the lshell_n[q]
    // codes from lshell[lshell_off[q]] instead of its member's prefix, unless lshell_n[q] is kNoList
    const ShellCode *lshell;
    const uint64_t *lshell_off;
    const uint32_t *lshell_n;
    const double *lshell_next;    // key after each list's last entry (as TessDesc::shell_next_d2)

    // reduced outputs, indexed by global generator index
    double *vol;
    double *cen;                  // 3 per cell
    uint32_t *nface;
    uint32_t *nvert;              // vertex count (0 = empty cell)
    int32_t *status;              // 0 ok, >0 needs a retry, <0 error

    // full cell export (optional: vtx == nullptr keeps only the neighbour ids in fnb, and with
    // fnb == nullptr too only the reductions are written)
    double *vtx;                  // arena of vertices, 3 doubles each
    uint8_t *fcnt;                // arena of face vertex counts
    int32_t *fnb;                 // arena of face neighbour ids (same indexing as fcnt)
    uint16_t *fidx;               // arena of face vertex indices
    uint64_t *arena_top;          // reserved [0] vertices, [1] faces, [2] face indices; [3..5] used
    uint64_t arena_cap[3];
    uint32_t arena_chunk[3];      // entries a wave reserves at a time (cells are carved from them)
    uint64_t *cell_off;           // 3 per cell: vertex, face, index offsets
    void *gmem;                   // the global-memory tier's polytopes: one slot per wave

    // seal mode (calculate_sealed, tessellation.rs:234-318): cell topology of the previous pass
    const uint32_t *seal_mask;    // per cell: walls (bit = wall index) of its curved-wall set
    const uint32_t *seal_off;     // per cell: its neighbour list in seal_nb (n + 1 offsets)
    const uint32_t *seal_nb;      // member-local neighbour indices in face order
    int32_t *seal_log;            // (cell, neighbour, wall id) per modifying clip, seal_slots
                                  // entries per neighbour from 3 * seal_slots * seal_off[i]
    uint32_t seal_slots;
    uint32_t *seal_logcnt;        // per cell: log entries written
};

// Status codes written per cell.
enum : int32_t {
    kCellOk = 0,
    kCellCapacity = 1,     // polytope exceeded this tier's LDS capacity: retry in a larger tier
    kCellShell = 2,        // walk ran past the shell prefix: retry with a longer prefix
    kCellArena = 3,        // output arena full: grow and retry
    kCellFace255 = -1,     // a face reached 256 vertices (the reference's u8 count would wrap)
};

// ---- per-cell shell lists for cells that reach across empty space (vd_cells.hip) -------------
// launch_far_pick: flags[q] = 1 for work-list entries whose cell stopped at the end of its shell
// prefix (status kCellShell, its r^2 in vol) and whose walk covers about min_entries bins or more;
// limit[q] = the number of prefix entries with key <= L = 4 r^2 (1 + kListMargin), the ones its
// rerun can visit (its radius is r at the stop and does not grow beyond rounding), and next[q] the
// key its list ends at: L when that cuts the prefix, else the prefix's own next key.
// launch_far_lists: block b takes chunk b - cp[s] of kListChunk entries of selected cell s
// (cp[s] <= b < cp[s + 1]; entries below limit[s]) and counts (out == nullptr) or writes, from
// offs[b] (the exclusive scan of the counts), the codes of the entries whose bin lies in the grid
// and holds a generator, in prefix order.
constexpr uint32_t kNoList = 0xffffffffu;
constexpr uint32_t kListChunk = 1024;
constexpr double kListMargin = 1e-9;
hipError_t launch_far_pick(const uint32_t *list, uint32_t n, const int32_t *status, const double *vol,
                           const uint32_t *gen_tess, const TessDesc *desc, double min_entries,
                           uint32_t *flags, uint32_t *limit, double *next, hipStream_t st);
hipError_t launch_far_lists(const uint32_t *cells, const uint32_t *limit, const uint32_t *cp, uint32_t nsel,
                            uint32_t blocks, const TessDesc *desc, const uint32_t *gen_tess, const double *gen,
                            const uint32_t *bin_start, uint32_t *counts, const uint32_t *offs, ShellCode *out,
                            hipStream_t st);

// ---- generator pipeline (vd_grid.hip) ------------------------------------------------------
hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st);
hipError_t launch_member_of(const uint64_t *offs, uint32_t nmem, uint32_t *tess, uint64_t n,
                            hipStream_t st);   // tess[g] = member t with offs[t] <= g < offs[t+1]
hipError_t launch_wall_flags(const double *xyz, uint64_t n, const TessDesc *desc,
                             const uint32_t *tess, const uint32_t *hflags, uint32_t *flags, hipStream_t st);
hipError_t launch_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *tmp,
                           size_t tmp_words, hipStream_t st);
size_t scan_tmp_words(uint32_t n);
hipError_t launch_compact(const double *xyz, const uint32_t *tess_in, const uint32_t *flags,
                          const uint32_t *pos, uint64_t n, double *out, uint32_t *tess_out,
                          hipStream_t st);
hipError_t launch_gather_u32(const uint32_t *src, const uint64_t *idx, uint32_t n, uint32_t *out,
                             hipStream_t st);
hipError_t launch_gather_nb(const int32_t *arena, const uint64_t *cell_off, const uint64_t *offs, uint64_t n,
                            int32_t *out, hipStream_t st);
hipError_t launch_bin_count(const double *xyz, uint64_t n, const TessDesc *desc, const uint32_t *tess,
                            uint32_t *gen_bin, uint32_t *counts, hipStream_t st);
hipError_t launch_bin_scatter(const uint32_t *gen_bin, uint64_t n, const uint32_t *bin_start,
                              uint32_t *fill, uint32_t *items, hipStream_t st);
hipError_t launch_bin_sort(const uint32_t *bin_start, uint64_t nbins, uint32_t *items, hipStream_t st);
hipError_t launch_fill_spos(const double *xyz, const uint32_t *items, uint64_t n, const TessDesc *desc,
                            const uint32_t *tess, double4 *spos, hipStream_t st);
hipError_t launch_tube_planes(const double *gen, uint64_t n, const TessDesc *desc, const uint32_t *tess,
                              uint32_t ntube, double *out, hipStream_t st);   // 7 doubles per slot
hipError_t launch_random_points(uint64_t seed, uint64_t start, uint64_t n, const double lo[3], const double hi[3],
                               double *out, hipStream_t st);   // n points, 24 B each
// Post-pass topology of the cells of the last launch (tessellation.rs:244-264, 330-350): per cell
// the wall mask of its curved-wall set and its count of neighbour faces (0 <= id < member count).
// cnt[i] is 0 for a cell whose mask is empty.
hipError_t launch_cell_topology(const TessDesc *desc, const uint32_t *tess, uint64_t n, const uint32_t *nface,
                                const uint64_t *cell_off, const int32_t *fnb, uint32_t *mask, uint32_t *cnt,
                                hipStream_t st);
// The neighbour lists of the cells with a non-empty mask, at off[i] (exclusive scan of cnt).
hipError_t launch_cell_neighbors(const TessDesc *desc, const uint32_t *tess, uint64_t n, const uint32_t *nface,
                                 const uint64_t *cell_off, const int32_t *fnb, const uint32_t *mask,
                                 const uint32_t *off, uint32_t *nb, hipStream_t st);
// prune_boundaries (tessellation.rs:324-385) per cell: count (ints, doubles) when out_i is null,
// else write the log at (ioff[i], poff[i]).
hipError_t launch_prune(const TessDesc *desc, const uint32_t *tess, uint64_t n, const uint32_t *nface,
                        const uint64_t *cell_off, const int32_t *fnb, const uint8_t *fcnt, const uint16_t *fidx,
                        const double *vtx, const uint32_t *mask, uint32_t *icnt, uint32_t *pcnt,
                        const uint32_t *ioff, const uint32_t *poff, int32_t *out_i, double *out_p, hipStream_t st);
hipError_t launch_relax(const double *gen, const double *cen, const uint32_t *nvert, uint64_t n,
                        double *out, hipStream_t st);
// flags[t] = 1 and rmax[t] = max r2[i] (as bits; r2 >= 0) over the listed cells i of member t
hipError_t launch_mark_members(const uint32_t *list, uint32_t n, const uint32_t *gen_tess, const double *r2,
                               uint32_t *flags, uint64_t *rmax, hipStream_t st);
hipError_t launch_collect(const int32_t *status, const uint32_t *list, uint32_t n,
                          uint32_t *out_cap, uint32_t *out_shell, uint32_t *counts, hipStream_t st);

// ---- shell order prefix (vd_shell.hip) ------------------------------------------------------
// Offsets of the box [-K, K] (w = prod(2K + 1) of them) with key <= T, sorted stably by key.
// shell_count: keys of the box and the number kept (read it from shell_counter after the stream
// syncs); shell_finish: sort and write the `count` kept entries to out (packed with pack),
// next[0] = bits of the first key after them (+inf if none). scratch holds shell_scratch_bytes(w).
size_t shell_scratch_bytes(uint32_t w);
hipError_t shell_count(const ShellPack &pack, const double cs[3], double T, uint32_t w, void *scratch,
                       hipStream_t st);
const uint32_t *shell_counter(const void *scratch, uint32_t w);
hipError_t shell_finish(const ShellPack &pack, const double cs[3], uint32_t w, uint32_t count, void *scratch,
                        size_t scratch_bytes, ShellCode *out, uint64_t *next, hipStream_t st);

// ---- cell kernel (vd_cells.hip) ------------------------------------------------------------
// tier: 0 = register fast path with a small LDS polytope (many waves per CU); 1 = the fast path
// with two vertices per lane (128 vertices); 2, 3 = generic path with LDS polytopes of 256 and 1024
// vertices (the last one wave per CU); 4 = the generic path with the polytope in global memory
// (a.gmem: grid_waves slots of tier_global_bytes() each), up to 65,535 vertices
// t0_waves: the tier-0 instantiation, by resident waves per SIMD (5, 6, 7 or 8; same results)
constexpr int kTier0WavesDefault = 8;
bool tier0_waves_valid(int w);
hipError_t launch_cells(int tier, const CellArgs &a, int grid_waves, hipStream_t st, bool seal = false,
                        bool probe = false, int t0_waves = kTier0WavesDefault);
hipError_t cells_occupancy(int tier, int *blocks_per_cu, int t0_waves = kTier0WavesDefault);
int tier_count();
int tier_max_vertices(int tier);
size_t tier_lds_bytes(int tier);
size_t tier_global_bytes();

}  // namespace vd
