// Device-side generator handling for Algorithm3DGrid: wall admission filter, stable compaction,
// stable counting-sort binning, bin-ordered position copy, and the Lloyd (relax) update. Every
// kernel handles a set of tessellations at once (member lookup through TessDesc / gen_tess).
//
// Reference: src/tessellation.rs:43-67 (set_generators), :175-200 (add_wall / prune),
// src/algorithm/algo_3d_grid.rs:98-130 (get_bin_index, set_generators), src/tessellation.rs:482-495
// (relax). All HBM-bound integer/byte work: one thread per element, coalesced where the access
// pattern allows (generators and flags), 256-thread blocks.
#include "vd_walls.h"

namespace vd {
namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(uint64_t n, int per = kBlock) {
    return (unsigned)((n + per - 1) / per);
}

__device__ inline uint64_t gid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }

// Member of every input point: binary search over the member offsets (nmem + 1 entries).
__global__ void k_member_of(const uint64_t *__restrict__ offs, uint32_t nmem, uint32_t *__restrict__ tess,
                            uint64_t n) {
    const uint64_t g = gid();
    if (g >= n) return;
    uint32_t lo = 0, hi = nmem;   // offs[lo] <= g < offs[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (offs[mid] <= g) lo = mid; else hi = mid;
    }
    tess[g] = lo;
}

// Keeps point g iff every wall contains it (tessellation.rs:43-67); hflags (optional) holds the
// verdict of the caller's walls, evaluated on the host.
__global__ void k_wall_flags(const double *__restrict__ xyz, uint64_t n, const TessDesc *__restrict__ desc,
                             const uint32_t *__restrict__ tess, const uint32_t *__restrict__ hflags,
                             uint32_t *__restrict__ flags) {
    const uint64_t g = gid();
    if (g >= n) return;
    const TessDesc &D = desc[tess[g]];
    const double x = xyz[3 * g], y = xyz[3 * g + 1], z = xyz[3 * g + 2];
    uint32_t in = hflags ? hflags[g] : 1u;
    for (int k = 0; k < D.nwalls; ++k)
        if (!wall_contains(D.walls[k], x, y, z)) { in = 0; break; }
    flags[g] = in;
}

// The cut plane of every tube wall for every generator (get_closest_point + cut,
// wall_3d.rs:408-456, 762-815, 908-961): slot k of generator g holds the plane of its member's
// k-th tube wall as point, normal and a flag (1: a plane is emitted, 0: g lies on the curve).
__global__ void k_tube_planes(const double *__restrict__ gen, uint64_t n, const TessDesc *__restrict__ desc,
                              const uint32_t *__restrict__ tess, uint32_t ntube, double *__restrict__ out) {
    const uint64_t g = gid();
    if (g >= n) return;
    const TessDesc &D = desc[tess[g]];
    const double x = gen[3 * g], y = gen[3 * g + 1], z = gen[3 * g + 2];
    uint32_t k = 0;
    for (int w = 0; w < D.nwalls && k < ntube; ++w) {
        const WallDev &W = D.walls[w];
        if (W.kind != kWallTube) continue;
        double cx, cy, cz, p[3] = {0.0, 0.0, 0.0}, nn[3] = {0.0, 0.0, 0.0};
        tube_closest(W, x, y, z, cx, cy, cz);
        const bool has = tube_plane(W, x, y, z, cx, cy, cz, p, nn);
        double *o = out + 7 * (g * ntube + k);
        o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
        o[3] = nn[0]; o[4] = nn[1]; o[5] = nn[2];
        o[6] = has ? 1.0 : 0.0;
        ++k;
    }
}

// ---- post-passes: calculate_sealed / prune_boundaries (tessellation.rs:234-385) -------------

// The walls whose id is in a cell's curved-wall set: for every face whose neighbour id is not a
// generator of the tessellation, the first wall with that id decides (walls.iter().find); if it is
// curved, the id joins the set (tessellation.rs:248-257), and every wall with that id gets its bit.
__device__ inline uint32_t curved_wall_mask(const TessDesc &D, const int32_t *fl, uint32_t F, uint32_t &nbc) {
    uint32_t m = 0;
    nbc = 0;
    for (uint32_t f = 0; f < F; ++f) {
        const int32_t fn = fl[f];
        if (fn >= 0 && (uint32_t)fn < D.ngen) { ++nbc; continue; }
        int w0 = -1;
        for (int w = 0; w < D.nwalls; ++w)
            if (D.walls[w].id == fn) { w0 = w; break; }
        if (w0 < 0 || wall_planar(D.walls[w0])) continue;
        for (int w = 0; w < D.nwalls; ++w)
            if (D.walls[w].id == fn) m |= 1u << w;
    }
    return m;
}

__global__ void k_cell_topology(const TessDesc *__restrict__ desc, const uint32_t *__restrict__ tess, uint64_t n,
                                const uint32_t *__restrict__ nface, const uint64_t *__restrict__ cell_off,
                                const int32_t *__restrict__ fnb, uint32_t *__restrict__ mask,
                                uint32_t *__restrict__ cnt) {
    const uint64_t g = gid();
    if (g >= n) return;
    const TessDesc &D = desc[tess[g]];
    uint32_t nbc = 0;
    const uint32_t m = curved_wall_mask(D, fnb + cell_off[3 * g + 1], nface[g], nbc);
    mask[g] = m;
    cnt[g] = m ? nbc : 0u;
}

__global__ void k_cell_neighbors(const TessDesc *__restrict__ desc, const uint32_t *__restrict__ tess, uint64_t n,
                                 const uint32_t *__restrict__ nface, const uint64_t *__restrict__ cell_off,
                                 const int32_t *__restrict__ fnb, const uint32_t *__restrict__ mask,
                                 const uint32_t *__restrict__ off, uint32_t *__restrict__ nb) {
    const uint64_t g = gid();
    if (g >= n || !mask[g]) return;
    const TessDesc &D = desc[tess[g]];
    const int32_t *fl = fnb + cell_off[3 * g + 1];
    uint32_t k = off[g];
    for (uint32_t f = 0; f < nface[g]; ++f) {
        const int32_t fn = fl[f];
        if (fn >= 0 && (uint32_t)fn < D.ngen) nb[k++] = (uint32_t)fn;   // face order, as listed
    }
}

// prune_boundaries for one cell: neighbours in first-occurrence face order (deduplicated,
// tessellation.rs:335-339), curved walls in wall order shared with the neighbour, and for each the
// vertices of the cell's wall faces (first occurrence order) that also lie on a face shared with
// the neighbour (Cell3DFaces::shared_vertices, cell_3d_faces.rs:536-568).
__global__ void k_prune(const TessDesc *__restrict__ desc, const uint32_t *__restrict__ tess, uint64_t n,
                        const uint32_t *__restrict__ nface, const uint64_t *__restrict__ cell_off,
                        const int32_t *__restrict__ fnb, const uint8_t *__restrict__ fcnt,
                        const uint16_t *__restrict__ fidx, const double *__restrict__ vtx,
                        const uint32_t *__restrict__ mask, uint32_t *__restrict__ icnt, uint32_t *__restrict__ pcnt,
                        const uint32_t *__restrict__ ioff, const uint32_t *__restrict__ poff,
                        int32_t *__restrict__ out_i, double *__restrict__ out_p) {
    const uint64_t g = gid();
    if (g >= n) return;
    const uint32_t mi = mask[g];
    uint32_t ni = 0, np = 0;
    if (mi) {
        const TessDesc &D = desc[tess[g]];
        const uint32_t li = (uint32_t)(g - D.gen_off);
        const uint32_t F = nface[g];
        const int32_t *fl = fnb + cell_off[3 * g + 1];
        const uint8_t *fc = fcnt + cell_off[3 * g + 1];
        const uint16_t *fi = fidx + cell_off[3 * g + 2];
        const double *vv = vtx + 3 * cell_off[3 * g];
        for (uint32_t f = 0; f < F; ++f) {
            const int32_t j = fl[f];
            if (!(j >= 0 && (uint32_t)j < D.ngen)) continue;
            bool seen = false;
            for (uint32_t e = 0; e < f && !seen; ++e) seen = fl[e] == j;
            if (seen) continue;
            const uint32_t mj = mask[D.gen_off + (uint32_t)j];
            for (int w = 0; w < D.nwalls; ++w) {
                const WallDev &W = D.walls[w];
                if (wall_planar(W) || !((mi >> w) & 1u) || !((mj >> w) & 1u)) continue;
                // shared_vertices(W.id, j)
                uint32_t c = 0;
                uint32_t s = 0;
                for (uint32_t fa = 0; fa < F; s += fc[fa], ++fa) {
                    if (fl[fa] != W.id) continue;
                    for (uint32_t q = 0; q < fc[fa]; ++q) {
                        const uint16_t v = fi[s + q];
                        // first occurrence among the wall faces' slots so far
                        bool dup = false;
                        uint32_t s2 = 0;
                        for (uint32_t fb = 0; fb <= fa && !dup; s2 += fc[fb], ++fb) {
                            if (fl[fb] != W.id) continue;
                            const uint32_t lim = fb == fa ? q : fc[fb];
                            for (uint32_t r = 0; r < lim && !dup; ++r) dup = fi[s2 + r] == v;
                        }
                        if (dup) continue;
                        bool onb = false;
                        uint32_t s3 = 0;
                        for (uint32_t fb = 0; fb < F && !onb; s3 += fc[fb], ++fb) {
                            if (fl[fb] != j) continue;
                            for (uint32_t r = 0; r < fc[fb] && !onb; ++r) onb = fi[s3 + r] == v;
                        }
                        if (!onb) continue;
                        if (out_p) {
                            double *o = out_p + poff[g] + np + 3 * c;
                            o[0] = vv[3 * v]; o[1] = vv[3 * v + 1]; o[2] = vv[3 * v + 2];
                        }
                        ++c;
                    }
                }
                if (c) {
                    if (out_i) {
                        int32_t *o = out_i + ioff[g] + ni;
                        o[0] = (int32_t)li; o[1] = j; o[2] = W.id; o[3] = (int32_t)c;
                    }
                    ni += 4;
                    np += 3 * c;
                }
            }
        }
    }
    if (!out_i) { icnt[g] = ni; pcnt[g] = np; }
}

// Keeps the flagged points in input order (indices are compacted, tessellation.rs:58-60).
__global__ void k_compact(const double *__restrict__ xyz, const uint32_t *__restrict__ tess_in,
                          const uint32_t *__restrict__ flags, const uint32_t *__restrict__ pos, uint64_t n,
                          double *__restrict__ out, uint32_t *__restrict__ tess_out) {
    const uint64_t g = gid();
    if (g >= n || !flags[g]) return;
    const uint64_t d = pos[g];
    out[3 * d] = xyz[3 * g]; out[3 * d + 1] = xyz[3 * g + 1]; out[3 * d + 2] = xyz[3 * g + 2];
    tess_out[d] = tess_in[g];
}

__global__ void k_gather_u32(const uint32_t *__restrict__ src, const uint64_t *__restrict__ idx, uint32_t n,
                             uint32_t *__restrict__ out) {
    const uint64_t k = gid();
    if (k < n) out[k] = src[idx[k]];
}

// The neighbour ids of cells [0, n) of one tessellation packed in cell order: cell k's ids start at
// cell_off[3k + 1] in the face arena and go to out[offs[k], offs[k + 1]).
__global__ void k_gather_nb(const int32_t *__restrict__ arena, const uint64_t *__restrict__ cell_off,
                            const uint64_t *__restrict__ offs, uint64_t n, int32_t *__restrict__ out) {
    const uint64_t k = gid();
    if (k >= n) return;
    const uint64_t src = cell_off[3 * k + 1], o = offs[k], c = offs[k + 1] - o;
    for (uint64_t j = 0; j < c; ++j) out[o + j] = arena[src + j];
}

// get_bin_index per axis (algo_3d_grid.rs:102-104): clamp((x - min) * s, 0, n - 1e-5) as usize.
// f64::clamp passes NaN through and `NaN as usize` is 0.
__device__ inline uint32_t axis_bin(double x, double lo, double s, double lim) {
    double v = (x - lo) * s;
    if (v < 0.0) v = 0.0;
    if (v > lim) v = lim;
    if (!(v == v)) return 0;
    return (uint32_t)v;
}

__global__ void k_bin_count(const double *__restrict__ xyz, uint64_t n, const TessDesc *__restrict__ desc,
                            const uint32_t *__restrict__ tess, uint32_t *__restrict__ gen_bin,
                            uint32_t *__restrict__ counts) {
    const uint64_t g = gid();
    if (g >= n) return;
    const TessDesc &D = desc[tess[g]];
    const GridGeom &G = D.grid;
    const uint32_t ix = axis_bin(xyz[3 * g], G.lo[0], G.s[0], G.lim[0]);
    const uint32_t iy = axis_bin(xyz[3 * g + 1], G.lo[1], G.s[1], G.lim[1]);
    const uint32_t iz = axis_bin(xyz[3 * g + 2], G.lo[2], G.s[2], G.lim[2]);
    const uint32_t b = (uint32_t)D.bin_off + ix + iy * (uint32_t)G.n[0] + iz * (uint32_t)G.n[0] * (uint32_t)G.n[1];
    gen_bin[g] = b;
    atomicAdd(&counts[b], 1u);
}

__global__ void k_bin_scatter(const uint32_t *__restrict__ gen_bin, uint64_t n,
                              const uint32_t *__restrict__ bin_start, uint32_t *__restrict__ fill,
                              uint32_t *__restrict__ items) {
    const uint64_t g = gid();
    if (g >= n) return;
    const uint32_t b = gen_bin[g];
    const uint32_t k = atomicAdd(&fill[b], 1u);
    items[bin_start[b] + k] = (uint32_t)g;
}

// The atomic scatter is unordered; each bin list must be ascending (the reference pushes in
// index order, algo_3d_grid.rs:121-129). Bins hold a handful of items, so a per-bin insertion
// sort by one thread is the cheap way to restore the order.
__global__ void k_bin_sort(const uint32_t *__restrict__ bin_start, uint64_t nbins,
                           uint32_t *__restrict__ items) {
    const uint64_t b = gid();
    if (b >= nbins) return;
    const uint32_t lo = bin_start[b], hi = bin_start[b + 1];
    for (uint32_t k = lo + 1; k < hi; ++k) {
        const uint32_t v = items[k];
        uint32_t j = k;
        while (j > lo && items[j - 1] > v) { items[j] = items[j - 1]; --j; }
        items[j] = v;
    }
}

// Bin-ordered copy of the generators; the 4th slot carries the member-local index.
__global__ void k_fill_spos(const double *__restrict__ xyz, const uint32_t *__restrict__ items, uint64_t n,
                            const TessDesc *__restrict__ desc, const uint32_t *__restrict__ tess,
                            double4 *__restrict__ spos) {
    const uint64_t k = gid();
    if (k >= n) return;
    const uint32_t g = items[k];
    const uint64_t local = g - desc[tess[g]].gen_off;
    double4 r;
    r.x = xyz[3 * (uint64_t)g]; r.y = xyz[3 * (uint64_t)g + 1]; r.z = xyz[3 * (uint64_t)g + 2];
    r.w = __longlong_as_double((long long)local);
    spos[k] = r;
}

// Tessellation::relax (tessellation.rs:482-495): the centroid of each non-empty cell, the old
// position for an empty one.
__global__ void k_relax(const double *__restrict__ gen, const double *__restrict__ cen,
                        const uint32_t *__restrict__ nvert, uint64_t n, double *__restrict__ out) {
    const uint64_t i = gid();
    if (i >= n) return;
    const double *src = nvert[i] == 0 ? gen : cen;
    out[3 * i] = src[3 * i]; out[3 * i + 1] = src[3 * i + 1]; out[3 * i + 2] = src[3 * i + 2];
}

// Candidate points of random_generators (tessellation.rs:96-122): point i takes outputs 3i, 3i + 1
// and 3i + 2 of the SplitMix64 stream of `seed` (Steele, Lea and Flood 2014: output k is the
// finaliser of seed + (k + 1) * 0x9e3779b97f4a7c15), each made a double in [0, 1) from its top 53
// bits and mapped to min + u * (max - min) per axis, the reference's expression.
__device__ inline uint64_t splitmix64_at(uint64_t seed, uint64_t k) {
    uint64_t z = seed + (k + 1) * 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void k_random_points(uint64_t seed, uint64_t start, uint64_t n, double lox, double loy, double loz,
                                double hix, double hiy, double hiz, double *__restrict__ out) {
    const uint64_t i = gid();
    if (i >= n) return;
    const uint64_t k = 3 * (start + i);
    const double lo[3] = {lox, loy, loz}, hi[3] = {hix, hiy, hiz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double u = (double)(splitmix64_at(seed, k + a) >> 11) * 0x1.0p-53;
        out[3 * i + a] = lo[a] + u * (hi[a] - lo[a]);
    }
}

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t v, uint64_t n) {
    const uint64_t i = gid();
    if (i < n) p[i] = v;
}

// Members of the listed cells and the largest radius r^2 among each member's listed cells (for the
// per-geometry shell extension after a launch). Non-negative doubles order as their bit patterns.
__global__ void k_mark_members(const uint32_t *__restrict__ list, uint32_t n, const uint32_t *__restrict__ gen_tess,
                               const double *__restrict__ r2, uint32_t *__restrict__ flags,
                               uint64_t *__restrict__ rmax) {
    const uint64_t g = gid();
    if (g >= n) return;
    const uint32_t i = list[g];
    const uint32_t t = gen_tess[i];
    flags[t] = 1u;
    atomicMax((unsigned long long *)&rmax[t], (unsigned long long)__double_as_longlong(r2[i]));
}

// Sorts the listed cells by outcome of a cell launch: capacity overflow -> out_cap, shell prefix
// too short -> out_shell; counts[0..3] = capacity, shell, arena-full, hard errors.
__global__ void k_collect(const int32_t *__restrict__ status, const uint32_t *__restrict__ list,
                          uint32_t n, uint32_t *__restrict__ out_cap, uint32_t *__restrict__ out_shell,
                          uint32_t *__restrict__ counts) {
    const uint64_t k = gid();
    if (k >= n) return;
    const uint32_t i = list[k];
    const int32_t s = status[i];
    if (s == kCellOk) return;
    if (s == kCellCapacity) out_cap[atomicAdd(&counts[0], 1u)] = i;
    else if (s == kCellShell) out_shell[atomicAdd(&counts[1], 1u)] = i;
    else if (s == kCellArena) atomicAdd(&counts[2], 1u);
    else atomicAdd(&counts[3], 1u);
}

// ---- exclusive scan of u32 (3 phases: tile scan, scan of tile sums, add) -------------------
constexpr int kScanTile = 2048;  // elements per 256-thread block (8 per thread)

__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *sh, uint32_t &total) {
    // Hillis-Steele over 256 threads in LDS.
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = 1; d < kBlock; d <<= 1) {
        uint32_t x = t >= d ? sh[t - d] : 0u;
        __syncthreads();
        sh[t] += x;
        __syncthreads();
    }
    total = sh[kBlock - 1];
    uint32_t incl = sh[t];
    __syncthreads();
    return incl - v;
}

__global__ void k_scan_tiles(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                             uint32_t n, uint32_t *__restrict__ sums) {
    __shared__ uint32_t sh[kBlock];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * 8;
    uint32_t v[8];
    uint32_t local = 0;
    for (int k = 0; k < 8; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        local += v[k];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(local, sh, total);
    for (int k = 0; k < 8; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void k_scan_add(uint32_t *__restrict__ out, uint32_t n, const uint32_t *__restrict__ offs) {
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * 8;
    uint32_t add = offs[blockIdx.x];
    for (int k = 0; k < 8; ++k)
        if (base + k < n) out[base + k] += add;
}

}  // namespace

size_t scan_tmp_words(uint32_t n) {
    // tile sums and their scans at every level
    size_t words = 0;
    uint64_t m = n;
    while (m > 1) {
        m = (m + kScanTile - 1) / kScanTile;
        words += 2 * m + 1;
    }
    return words + 2;
}

// Exclusive scan of n values; out may equal in. tmp holds scan_tmp_words(n) u32.
hipError_t launch_scan_u32(const uint32_t *in, uint32_t *out, uint32_t n, uint32_t *tmp,
                           size_t tmp_words, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (tmp_words < scan_tmp_words(n)) return hipErrorInvalidValue;
    uint32_t tiles = (uint32_t)((n + kScanTile - 1) / kScanTile);
    uint32_t *sums = tmp;
    uint32_t *sums_scan = tmp + tiles;
    k_scan_tiles<<<tiles, kBlock, 0, st>>>(in, out, n, sums);
    if (tiles > 1) {
        hipError_t e = launch_scan_u32(sums, sums_scan, tiles, tmp + 2 * tiles + 1,
                                       tmp_words - (2 * tiles + 1), st);
        if (e != hipSuccess) return e;
        k_scan_add<<<tiles, kBlock, 0, st>>>(out, n, sums_scan);
    }
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_fill_u32<<<blocks_for(n), kBlock, 0, st>>>(p, v, n);
    return hipGetLastError();
}

hipError_t launch_member_of(const uint64_t *offs, uint32_t nmem, uint32_t *tess, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_member_of<<<blocks_for(n), kBlock, 0, st>>>(offs, nmem, tess, n);
    return hipGetLastError();
}

hipError_t launch_wall_flags(const double *xyz, uint64_t n, const TessDesc *desc, const uint32_t *tess,
                             const uint32_t *hflags, uint32_t *flags, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_wall_flags<<<blocks_for(n), kBlock, 0, st>>>(xyz, n, desc, tess, hflags, flags);
    return hipGetLastError();
}

hipError_t launch_compact(const double *xyz, const uint32_t *tess_in, const uint32_t *flags, const uint32_t *pos,
                          uint64_t n, double *out, uint32_t *tess_out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_compact<<<blocks_for(n), kBlock, 0, st>>>(xyz, tess_in, flags, pos, n, out, tess_out);
    return hipGetLastError();
}

hipError_t launch_gather_u32(const uint32_t *src, const uint64_t *idx, uint32_t n, uint32_t *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_gather_u32<<<blocks_for(n), kBlock, 0, st>>>(src, idx, n, out);
    return hipGetLastError();
}

hipError_t launch_gather_nb(const int32_t *arena, const uint64_t *cell_off, const uint64_t *offs, uint64_t n,
                            int32_t *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_gather_nb<<<blocks_for(n), kBlock, 0, st>>>(arena, cell_off, offs, n, out);
    return hipGetLastError();
}

hipError_t launch_bin_count(const double *xyz, uint64_t n, const TessDesc *desc, const uint32_t *tess,
                            uint32_t *gen_bin, uint32_t *counts, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_bin_count<<<blocks_for(n), kBlock, 0, st>>>(xyz, n, desc, tess, gen_bin, counts);
    return hipGetLastError();
}

hipError_t launch_bin_scatter(const uint32_t *gen_bin, uint64_t n, const uint32_t *bin_start, uint32_t *fill,
                              uint32_t *items, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_bin_scatter<<<blocks_for(n), kBlock, 0, st>>>(gen_bin, n, bin_start, fill, items);
    return hipGetLastError();
}

hipError_t launch_bin_sort(const uint32_t *bin_start, uint64_t nbins, uint32_t *items, hipStream_t st) {
    if (nbins == 0) return hipSuccess;
    k_bin_sort<<<blocks_for(nbins), kBlock, 0, st>>>(bin_start, nbins, items);
    return hipGetLastError();
}

hipError_t launch_fill_spos(const double *xyz, const uint32_t *items, uint64_t n, const TessDesc *desc,
                            const uint32_t *tess, double4 *spos, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_fill_spos<<<blocks_for(n), kBlock, 0, st>>>(xyz, items, n, desc, tess, spos);
    return hipGetLastError();
}

hipError_t launch_random_points(uint64_t seed, uint64_t start, uint64_t n, const double lo[3], const double hi[3],
                               double *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_random_points<<<blocks_for(n), kBlock, 0, st>>>(seed, start, n, lo[0], lo[1], lo[2], hi[0], hi[1], hi[2], out);
    return hipGetLastError();
}

hipError_t launch_tube_planes(const double *gen, uint64_t n, const TessDesc *desc, const uint32_t *tess,
                              uint32_t ntube, double *out, hipStream_t st) {
    if (!n || !ntube) return hipSuccess;
    k_tube_planes<<<blocks_for(n), kBlock, 0, st>>>(gen, n, desc, tess, ntube, out);
    return hipGetLastError();
}

hipError_t launch_cell_topology(const TessDesc *desc, const uint32_t *tess, uint64_t n, const uint32_t *nface,
                                const uint64_t *cell_off, const int32_t *fnb, uint32_t *mask, uint32_t *cnt,
                                hipStream_t st) {
    if (!n) return hipSuccess;
    k_cell_topology<<<blocks_for(n), kBlock, 0, st>>>(desc, tess, n, nface, cell_off, fnb, mask, cnt);
    return hipGetLastError();
}

hipError_t launch_cell_neighbors(const TessDesc *desc, const uint32_t *tess, uint64_t n, const uint32_t *nface,
                                 const uint64_t *cell_off, const int32_t *fnb, const uint32_t *mask,
                                 const uint32_t *off, uint32_t *nb, hipStream_t st) {
    if (!n) return hipSuccess;
    k_cell_neighbors<<<blocks_for(n), kBlock, 0, st>>>(desc, tess, n, nface, cell_off, fnb, mask, off, nb);
    return hipGetLastError();
}

hipError_t launch_prune(const TessDesc *desc, const uint32_t *tess, uint64_t n, const uint32_t *nface,
                        const uint64_t *cell_off, const int32_t *fnb, const uint8_t *fcnt, const uint16_t *fidx,
                        const double *vtx, const uint32_t *mask, uint32_t *icnt, uint32_t *pcnt,
                        const uint32_t *ioff, const uint32_t *poff, int32_t *out_i, double *out_p, hipStream_t st) {
    if (!n) return hipSuccess;
    k_prune<<<blocks_for(n), kBlock, 0, st>>>(desc, tess, n, nface, cell_off, fnb, fcnt, fidx, vtx, mask, icnt,
                                              pcnt, ioff, poff, out_i, out_p);
    return hipGetLastError();
}

hipError_t launch_relax(const double *gen, const double *cen, const uint32_t *nvert, uint64_t n, double *out,
                        hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_relax<<<blocks_for(n), kBlock, 0, st>>>(gen, cen, nvert, n, out);
    return hipGetLastError();
}

hipError_t launch_mark_members(const uint32_t *list, uint32_t n, const uint32_t *gen_tess, const double *r2,
                               uint32_t *flags, uint64_t *rmax, hipStream_t st) {
    if (!n) return hipSuccess;
    k_mark_members<<<blocks_for(n), kBlock, 0, st>>>(list, n, gen_tess, r2, flags, rmax);
    return hipGetLastError();
}

hipError_t launch_collect(const int32_t *status, const uint32_t *list, uint32_t n, uint32_t *out_cap,
                          uint32_t *out_shell, uint32_t *counts, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_collect<<<blocks_for(n), kBlock, 0, st>>>(status, list, n, out_cap, out_shell, counts);
    return hipGetLastError();
}

}  // namespace vd
