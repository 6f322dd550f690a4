// Cell kernel for gfx950: Tessellation<3, Cell3DFaces, Algorithm3DGrid>::compute_cell for every
// listed generator, one cell per 64-lane wavefront, persistent over an atomic work queue.
//
// Reference path (mdt-re/vorothree):
//   src/tessellation.rs:414-475          compute_cell (walls, radius, neighbour visitor)
//   src/algorithm/algo_3d_grid.rs:145-202 visit_neighbors (shell walk, early exit, bin test)
//   src/cell/cell_3d_faces.rs:37-74       box cell
//   src/cell/cell_3d_faces.rs:288-503     clip_with_scratch (half-space clip + lid)
//   src/cell/cell_3d_faces.rs:106-205     volume / centroid (fan triangulation, left fold)
//   src/cell/cell_3d_faces.rs:271-286     max_radius_sq
//   src/wall/wall_3d.rs:38-42, 80-100     plane / sphere wall cut
//
// Results are bit-identical to the sequential algorithm: the same clip sequence, the same vertex
// and face numbering, the same expression order (built with -ffp-contract=off), IEEE division and
// square root, and the same left folds.
//
// Layout: the cell's polytope lives in LDS (structure of arrays, ping-pong buffers):
//   vx/vy/vz[V]      vertex coordinates (f64)
//   sedge[S]         per face slot: vertex (low 16 bits) and its successor in the face (high 16)
//   sface[S]         per face slot: face number
//   fstart[F+1]      first slot of each face (fstart[F] = S); fnb[F] neighbour id per face
// "Slots" are the flattened face_indices of the reference; sedge carries the successor so a clip
// reads each edge with one LDS access. In a fast clip (V <= 64 NV), lane v reads vertex v (and
// v + 64, ...) into registers once, at the clip's start.
//
// Parallel clip (one plane): lanes over vertices compute the signed distances and the kept-vertex
// numbering (ballot prefix counts); lanes over slots classify each edge, prefix-sum the output
// entries per face, and list the plane-crossing slots; lanes over crossings find the first
// occurrence of each edge key in slot order (that occurrence's direction fixes the intersection's
// arithmetic, exactly as the reference's insertion-ordered search does) and number the new
// vertices in that order; lanes over the new slots find the lid segments; the lid cycle is walked
// by the whole wave in lock step. Polytopes with V <= 64, S <= 192 and F <= 64 take a fast path
// that keeps the slot data of a clip in registers (three 64-slot chunks) and uses DPP scans;
// larger ones take the generic multi-chunk path. Both produce the same polytope.
//
// Neighbour walk: the shell order is consumed 64 entries at a time (one entry per lane: break
// test, bin range, bin distance). A window of up to 64 candidates is gathered in visiting order
// under the current search radius R^2. Candidates are then clipped in order; a clip whose
// distance pass finds every vertex inside leaves the cell unchanged, exactly like the reference.
// After a cut the radius shrinks, so the rest of the window is re-filtered with the new radius
// (break test and bin test per entry, distance test per candidate) instead of being rebuilt; if a
// cut ever grows the radius (possible by rounding), the window is rebuilt from the position after
// the cut, because entries dropped earlier might then pass.
#include "vd_walls.h"

#include <cfloat>

namespace vd {
namespace {

constexpr double kTol = 1e-9;

__device__ inline int lane_id() { return (int)__lane_id(); }

// The lane index computed again where it is used (v_mbcnt in a volatile block, no memory access):
// LDS addresses derived from it in once-per-cell code then stay there, instead of being hoisted to
// the kernel's start where they would hold registers, or scratch, through the whole walk.
__device__ inline int lane_id_here() {
    int r;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(r));
    return r;
}

// Every cell-kernel workgroup is one wavefront: its lanes execute each LDS instruction together and
// the LDS unit processes one wave's instructions in order, so a value one lane writes is seen by a
// later read of another lane without s_barrier. The wavefront-scope fence emits no instruction; it
// keeps the compiler from moving LDS accesses across the point.
#ifndef VD_WAVE_SYNC_BARRIER
__device__ inline void wave_sync() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront"); }
#else
__device__ inline void wave_sync() { __syncthreads(); }
#endif
// The same ordering point for a tier whose polytope lives in global memory: a workgroup-scope
// fence, which waits for the wave's outstanding vector memory accesses, so a value one lane
// stores is seen by a later load of another lane.
template <class C>
__device__ inline void poly_sync() {
    if constexpr (C::GLOBAL) __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    else wave_sync();
}

// Set bits of m below this lane (v_mbcnt_lo / v_mbcnt_hi): the prefix count of a ballot.
__device__ inline int below_lane(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}
__device__ inline uint64_t ballot(bool p) { return (uint64_t)__ballot(p); }
// Lanes [0, n) as a wave-uniform mask (n <= 0: none, n >= 64: all).
__device__ inline uint64_t lanes_below(int n) {
    return n >= 64 ? ~0ull : n <= 0 ? 0ull : (1ull << n) - 1ull;
}
__device__ inline int popc(uint64_t x) { return __popcll(x); }
__device__ inline int ffs64(uint64_t x) { return x ? (int)__builtin_ctzll(x) : 64; }
__device__ inline int rl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ inline uint32_t rlu(uint32_t v, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}
__device__ inline double rld(double v, int l) {
    long long b = __double_as_longlong(v);
    uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b & 0xffffffffll), l);
    uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((unsigned long long)b >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ inline uint32_t first_lane_u32(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
// Wave-uniform values moved to scalar registers. Loads through the launch's descriptors land in
// vector registers (the compiler cannot prove that the persistent loop's stores leave those
// arrays alone), and the walk would hold them in vector registers across every clip.
__device__ inline double uni_d(double v) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint64_t lo = first_lane_u32((uint32_t)b), hi = first_lane_u32((uint32_t)(b >> 32));
    return __longlong_as_double((long long)((hi << 32) | lo));
}
template <class T>
__device__ inline T *uni_ptr(T *p) {
    const uint64_t b = (uint64_t)p;
    const uint64_t lo = first_lane_u32((uint32_t)b), hi = first_lane_u32((uint32_t)(b >> 32));
    return reinterpret_cast<T *>((hi << 32) | lo);
}
__device__ inline GridGeom uni_grid(const GridGeom &g) {
    GridGeom u;
    for (int a = 0; a < 3; ++a) {
        u.lo[a] = uni_d(g.lo[a]);
        u.s[a] = uni_d(g.s[a]);
        u.cs[a] = uni_d(g.cs[a]);
        u.lim[a] = uni_d(g.lim[a]);
        u.n[a] = (int32_t)first_lane_u32((uint32_t)g.n[a]);
    }
    return u;
}
__device__ inline ShellPack uni_pack(const ShellPack &p) {
    static_assert(sizeof(ShellPack) == 8, "two words");
    uint32_t w[2];
    __builtin_memcpy(w, &p, 8);
    w[0] = first_lane_u32(w[0]);
    w[1] = first_lane_u32(w[1]);
    ShellPack u;
    __builtin_memcpy(&u, w, 8);
    return u;
}

__device__ inline uint64_t bcast_u64(uint64_t v) {  // lane 0's value to every lane
    const uint32_t lo = first_lane_u32((uint32_t)v), hi = first_lane_u32((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
// Value of v in lane src (per-lane source, all lanes must be active).
__device__ inline uint32_t lane_get(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}
__device__ inline double lane_get_d(double v, int src) {
    const long long b = __double_as_longlong(v);
    const uint32_t lo = lane_get((uint32_t)(b & 0xffffffffll), src);
    const uint32_t hi = lane_get((uint32_t)((unsigned long long)b >> 32), src);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

template <int CTRL, int ROW>
__device__ inline uint32_t dpp0(uint32_t x) {  // x from the DPP source lane, 0 where none
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW, 0xf, false);
}

// Inclusive prefix sum over the 64 lanes (row shifts, then row broadcasts 15 and 31).
__device__ inline uint32_t wave_incl_scan(uint32_t x) {
    x += dpp0<0x111, 0xf>(x);
    x += dpp0<0x112, 0xf>(x);
    x += dpp0<0x114, 0xf>(x);
    x += dpp0<0x118, 0xf>(x);
    x += dpp0<0x142, 0xa>(x);
    x += dpp0<0x143, 0xc>(x);
    return x;
}

// Maximum with the reference's update rule `if d2 > m { m = d2 }` (NaN never wins).
__device__ inline double gt_max(double m, double d2) { return d2 > m ? d2 : m; }

__device__ inline uint32_t umax32(uint32_t a, uint32_t b) { return a > b ? a : b; }

// Maximum over the 64 lanes (row shifts, then row broadcasts 15 and 31; 0 fills lanes without a
// DPP source); each step is one v_max_u32 with a DPP operand.
__device__ inline uint32_t wave_max_u32(uint32_t x) {
    x = umax32(x, dpp0<0x111, 0xf>(x));
    x = umax32(x, dpp0<0x112, 0xf>(x));
    x = umax32(x, dpp0<0x114, 0xf>(x));
    x = umax32(x, dpp0<0x118, 0xf>(x));
    x = umax32(x, dpp0<0x142, 0xa>(x));
    x = umax32(x, dpp0<0x143, 0xc>(x));
    return rlu(x, 63);
}

// Maximum over the wave of values built by gt_max from +0.0: never NaN and never -0.0, so their
// IEEE order is the unsigned order of their bit patterns. The high words are reduced first, then
// the low words of the lanes that hold the largest high word; the result has the bits of the
// largest value.
__device__ inline double wave_max(double x) {
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const uint32_t hi = wave_max_u32((uint32_t)(b >> 32));
    const uint32_t lo = wave_max_u32((uint32_t)(b >> 32) == hi ? (uint32_t)b : 0u);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

__device__ inline uint32_t axis_bin(double x, double lo, double s, double lim) {
    double v = (x - lo) * s;  // algo_3d_grid.rs:102-104
    if (v < 0.0) v = 0.0;
    if (v > lim) v = lim;
    if (!(v == v)) return 0;
    return (uint32_t)v;
}

template <int VM_, int SM_, int FM_, int KM_, bool FAST_, int WAVES_, bool GLOBAL_ = false, bool COMPACT_ = false>
struct Cap {
    static constexpr int VM = VM_;      // vertices
    static constexpr int SM = SM_;      // face slots (flattened face indices)
    static constexpr int FM = FM_;      // faces
    static constexpr int KM = KM_;      // plane-crossing slots per clip
    static constexpr bool FAST = FAST_; // register fast path (VM <= 64) instead of the generic one
    static constexpr int WAVES = WAVES_;  // occupancy target (waves per SIMD) for the compiler
    static constexpr bool GLOBAL = GLOBAL_;  // the polytope lives in global memory, not LDS
    // LDS under 5 KiB: the window staging shares the clip scratch's bytes and the volume fold stages
    // half a slot chunk at a time (for 8 resident waves per SIMD)
    static constexpr bool COMPACT = COMPACT_;
    static constexpr int TRI = COMPACT_ ? 32 : 64;   // fold terms staged at a time
    static constexpr int NV = FAST_ ? (VM_ + 63) / 64 : 1;   // register vertices per lane
    // polytope buffers: the fast path rewrites the polytope in place, the generic one ping-pongs
    static constexpr int NB = FAST_ ? 1 : 2;
};

// Scratch of one clip. The fast path's fields live in phases: A (steps 1-4: distances, crossing
// keys, output prefixes), B (steps 3-5: face bookkeeping) and C (steps 5-7: new slot lists, lid
// segments). A and C are never live together, so they share bytes. The generic path keeps one
// field per name.
template <class C, bool FAST = C::FAST>
struct ClipScratch;

template <class C>
struct ClipScratch<C, true> {
    union {
        struct {
            double dist[C::VM];
            uint32_t ckey[C::KM];
            uint32_t cab[C::KM];
            uint16_t P[C::SM + 1];
        };
        struct {
            uint16_t fidxn[C::SM];
            uint16_t nextp[C::SM];
            uint32_t seg[C::KM];    // lid segments: one per cut face
            uint32_t lmap[C::KM];   // by new vertex minus the kept count (at most K new vertices)
        };
    };
    int32_t fbase[C::FM];
    int16_t fnew[C::FM];
    uint16_t fnoff[C::FM];
    uint16_t nend[C::FM];
};

template <class C>
struct ClipScratch<C, false> {
    double dist[C::VM];
    uint32_t ckey[C::KM];
    uint32_t cab[C::KM];
    uint32_t lmap[C::VM];
    uint32_t seg[C::SM];
    int32_t fbase[C::FM];
    uint16_t remap[C::VM];
    uint16_t P[C::SM + 1];
    uint16_t cxid[C::KM];
    uint16_t fidxn[C::SM];
    uint16_t nextp[C::SM];
    int16_t fnew[C::FM];
    uint16_t fnoff[C::FM];
    uint16_t nend[C::FM];
    uint16_t lid[C::SM];
};

// One gather's window staging: written, then read back into registers.
struct WinStage {
    uint32_t q[64];     // bin slot of window lane w's candidate
    uint16_t ent[64];   // batch lane of its shell entry
    uint16_t off[64];   // its offset within the entry's bin
};

// Bytes of phases that never overlap: one clip's scratch and, at the end of a cell, the fold terms
// of one 64-slot chunk (half a chunk in the compact layout), and in the compact layout also one
// gather's window staging (read back into registers before the next clip). The default layout
// keeps the window staging apart: sharing it saves 512 bytes but measured slower on walks through
// sparse regions, where gathers dominate.
template <class C, bool COMPACT = C::COMPACT>
struct MemScratch {
    WinStage win;
    union {
        ClipScratch<C> sc;
        double tri[4][C::TRI + 1];   // +1 keeps the four rows in different LDS banks
    };
};
template <class C>
struct MemScratch<C, true> {
    union {
        WinStage win;
        ClipScratch<C> sc;
        double tri[4][C::TRI + 1];
    };
};

// Tier 0's LDS per wave: 5.5 KiB in the default layout (up to 7 resident waves per SIMD), under
// 5 KiB in the compact one (8).
template <class C>
struct Mem : MemScratch<C> {
    double vx[C::NB][C::VM], vy[C::NB][C::VM], vz[C::NB][C::VM];
    uint32_t sedge[C::NB][C::SM];
    int32_t fnb[C::NB][C::FM];
    uint16_t sface[C::NB][C::SM];
    uint16_t fstart[C::NB][C::FM + 1];
};

enum ClipRc : int { kNoChange = 0, kModified = 1, kOverflow = -1, kFace255 = -2, kGeneric = -3 };

struct Plane { double px, py, pz, nx, ny, nz; };

// The polytope state of one wave; all members are wave-uniform. After a clip that changed the
// polytope, vertices [0, nk) are bit-exact copies of the kept vertices and [nk, V) are new.
struct Poly { int cur, V, F, S, nk; };

// The buffer holding the current polytope: always 0 in the single-buffer (fast) tiers, which lets
// the compiler fold every LDS address of a 64-lane chunk into one lane base and immediate offsets.
template <class C>
__device__ inline int cur_buf(const Poly &P) { return C::NB == 1 ? 0 : P.cur; }

// A wave's reserved arena ranges [at, end) for vertices, faces and face indices. Cells are carved
// from them, so the shared arena counters see one atomic per chunk instead of three per cell
// (with thousands of waves, same-address atomics per cell cost more than the cell's stores).
struct ArenaChunks { uint64_t at[3] = {0, 0, 0}, end[3] = {0, 0, 0}, used[3] = {0, 0, 0}; };

// Offset of n entries of arena k from the wave's chunk, reserving a new chunk (or a range of its
// own for a cell larger than a chunk) when the current one cannot hold them. Wave-uniform.
__device__ inline uint64_t arena_take(const CellArgs &A, ArenaChunks &ac, int k, uint64_t n) {
    if (ac.at[k] + n > ac.end[k]) {
        const uint64_t sz = n > A.arena_chunk[k] ? n : (uint64_t)A.arena_chunk[k];
        uint64_t base = 0;
        if (lane_id() == 0) base = atomicAdd((unsigned long long *)&A.arena_top[k], (unsigned long long)sz);
        base = bcast_u64(base);
        ac.at[k] = base;
        ac.end[k] = base + sz;
    }
    const uint64_t o = ac.at[k];
    ac.at[k] += n;
    ac.used[k] += n;
    return o;
}

// Lane l's copies of vertices l, 64 + l, ... during a fast clip (NV per lane).
template <int NV>
struct VRegN { double x[NV], y[NV], z[NV]; };

template <class C>
using VReg = VRegN<C::NV>;

// A vertex set as NV 64-bit ballots (vertex v in word v / 64, bit v % 64).
template <int NV>
struct VMask {
    uint64_t w[NV];
    __device__ bool test(uint32_t v) const {
        if constexpr (NV == 1) return (w[0] >> v) & 1ull;
        else return v < 64 ? ((w[0] >> v) & 1ull) : ((w[1] >> (v - 64)) & 1ull);
    }
    __device__ int below(uint32_t v) const {   // members with an index below v
        if constexpr (NV == 1) return popc(w[0] & ((1ull << v) - 1ull));
        else return v < 64 ? popc(w[0] & ((1ull << v) - 1ull)) : popc(w[0]) + popc(w[1] & ((1ull << (v - 64)) - 1ull));
    }
    __device__ int count() const {
        int c = 0;
#pragma unroll
        for (int k = 0; k < NV; ++k) c += popc(w[k]);
        return c;
    }
};

template <class C>
__device__ void poly_box(Mem<C> &m, Poly &P, const double lo[3], const double hi[3]) {
    // cell_3d_faces.rs:37-74: vertex order (lo,lo,lo) (hi,lo,lo) (hi,hi,lo) (lo,hi,lo), then z=hi;
    // faces z-, z+, y-, y+, x-, x+ with the listed index cycles; neighbours -5 -6 -3 -4 -1 -2.
    const int l = lane_id_here();
    if (l < 8) {
        const int xb = (l == 1 || l == 2 || l == 5 || l == 6);
        const int yb = (l == 2 || l == 3 || l == 6 || l == 7);
        const int zb = l >> 2;
        m.vx[0][l] = xb ? hi[0] : lo[0];
        m.vy[0][l] = yb ? hi[1] : lo[1];
        m.vz[0][l] = zb ? hi[2] : lo[2];
    }
    if (l < 24) {
        // face f = l/4, position i = l%4, successor (i+1)%4; one nibble per index:
        // z- 3,2,1,0 | z+ 4,5,6,7 | y- 0,1,5,4 | y+ 2,3,7,6 | x- 0,4,7,3 | x+ 1,2,6,5
        const uint64_t lo16 = 0x6732'4510'7654'0123ull;  // faces 0-3, nibble k = index k
        const uint32_t hi8 = 0x5621'3740u;               // faces 4-5
        const int f = l >> 2, i = l & 3;
        const int ka = 4 * f + i, kb = 4 * f + ((i + 1) & 3);
        const uint32_t a = (uint32_t)((ka < 16 ? lo16 >> (4 * ka) : (uint64_t)hi8 >> (4 * (ka - 16))) & 0xf);
        const uint32_t b = (uint32_t)((kb < 16 ? lo16 >> (4 * kb) : (uint64_t)hi8 >> (4 * (kb - 16))) & 0xf);
        m.sedge[0][l] = a | (b << 16);
        m.sface[0][l] = (uint16_t)f;
    }
    if (l < 7) m.fstart[0][l] = (uint16_t)(4 * l);
    if (l < 6) m.fnb[0][l] = -(int32_t)((0x214365u >> (4 * l)) & 0xf);  // -5 -6 -3 -4 -1 -2
    P.cur = 0; P.V = 8; P.F = 6; P.S = 24; P.nk = 0;
    poly_sync<C>();
}

// Half-space clip keeping (x - p).n <= 1e-9 (cell_3d_faces.rs:288-503). Returns kNoChange,
// kModified (r2 = max |v - g|^2 over the new vertices, or 0 without a generator), or a negative
// code when this tier's capacity is exceeded.
template <class C>
__device__ int clip_generic(Mem<C> &m, Poly &P, const Plane pl, int32_t nid, bool withgen,
                            double gx, double gy, double gz, double &r2) {
    const int l = lane_id();
    const int c = P.cur, o = c ^ 1;

    // 1. signed distances, early outs, kept-vertex numbering (:303-362)
    uint64_t gtm = 0, ltm = 0;
    int nk = 0;
    double md2 = 0.0;
    for (int base = 0; base < P.V; base += 64) {
        const int v = base + l;
        const bool act = v < P.V;
        double x = 0.0, y = 0.0, z = 0.0, d = 0.0;
        if (act) {
            x = m.vx[c][v]; y = m.vy[c][v]; z = m.vz[c][v];
            d = (x - pl.px) * pl.nx + (y - pl.py) * pl.ny + (z - pl.pz) * pl.nz;
            m.sc.dist[v] = d;
        }
        gtm |= ballot(act && d > kTol);
        ltm |= ballot(act && d < -kTol);
        const bool in = act && d <= kTol;
        const uint64_t im = ballot(in);
        if (in) {
            const int ni = nk + below_lane(im);
            m.sc.remap[v] = (uint16_t)ni;
            m.vx[o][ni] = x; m.vy[o][ni] = y; m.vz[o][ni] = z;
            if (withgen) {
                const double dx = x - gx, dy = y - gy, dz = z - gz;
                md2 = gt_max(md2, dx * dx + dy * dy + dz * dz);
            }
        }
        nk += popc(im);
    }
    if (gtm == 0) return kNoChange;     // all inside (:317-319)
    if (ltm == 0) {                     // all outside (:320-326)
        P.V = P.F = P.S = P.nk = 0;
        r2 = 0.0;
        return kModified;
    }
    poly_sync<C>();

    // 2. per slot: output entry count and plane crossings (:378-452)
    int K = 0, ptot = 0;
    for (int base = 0; base < P.S; base += 64) {
        const int s = base + l;
        const bool act = s < P.S;
        const uint32_t e = act ? m.sedge[c][s] : 0u;
        const uint32_t a = e & 0xffffu, b = e >> 16;
        const double da = m.sc.dist[a], db = m.sc.dist[b];
        const bool ain = da <= kTol, bin = db <= kTol;
        const uint64_t b0 = ballot(act && ain);            // in->in, in->out: 1 entry
        const uint64_t b1 = ballot(act && !ain && bin);    // out->in: 2 entries
        if (act) m.sc.P[s] = (uint16_t)(ptot + below_lane(b0) + 2 * below_lane(b1));
        ptot += popc(b0) + 2 * popc(b1);
        const bool cr = act && (ain != bin);
        const uint64_t cm = ballot(cr);
        if (cr) {
            const int ci = K + below_lane(cm);
            if (ci < C::KM) {
                m.sc.ckey[ci] = a < b ? ((a << 16) | b) : ((b << 16) | a);
                m.sc.cab[ci] = a | (b << 16);
            }
        }
        K += popc(cm);
    }
    if (l == 0) m.sc.P[P.S] = (uint16_t)ptot;
    if (K > C::KM || ptot > 0xffff) return kOverflow;
    poly_sync<C>();

    // 3. per face: surviving faces (>= 3 entries, :454-456) and their new offsets
    int Fn = 0, Spre = 0;
    bool big = false;
    for (int base = 0; base < P.F; base += 64) {
        const int f = base + l;
        const bool act = f < P.F;
        const int fs = act ? m.fstart[c][f] : 0, fe = act ? m.fstart[c][f + 1] : 0;
        const int tot = act ? (int)m.sc.P[fe] - (int)m.sc.P[fs] : 0;
        const bool sv = act && tot >= 3;
        const uint64_t smk = ballot(sv);
        const int nf = Fn + below_lane(smk);
        const uint32_t t = sv ? (uint32_t)tot : 0u;
        const uint32_t inc = wave_incl_scan(t);
        const int off = Spre + (int)(inc - t);
        if (act) { m.sc.fnew[f] = (int16_t)(sv ? nf : -1); m.sc.fnoff[f] = (uint16_t)off; }
        if (sv && nf < C::FM) { m.fstart[o][nf] = (uint16_t)off; m.fnb[o][nf] = m.fnb[c][f]; }
        big |= ballot(sv && tot > 255) != 0;
        Fn += popc(smk);
        Spre += (int)rlu(inc, 63);
    }
    if (big) return kFace255;
    if (Fn + 1 > C::FM || Spre > C::SM) return kOverflow;

    // 4. first occurrence of each crossing edge in slot order numbers the new vertices (:391-417)
    int nint = 0;
    for (int base = 0; base < K; base += 64) {
        const int ci = base + l;
        const bool act = ci < K;
        const uint32_t key = act ? m.sc.ckey[ci] : 0xffffffffu;
        int first = ci;
        if (K <= 64) {
            for (int cj = 0; cj < K; ++cj) {
                const uint32_t kj = rlu(key, cj);
                if (cj < first && kj == key) first = cj;
            }
        } else {
            const int lim = min(K, base + 64);
            for (int cj = 0; cj < lim; ++cj) {
                const uint32_t kj = m.sc.ckey[cj];
                if (cj < first && kj == key) first = cj;
            }
        }
        const bool isf = act && first == ci;
        const uint64_t fm = ballot(isf);
        const int xid = nk + nint + below_lane(fm);
        nint += popc(fm);
        if (act) m.sc.cxid[ci] = (uint16_t)(isf ? xid : (0x8000 | first));
        if (isf && xid < C::VM) {
            const uint32_t ab = m.sc.cab[ci];
            const uint32_t a = ab & 0xffffu, b = ab >> 16;
            const double da = m.sc.dist[a], db = m.sc.dist[b];
            double t = da / (da - db);
            if (t < 0.0) t = 0.0;   // f64::clamp(0, 1)
            if (t > 1.0) t = 1.0;
            const double ax = m.vx[c][a], ay = m.vy[c][a], az = m.vz[c][a];
            const double bx = m.vx[c][b], by = m.vy[c][b], bz = m.vz[c][b];
            const double x = ax + t * (bx - ax);
            const double y = ay + t * (by - ay);
            const double z = az + t * (bz - az);
            m.vx[o][xid] = x; m.vy[o][xid] = y; m.vz[o][xid] = z;
            if (withgen) {
                const double dx = x - gx, dy = y - gy, dz = z - gz;
                md2 = gt_max(md2, dx * dx + dy * dy + dz * dz);
            }
        }
    }
    if (nk + nint > C::VM) return kOverflow;
    poly_sync<C>();
    for (int ci = l; ci < K; ci += 64) {
        const uint16_t v = m.sc.cxid[ci];
        if (v & 0x8000) m.sc.cxid[ci] = m.sc.cxid[v & 0x7fff];
    }
    poly_sync<C>();

    // 5. emit the surviving faces' entries (:386-451, :466)
    int Kb = 0;
    for (int base = 0; base < P.S; base += 64) {
        const int s = base + l;
        const bool act = s < P.S;
        const uint32_t e = act ? m.sedge[c][s] : 0u;
        const uint32_t a = e & 0xffffu, b = e >> 16;
        const double da = m.sc.dist[a], db = m.sc.dist[b];
        const bool ain = da <= kTol, bin = db <= kTol;
        const bool cr = act && (ain != bin);
        const uint64_t cm = ballot(cr);
        const int ci = Kb + below_lane(cm);
        Kb += popc(cm);
        const int f = act ? m.sface[c][s] : 0;
        const int nf = act ? m.sc.fnew[f] : -1;
        if (act && nf >= 0 && (ain || bin)) {
            const int pos = (int)m.sc.fnoff[f] + ((int)m.sc.P[s] - (int)m.sc.P[m.fstart[c][f]]);
            if (ain && bin) {
                m.sc.fidxn[pos] = m.sc.remap[b];
                m.sface[o][pos] = (uint16_t)nf;
            } else {
                m.sc.fidxn[pos] = m.sc.cxid[ci];
                m.sface[o][pos] = (uint16_t)nf;
                if (!ain) {
                    m.sc.fidxn[pos + 1] = m.sc.remap[b];
                    m.sface[o][pos + 1] = (uint16_t)nf;
                }
            }
        }
    }
    if (l == 0) m.fstart[o][Fn] = (uint16_t)Spre;
    poly_sync<C>();

    // 6. successor of every new slot; lid segments (v, u) for intersection pairs (:459-465)
    int M = 0;
    for (int base = 0; base < Spre; base += 64) {
        const int p = base + l;
        const bool act = p < Spre;
        const int fp = act ? m.sface[o][p] : 0;
        const int fs = m.fstart[o][fp], fe = m.fstart[o][fp + 1];
        const int nx = (p + 1 == fe) ? fs : p + 1;
        const uint32_t ua = act ? m.sc.fidxn[p] : 0u, vb = act ? m.sc.fidxn[nx] : 0u;
        if (act) m.sedge[o][p] = ua | (vb << 16);
        const bool sg = act && (int)ua >= nk && (int)vb >= nk;
        const uint64_t gm = ballot(sg);
        if (sg) m.sc.seg[M + below_lane(gm)] = vb | (ua << 16);
        M += popc(gm);
    }

    // 7. lid polygon (:471-495): map[first] = second (later segments overwrite), walk from seg[0]
    if (M > 0) {
        for (int t = l; t < nint; t += 64) m.sc.lmap[t] = 0u;
        poly_sync<C>();
        for (int t = l; t < M; t += 64) {
            const uint32_t sg = m.sc.seg[t];
            atomicMax(&m.sc.lmap[(int)(sg & 0xffffu) - nk], ((uint32_t)(t + 1) << 16) | (sg >> 16));
        }
        poly_sync<C>();
        const uint32_t s0 = first_lane_u32(m.sc.seg[0]);   // wave-uniform: the walk is a scalar loop
        const int start = (int)(s0 & 0xffffu);
        int cv = (int)(s0 >> 16);
        if (l == 0) m.sc.lid[0] = (uint16_t)start;
        int len = 1;
        while (cv != start && len <= M) {
            if (len >= C::SM) return kOverflow;
            if (l == 0) m.sc.lid[len] = (uint16_t)cv;
            ++len;
            const uint32_t mv = m.sc.lmap[cv - nk];
            cv = mv ? (int)(mv & 0xffffu) : 0xffff;
            if (cv == 0xffff) break;
        }
        if (len >= 3) {
            if (Spre + len > C::SM || len > 255) return len > 255 ? kFace255 : kOverflow;
            poly_sync<C>();
            for (int t = l; t < len; t += 64) {
                const uint32_t u = m.sc.lid[t], w = m.sc.lid[(t + 1) % len];
                m.sedge[o][Spre + t] = u | (w << 16);
                m.sface[o][Spre + t] = (uint16_t)Fn;
            }
            if (l == 0) {
                m.fstart[o][Fn] = (uint16_t)Spre;
                m.fnb[o][Fn] = nid;
                m.fstart[o][Fn + 1] = (uint16_t)(Spre + len);
            }
            Fn += 1;
            Spre += len;
        }
    }
    poly_sync<C>();
    P.cur = o;
    P.V = nk + nint;
    P.F = Fn;
    P.S = Spre;
    P.nk = nk;
    r2 = withgen ? wave_max(md2) : 0.0;
    return kModified;
}


// Fast clip for V <= 64 * NV, S <= 192 * NV, F <= 64: vertex distances from register copies read
// at its start (NV vertices per lane), slot data of the clip held in registers (3 * NV chunks of
// 64 slots), crossings and lid cycle handled with lane shuffles. Same result as clip_generic. Returns kGeneric
// (with the current polytope untouched) when the faces, crossings or the lid do not fit its
// register layout.
template <class C>
__device__ int clip_fast(Mem<C> &m, Poly &P, const Plane pl, int32_t nid, bool withgen, double gx, double gy,
                         double gz, double &r2) {
    constexpr int NV = C::NV;
    constexpr int NCH = 3 * NV;                 // slot chunks
    constexpr uint32_t SH = NV == 1 ? 6 : 7;    // bits of a vertex index in the packed slot word
    constexpr uint32_t VMSK = (1u << SH) - 1u;
    const int l = lane_id_here();
    // One buffer: every read of an old entry comes before the write that can replace it (the slot
    // and face words are in registers after step 2; old vertices are last read in step 4, so the
    // kept ones are stored after it). A clip that fails part way leaves a mixed polytope, but every
    // failure moves the cell to the next tier, which starts it again.
    static_assert(C::NB == 1, "the fast path rewrites the polytope in place");
    const int c = 0, o = 0;
    if (NV > 1 && P.F > 64) return kGeneric;

    // 1. signed distances (:303-315), early outs (:317-326), kept vertices (:345-362). Each lane
    //    reads its vertices from LDS in whole 64-lane chunks of the VM-entry arrays (read here
    //    rather than kept in registers across the walk, which frees registers there). Lanes past
    //    the polytope compute a distance from stale entries; every ballot is masked with the
    //    polytope's lanes (a scalar mask), so those lanes take no part. Ballots are of single
    //    compares, combined with scalar operations (a ballot of a combined predicate costs two more
    //    vector instructions).
    static_assert(64 * NV <= C::VM, "whole vertex chunks stay inside the vertex arrays");
    VReg<C> r;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        r.x[k] = m.vx[0][64 * k + l];
        r.y[k] = m.vy[0][64 * k + l];
        r.z[k] = m.vz[0][64 * k + l];
    }
    double d[NV];
    VMask<NV> gtm, ltm, inm, vam;
    bool anyg = false, anyl = false;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        vam.w[k] = lanes_below(P.V - 64 * k);
        d[k] = (r.x[k] - pl.px) * pl.nx + (r.y[k] - pl.py) * pl.ny + (r.z[k] - pl.pz) * pl.nz;
        gtm.w[k] = ballot(d[k] > kTol) & vam.w[k];
        anyg |= gtm.w[k] != 0;
    }
    if (!anyg) return kNoChange;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        ltm.w[k] = ballot(d[k] < -kTol) & vam.w[k];
        anyl |= ltm.w[k] != 0;
    }
    if (!anyl) {
        P.V = P.F = P.S = P.nk = 0;
        r2 = 0.0;
        return kModified;
    }
    // kin: this lane's vertex is kept (the lane predicate, so tests of it need no bit extraction)
    bool kin[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        kin[k] = d[k] <= kTol && 64 * k + l < P.V;
        inm.w[k] = ballot(kin[k]);
    }
    const int nk = inm.count();
    // one vertex per lane: lane v's new index if it is kept (read by lane shuffle in step 5)
    const uint32_t kidx = NV == 1 ? (uint32_t)below_lane(inm.w[0]) : 0u;
    double md2 = 0.0;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        m.sc.dist[64 * k + l] = d[k];   // never read past the polytope
        if (withgen && kin[k]) {
            const double dx = r.x[k] - gx, dy = r.y[k] - gy, dz = r.z[k] - gz;
            md2 = gt_max(md2, dx * dx + dy * dy + dz * dz);
        }
    }

    // 2. slots: output entry counts (prefix in slot order) and plane crossings (:378-452).
    //    Per chunk and lane two registers: ea = a | b << SH | face << 2 SH | a kept << 30 | b kept << 31 and
    //    eq = output prefix | crossing index << 16.
    //    Slot, key and step-6 reads below index whole 64-lane chunks of arrays of SM (or KM = 64)
    //    entries without a guard, then select: a guarded read is a branch whose predicate the
    //    compiler then carries through a vector register.
    static_assert(64 * NCH <= C::SM && C::SM % 64 == 0 && C::KM >= 64, "chunked reads stay inside the slot arrays");
    const int nch = (P.S + 63) >> 6;
    uint32_t ea[NCH];
    uint32_t eq[NCH];
    int K = 0, ptot = 0;
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        ea[ch] = 0u;
        eq[ch] = 0u;
        if (ch < nch) {
            const int s = 64 * ch + l;
            const bool sa = s < P.S;
            const uint32_t e0 = m.sedge[c][s], f0 = m.sface[c][s];
            const uint32_t e = sa ? e0 : 0u;
            const uint32_t a = e & 0xffffu, b = e >> 16;
            const bool ain = inm.test(a), bin = inm.test(b);
            // (the face field of a lane past the slots is never read)
            ea[ch] = a | (b << SH) | (f0 << (2 * SH)) | ((uint32_t)ain << 30) | ((uint32_t)bin << 31);
            const uint64_t sam = ballot(sa), ainm = ballot(ain), binm = ballot(bin);
            const uint64_t b0 = sam & ainm;
            const uint64_t b1 = sam & ~ainm & binm;
            const uint32_t pp = (uint32_t)(ptot + below_lane(b0) + 2 * below_lane(b1));
            ptot += popc(b0) + 2 * popc(b1);
            m.sc.P[s] = (uint16_t)pp;   // past P.S never read; P[P.S] is written below
            const bool cr = sa && (ain != bin);
            const uint64_t cm = sam & (ainm ^ binm);
            const int ci = K + below_lane(cm);
            eq[ch] = pp | ((uint32_t)(ci & 0xffff) << 16);
            if (cr && ci < 64) {
                m.sc.ckey[ci] = a < b ? ((a << 16) | b) : ((b << 16) | a);
                m.sc.cab[ci] = a | (b << 16);
            }
            K += popc(cm);
        }
    }
    if (K > 64) return kGeneric;
    if (l == 0) m.sc.P[P.S] = (uint16_t)ptot;
    poly_sync<C>();

    // 3. faces: survivors (>= 3 entries, :454-456) and their new offsets
    const bool fa = l < P.F;
    const int fs = fa ? m.fstart[c][l] : 0, fe = fa ? m.fstart[c][l + 1] : 0;
    const int pfs = fa ? (int)m.sc.P[fs] : 0;
    const int tot = fa ? (int)m.sc.P[fe] - pfs : 0;
    const bool sv = fa && tot >= 3;
    const uint64_t svm = ballot(tot >= 3);   // tot = 0 past the faces
    int Fn = popc(svm);
    // a face's output offset is the slot prefix of its first slot unless a face before it drops
    // 1 or 2 entries (a face left with fewer than 3): a scan only then
    int off = pfs, Spre = ptot;
    if (ballot(tot > 0) & ~svm) {
        const uint32_t tv = sv ? (uint32_t)tot : 0u;
        const uint32_t inc = wave_incl_scan(tv);
        off = (int)(inc - tv);
        Spre = (int)rlu(inc, 63);
    }
    if (ballot(tot > 255)) return kFace255;
    if (Fn + 1 > C::FM || Spre > C::SM) return kOverflow;
    if (fa) {
        const int nf = below_lane(svm);
        m.sc.fnew[l] = (int16_t)(sv ? nf : -1);
        m.sc.fbase[l] = off - pfs;
        m.sc.fnoff[l] = (uint16_t)off;
        m.sc.nend[l] = (uint16_t)(off + tot);
        if (sv) {
            m.fstart[o][nf] = (uint16_t)off;
            m.fnb[o][nf] = m.fnb[c][l];
        }
    }
    if (l == 0) m.fstart[o][Fn] = (uint16_t)Spre;

    // 4. crossings: the first occurrence of an edge key in slot order creates the vertex
    //    (:391-417); new vertices are numbered in that order after the kept ones. Each pass takes
    //    the lowest crossing not numbered yet, which is the first occurrence of its key, and gives
    //    every crossing with that key the next new index: one pass per new vertex. Keys are below
    //    0xffff0000 (vertex indices are 16-bit and a < b), so lanes without a crossing never match.
    const bool ka = l < K;
    const uint32_t key0 = m.sc.ckey[l];
    const uint32_t key = ka ? key0 : 0xffffffffu;
    uint64_t rest = ballot(ka), fm = 0;
    int nint = 0, xid = 0;
    while (rest) {     // every pass clears lane f (its own key matches): at most K <= 64 passes
        const int f = ffs64(rest);
        const bool mine = key == rlu(key, f);
        const uint64_t same = ballot(mine);
        rest &= ~same;
        fm |= 1ull << f;
        if (mine) xid = nint;
        ++nint;
    }
    xid += nk;   // read from crossing lanes only
    const bool isf = (fm >> l) & 1ull;
    const int myx = xid;
    if (nk + nint > C::VM) return kOverflow;
    double x = 0.0, y = 0.0, z = 0.0;
    if (isf) {
        const uint32_t ab = m.sc.cab[l];
        const uint32_t a = ab & 0xffffu, b = ab >> 16;
        const double da = m.sc.dist[a], db = m.sc.dist[b];
        double t = da / (da - db);
        if (t < 0.0) t = 0.0;   // f64::clamp(0, 1)
        if (t > 1.0) t = 1.0;
        const double ax = m.vx[c][a], ay = m.vy[c][a], az = m.vz[c][a];
        const double bx = m.vx[c][b], by = m.vy[c][b], bz = m.vz[c][b];
        x = ax + t * (bx - ax);
        y = ay + t * (by - ay);
        z = az + t * (bz - az);
        if (withgen) {
            const double dx = x - gx, dy = y - gy, dz = z - gz;
            md2 = gt_max(md2, dx * dx + dy * dy + dz * dz);
        }
    }
    poly_sync<C>();
    // the old vertices are read: store the kept ones at [0, nk) and the new ones after them
    int kbase = 0;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        if (kin[k]) {
            const int ni = kbase + below_lane(inm.w[k]);
            m.vx[o][ni] = r.x[k]; m.vy[o][ni] = r.y[k]; m.vz[o][ni] = r.z[k];
        }
        kbase += popc(inm.w[k]);
    }
    if (isf) { m.vx[o][myx] = x; m.vy[o][myx] = y; m.vz[o][myx] = z; }
    poly_sync<C>();

    // 5. emit the entries of every surviving face (:386-451), with each entry's successor slot
#pragma unroll
    for (int ch = 0; ch < NCH; ++ch) {
        if (ch < nch) {
            const int s = 64 * ch + l;
            const uint32_t b = (ea[ch] >> SH) & VMSK;
            const bool ain = (ea[ch] >> 30) & 1u, bin = ea[ch] >> 31;
            const int xv = (int)lane_get((uint32_t)xid, (int)(eq[ch] >> 16) & 63);
            const uint32_t kb = NV == 1 ? lane_get(kidx, (int)b) : 0u;   // full wave: b's new index
            if (s < P.S && (ain || bin)) {
                const int f = (int)((ea[ch] >> (2 * SH)) & ((1u << (30 - 2 * SH)) - 1u));
                const int nf = m.sc.fnew[f];
                if (nf >= 0) {
                    const int pos = m.sc.fbase[f] + (int)(eq[ch] & 0xffffu);
                    const int ns = m.sc.fnoff[f], ne = m.sc.nend[f];
                    const uint16_t rb = (uint16_t)(NV == 1 ? kb : (uint32_t)inm.below(b));
                    m.sc.fidxn[pos] = (ain && bin) ? rb : (uint16_t)xv;
                    m.sface[o][pos] = (uint16_t)nf;
                    m.sc.nextp[pos] = (uint16_t)(pos + 1 == ne ? ns : pos + 1);
                    if (!ain) {
                        m.sc.fidxn[pos + 1] = rb;
                        m.sface[o][pos + 1] = (uint16_t)nf;
                        m.sc.nextp[pos + 1] = (uint16_t)(pos + 2 == ne ? ns : pos + 2);
                    }
                }
            }
        }
    }
    poly_sync<C>();

    // 6. successor of every new slot; lid segments (v, u) for intersection pairs (:459-465)
    int M = 0;
    for (int base = 0; base < Spre; base += 64) {
        const int p = base + l;            // < SM: Spre <= SM, a multiple of 64
        const bool pa = p < Spre;
        const uint32_t u0 = m.sc.fidxn[p], n0 = m.sc.nextp[p];
        const uint32_t ua = pa ? u0 : 0u;
        const int np = pa ? (int)n0 : 0;   // stale past Spre: read at 0 instead
        const uint32_t v0 = m.sc.fidxn[np];
        const uint32_t vb = pa ? v0 : 0u;
        m.sedge[o][p] = ua | (vb << 16);   // past Spre: the lid's slots, or unused
        const bool sg = pa && (int)ua >= nk && (int)vb >= nk;
        const uint64_t gm = ballot(sg);
        if (sg && M + below_lane(gm) < C::KM) m.sc.seg[M + below_lane(gm)] = vb | (ua << 16);
        M += popc(gm);
    }
    if (M > C::KM) return kGeneric;   // a convex face has one such segment: at most K / 2

    // 7. lid polygon (:471-495); lane j holds the map entry of new vertex nk + j
    if (M > 0) {
        m.sc.lmap[l] = 0u;   // KM >= 64 entries
        poly_sync<C>();
        for (int t = l; t < M; t += 64) {
            const uint32_t sg = m.sc.seg[t];
            atomicMax(&m.sc.lmap[(int)(sg & 0xffffu) - nk], ((uint32_t)(t + 1) << 16) | (sg >> 16));
        }
        poly_sync<C>();
        const uint32_t mv = (l < nint) ? m.sc.lmap[l] : 0u;
        // the walk runs over lane indices (vertex - nk): lane j holds the lane of new vertex
        // nk + j's successor, or 0xffff where the map has none
        const uint32_t nxl = mv ? (mv & 0xffffu) - (uint32_t)nk : 0xffffu;
        const uint32_t s0 = first_lane_u32(m.sc.seg[0]);   // wave-uniform: the walk is a scalar loop
        const int start = (int)(s0 & 0xffffu) - nk;
        int cv = (int)(s0 >> 16) - nk;
        uint32_t lidv = (uint32_t)start;
        int len = 1;
        while (cv != start && len <= M) {
            if (len >= 64) return kGeneric;
            if (l == len) lidv = (uint32_t)cv;
            ++len;
            cv = (int)rlu(nxl, cv);
            if (cv == 0xffff) break;
        }
        lidv += (uint32_t)nk;
        if (len >= 3) {
            if (Spre + len > C::SM || Fn + 1 > C::FM) return kOverflow;
            const uint32_t nxt = lane_get(lidv, (l + 1 >= len) ? 0 : l + 1);
            if (l < len) {
                m.sedge[o][Spre + l] = lidv | (nxt << 16);
                m.sface[o][Spre + l] = (uint16_t)Fn;
            }
            if (l == 0) {
                m.fstart[o][Fn] = (uint16_t)Spre;
                m.fnb[o][Fn] = nid;
                m.fstart[o][Fn + 1] = (uint16_t)(Spre + len);
            }
            Fn += 1;
            Spre += len;
        }
    }
    poly_sync<C>();
    P.cur = o;
    P.V = nk + nint;
    P.F = Fn;
    P.S = Spre;
    P.nk = nk;
    r2 = withgen ? wave_max(md2) : 0.0;
    return kModified;
}

template <class C>
__device__ inline int clip_any(Mem<C> &m, Poly &P, const Plane pl, int32_t nid, bool withgen, double gx, double gy,
                               double gz, double &r2) {
    if constexpr (C::FAST) {
        // a polytope that leaves the register layout moves the cell to a generic tier
        const int rc = clip_fast<C>(m, P, pl, nid, withgen, gx, gy, gz, r2);
        return rc == kGeneric ? kOverflow : rc;
    } else {
        return clip_generic<C>(m, P, pl, nid, withgen, gx, gy, gz, r2);
    }
}

template <class C>
__device__ double poly_max_radius_sq(Mem<C> &m, const Poly &P, double gx, double gy, double gz) {
    double md = 0.0;  // cell_3d_faces.rs:271-286
    for (int v = lane_id_here(); v < P.V; v += 64) {
        const int c = cur_buf<C>(P);
        const double dx = m.vx[c][v] - gx, dy = m.vy[c][v] - gy, dz = m.vz[c][v] - gz;
        md = gt_max(md, dx * dx + dy * dy + dz * dz);
    }
    return wave_max(md);
}

// Volume and centroid (cell_3d_faces.rs:106-205). Every face has >= 3 entries, so slot p of face
// f at position 1 <= i <= count-2 is triangle p - 2f - 1 of the fan order, and the triangles of one
// 64-slot chunk are consecutive. Per chunk, triangle terms are computed lane-parallel and staged in
// LDS; lanes 0-3 then fold det, det*x, det*y, det*z left in triangle order, chunk after chunk.
template <class C>
__device__ void poly_reduce(Mem<C> &m, const Poly &P, double &vol, double cen[3]) {
    const int l = lane_id_here();
    const int c = cur_buf<C>(P);
    double acc = 0.0;
    for (int base = 0; base < P.S; base += 64) {
        const int p = base + l;
        double det = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
        bool tri = false;
        if (p < P.S) {
            const int f = m.sface[c][p];
            const int fs = m.fstart[c][f], fe = m.fstart[c][f + 1];
            const int cnt = fe - fs, i = p - fs;
            tri = cnt >= 3 && i >= 1 && i <= cnt - 2;
            if (tri) {
                const uint32_t e = m.sedge[c][p];
                const int i0 = (int)(m.sedge[c][fs] & 0xffffu);
                const int i1 = (int)(e & 0xffffu), i2 = (int)(e >> 16);
                const double ax = m.vx[c][i0], ay = m.vy[c][i0], az = m.vz[c][i0];
                const double bx = m.vx[c][i1], by = m.vy[c][i1], bz = m.vz[c][i1];
                const double qx = m.vx[c][i2], qy = m.vy[c][i2], qz = m.vz[c][i2];
                const double cx = by * qz - bz * qy;
                const double cy = bz * qx - bx * qz;
                const double cz = bx * qy - by * qx;
                det = ax * cx + ay * cy + az * cz;
                wx = det * (ax + bx + qx);
                wy = det * (ay + by + qy);
                wz = det * (az + bz + qz);
            }
        }
        const uint64_t tm = ballot(tri);
        const int nt = popc(tm), j = below_lane(tm);   // the chunk's triangles in order
        for (int j0 = 0; j0 < nt; j0 += C::TRI) {      // staged C::TRI at a time
            if (tri && j >= j0 && j < j0 + C::TRI) {
                m.tri[0][j - j0] = det; m.tri[1][j - j0] = wx; m.tri[2][j - j0] = wy; m.tri[3][j - j0] = wz;
            }
            poly_sync<C>();
            if (l < 4) {
                const int ne = nt - j0 < C::TRI ? nt - j0 : C::TRI;
                for (int k = 0; k < ne; ++k) acc += m.tri[l][k];
            }
            poly_sync<C>();
        }
    }
    const double tv = rld(acc, 0), sx = rld(acc, 1), sy = rld(acc, 2), sz = rld(acc, 3);
    vol = fabs(tv / 6.0);
    if (fabs(tv) < 1e-9) {
        cen[0] = cen[1] = cen[2] = 0.0;
    } else {
        const double f = 1.0 / (4.0 * tv);
        cen[0] = sx * f; cen[1] = sy * f; cen[2] = sz * f;
    }
}

// Candidate pre-test: lane w tests its candidate's bisector plane (p = g + d / 2, n = d, exactly
// the plane the clip builds) against vertices [v0, v1) of the current polytope, with the clip's
// own expression for the signed distance. A candidate with no vertex beyond kTol would leave the
// polytope unchanged (kNoChange), so the walk may skip its clip. Kept vertices of a later cut are
// bit-exact copies, so after a cut only the new vertices [nk, V) need testing.
template <class C>
__device__ inline uint64_t pretest(Mem<C> &m, const Poly &P, bool act, double d0, double d1, double d2,
                                   double gx, double gy, double gz, int v0, int v1) {
    bool cut = false;
    if (act) {
        const double px = gx + d0 * 0.5, py = gy + d1 * 0.5, pz = gz + d2 * 0.5;
        const int c = cur_buf<C>(P);
        const double *vx = m.vx[c], *vy = m.vy[c], *vz = m.vz[c];
        for (int v = v0; v < v1; ++v) {
            const double d = (vx[v] - px) * d0 + (vy[v] - py) * d1 + (vz[v] - pz) * d2;
            cut |= d > kTol;
        }
    }
    return ballot(cut);
}

// Consecutive no-change clips after which the fast tier switches the pre-test on for the cell.
#ifndef VD_PRETEST_AFTER
#define VD_PRETEST_AFTER 32
#endif
constexpr int kPretestAfter = VD_PRETEST_AFTER;

// The pre-test costs one pass over the polytope's V vertices per window; a candidate it does not
// clear costs a clip, which exits after its distance pass when nothing is cut. So a window of fewer
// than V / kPretestMinDiv candidates (walks through sparse space gather few per batch of shell
// entries) is clipped without it.
#ifndef VD_PRETEST_MIN_DIV
#define VD_PRETEST_MIN_DIV 4
#endif
constexpr int kPretestMinDiv = VD_PRETEST_MIN_DIV;

// Target bin of shell offset (dx, dy, dz) for a generator in bin (ix, iy, iz): false if outside
// the grid, else its index (algo_3d_grid.rs:175-180,192).
__device__ inline bool entry_bin_index(int dx, int dy, int dz, const GridGeom &G, int ix, int iy, int iz,
                                       uint32_t &b) {
    const int bx = ix + dx, by = iy + dy, bz = iz + dz;
    if (!(bx >= 0 && bx < G.n[0] && by >= 0 && by < G.n[1] && bz >= 0 && bz < G.n[2])) return false;
    b = (uint32_t)bx + (uint32_t)by * (uint32_t)G.n[0] + (uint32_t)bz * (uint32_t)G.n[0] * (uint32_t)G.n[1];
    return true;
}

// Squared distance from a generator at relative position (rx, ry, rz) in its bin to the bin at
// offset (dx, dy, dz) (algo_3d_grid.rs:182-190).
__device__ inline double entry_bin_dist(int dx, int dy, int dz, const GridGeom &G, double rx, double ry, double rz) {
    double ex = dx > 0 ? ((double)dx - rx) * G.cs[0] : dx < 0 ? ((double)(-(dx + 1)) + rx) * G.cs[0] : 0.0;
    double ey = dy > 0 ? ((double)dy - ry) * G.cs[1] : dy < 0 ? ((double)(-(dy + 1)) + ry) * G.cs[1] : 0.0;
    double ez = dz > 0 ? ((double)dz - rz) * G.cs[2] : dz < 0 ? ((double)(-(dz + 1)) + rz) * G.cs[2] : 0.0;
    ex = fmax(ex, 0.0); ey = fmax(ey, 0.0); ez = fmax(ez, 0.0);
    return ex * ex + ey * ey + ez * ez;
}

// Both: false if outside the grid, else the bin's index and distance.
__device__ inline bool entry_bin(int dx, int dy, int dz, const GridGeom &G, int ix, int iy, int iz, double rx,
                                 double ry, double rz, uint32_t &b, double &bdist) {
    if (!entry_bin_index(dx, dy, dz, G, ix, iy, iz, b)) return false;
    bdist = entry_bin_dist(dx, dy, dz, G, rx, ry, rz);
    return true;
}

// Skip-ahead over a stretch of the walk with no candidates (cells that reach into empty space):
// probes kProbe batches of 64 entries from e0 with independent loads and returns how many whole
// batches are inert under the limit lim = 4 r^2: no entry ends the walk (key > lim or end of the
// prefix) and no entry has a non-empty in-grid bin within the distance test. Skipping them is
// exact: the walk would visit them without a clip, so the radius and the order are unchanged.
constexpr int kProbe = 8;

__device__ inline int probe_inert(const ShellCode *shell, uint32_t nshell, const ShellPack &pk, uint32_t e0,
                                  const GridGeom &G, const uint32_t *bin_start, int ix, int iy, int iz,
                                  double rx, double ry, double rz, double lim) {
    const int l = lane_id();
    ShellCode sc[kProbe];
#pragma unroll
    for (int j = 0; j < kProbe; ++j) {
        const uint32_t e = e0 + 64u * (uint32_t)j + (uint32_t)l;
        sc[j] = shell[e < nshell ? e : 0u];
    }
    uint32_t lo[kProbe], hi[kProbe];
    bool live[kProbe];
#pragma unroll
    for (int j = 0; j < kProbe; ++j) {
        const uint32_t e = e0 + 64u * (uint32_t)j + (uint32_t)l;
        int dx, dy, dz;
        shell_unpack(sc[j], pk, dx, dy, dz);
        uint32_t b = 0;
        double bdist = 0.0;
        live[j] = e >= nshell || shell_key(dx, dy, dz, G.cs) > lim;
        lo[j] = hi[j] = 0u;
        if (!live[j] && entry_bin(dx, dy, dz, G, ix, iy, iz, rx, ry, rz, b, bdist) && bdist <= lim) {
            lo[j] = bin_start[b];
            hi[j] = bin_start[b + 1];
        }
    }
    int first = kProbe;
#pragma unroll
    for (int j = kProbe - 1; j >= 0; --j)
        if (ballot(live[j] || hi[j] > lo[j])) first = j;
    return first;
}

// One cell: compute_cell (tessellation.rs:414-475) then the reductions and the exports. The wall
// planes and the neighbour candidates each have a clip site (two inlined copies of the clip).
// PROBE: skip inert runs of shell batches in bulk (probe_inert). The generic tiers always do; the
// fast tiers only in the instantiation that reruns cells whose walk ran past the shell prefix,
// because the probe's registers slow the common walk down.
template <class C, bool SEAL, bool PROBE>
__device__ void process_cell(Mem<C> &m, const CellArgs &A, uint32_t i, uint32_t q, ArenaChunks &ac) {
    const int l = lane_id_here();
    const TessDesc &D = A.desc[A.gen_tess[i]];   // this cell's tessellation (wave-uniform)
    const double gx = uni_d(A.gen[3 * (size_t)i]), gy = uni_d(A.gen[3 * (size_t)i + 1]),
                 gz = uni_d(A.gen[3 * (size_t)i + 2]);
    const uint32_t li = first_lane_u32((uint32_t)(i - D.gen_off));   // index within its tessellation
    const GridGeom G = uni_grid(D.grid);
    // a per-cell list (k_far_lists) replaces the member's prefix for cells that reach across
    // empty space: the same candidates in the same order. Only reruns of such cells have lists,
    // and they run in the skip-ahead instantiation.
    const bool own = PROBE && A.lshell && A.lshell_n[q] != kNoList;
    const ShellCode *shell = uni_ptr(own ? A.lshell + A.lshell_off[q] : D.shell);
    const ShellPack pk = uni_pack(D.pack);
    const uint32_t nshell = first_lane_u32(own ? A.lshell_n[q] : D.nshell);
    const double shell_next = uni_d(own ? A.lshell_next[q] : D.shell_next_d2);
    const uint32_t *bin_start = uni_ptr(A.bin_start + D.bin_off);
    const int nwalls = D.nwalls;
    Poly P;
    poly_box<C>(m, P, D.box_lo, D.box_hi);
    int status = kCellOk;
    double r2 = 0.0;

    // walk state (algo_3d_grid.rs:145-202)
    int ix = 0, iy = 0, iz = 0;
    double rx = 0.0, ry = 0.0, rz = 0.0;
    uint32_t e0 = 0;        // first shell entry of the loaded batch (lane k holds e0 + k)
    int k0 = 0, o0 = 0;     // next unvisited entry of the batch, items of it already visited
    bool load = true;
    double bd2 = 0.0, bdist = 0.0;
    uint32_t bs = 0, bc = 0;
    bool binr = false, bend = false;
    // window state: lane w holds candidate w
    bool win = false;       // a window is open
    double r2b = 0.0;       // radius when the window was gathered
    uint32_t wtot = 0;      // candidates the window would have held without the 64 cap
    int kbc = 64;           // first batch entry that ends the walk under the current radius
    int last = -1;          // window lanes <= last are visited
    uint32_t wko = 0;       // batch lane of the candidate's entry | offset in its bin << 8
    uint32_t wj = 0xffffffffu;
    double wd0 = 0.0, wd1 = 0.0, wd2 = 0.0, wdsq = 0.0;
    uint64_t wvm = 0;       // window lanes whose entry is still live
    int inert = 0;          // consecutive whole batches that held no candidate
    bool pre = !C::FAST;    // pre-test the window's candidates (always in the generic tiers)
    int noc = 0;            // consecutive candidate clips without a change
    uint64_t wcm = ~0ull;   // window lanes whose candidate may cut (pre-test), all without it

    // walls (tessellation.rs:429-436): every plane each wall's cut emits, in order; the cell is
    // done once a wall has left it empty
    uint32_t tk = 0;        // tube walls seen so far (their planes come from a pre-pass)
    for (int wi = 0; wi < nwalls; ++wi) {
        const WallDev &W = D.walls[wi];
        for (uint32_t wp = 0;; ++wp) {
            WallPlane q{};
            bool has;
            if (W.kind == kWallTube) {
                const double *tp = A.tube_planes + 7 * ((size_t)i * A.ntube + tk);
                has = wp == 0 && tp[6] != 0.0;
                q = WallPlane{tp[0], tp[1], tp[2], tp[3], tp[4], tp[5]};
            } else if (W.kind == kWallHost) {
                // the caller's cut planes for this generator, evaluated on the host
                const uint64_t *ho = A.host_off + (size_t)i * A.nhost + W.npts;
                const uint64_t pq = ho[0] + wp;
                has = pq < ho[1];
                if (has) {
                    const double *hp = A.host_planes + 6 * pq;
                    q = WallPlane{hp[0], hp[1], hp[2], hp[3], hp[4], hp[5]};
                }
            } else {
                has = wall_plane(W, wp, gx, gy, gz, q);
            }
            if (!has) break;
            double wr2 = 0.0;
            const int rc = clip_any<C>(m, P, Plane{q.px, q.py, q.pz, q.nx, q.ny, q.nz}, W.id, false,
                                       gx, gy, gz, wr2);
            if (rc < 0) {
                status = (rc == kFace255) ? kCellFace255 : kCellCapacity;
                break;
            }
        }
        if (W.kind == kWallTube) ++tk;
        if (status != kCellOk || P.V == 0) break;
    }

    // the neighbour walk, for a cell the walls left non-empty (tessellation.rs:438-472)
    const bool walk = P.V != 0 && status == kCellOk;
    if (walk) {
        r2 = poly_max_radius_sq<C>(m, P, gx, gy, gz);   // tessellation.rs:438
        ix = (int)first_lane_u32(axis_bin(gx, G.lo[0], G.s[0], G.lim[0]));
        iy = (int)first_lane_u32(axis_bin(gy, G.lo[1], G.s[1], G.lim[1]));
        iz = (int)first_lane_u32(axis_bin(gz, G.lo[2], G.s[2], G.lim[2]));
        rx = uni_d((gx - G.lo[0]) * G.s[0] - (double)ix);
        ry = uni_d((gy - G.lo[1]) * G.s[1] - (double)iy);
        rz = uni_d((gz - G.lo[2]) * G.s[2] - (double)iz);
    }
    while (walk) {
        Plane pl;
        int32_t nid = 0;
        bool withgen = false;
        int cl = -1;
        {
            if (!win) {
                if (load) {
                    if ((!C::FAST || PROBE) && inert >= 2) {
                        // a run of batches without candidates: skip inert batches in bulk
                        const double plim = 4.0 * r2;
                        for (;;) {
                            const int f = probe_inert(shell, nshell, pk, e0, G, bin_start, ix, iy, iz, rx, ry,
                                                      rz, plim);
                            e0 += 64u * (uint32_t)f;
                            if (f < kProbe) break;
                        }
                    }
                    const uint32_t e = e0 + (uint32_t)l;
                    bend = e >= nshell;
                    binr = false;
                    bc = 0;
                    bdist = 0.0;
                    int dx = 0, dy = 0, dz = 0;
                    if (!bend) {
                        shell_unpack(shell[e], pk, dx, dy, dz);
                        bd2 = shell_key(dx, dy, dz, G.cs);
                        uint32_t b = 0;
                        binr = entry_bin_index(dx, dy, dz, G, ix, iy, iz, b);
                        if (binr) {
                            bs = bin_start[b];
                            bc = bin_start[b + 1] - bs;
                        }
                    } else {
                        bd2 = (e == nshell) ? shell_next : DBL_MAX;
                    }
                    // the bin distance decides only whether a bin's generators are taken: empty
                    // bins (most of a batch where generators are sparse) skip it
                    if (bc) bdist = entry_bin_dist(dx, dy, dz, G, rx, ry, rz);
                    load = false;
                }
                // gather the next window of candidates under the current radius
                r2b = r2;
                const double lim = 4.0 * r2;
                const bool committed = (l == k0) && (o0 > 0);
                const bool brk = (l >= k0) && !committed && (bend || bd2 > lim);   // :170
                kbc = ffs64(ballot(brk));
                const bool take = (l >= k0) && (l < kbc) && (committed || (binr && bdist <= lim));
                const uint32_t skip = (l == k0) ? (uint32_t)o0 : 0u;
                const uint32_t cnt = take ? bc - skip : 0u;
                const uint32_t inc = wave_incl_scan(cnt);
                const uint32_t pos = inc - cnt;
                wtot = rlu(inc, 63);
                inert = (wtot == 0u && kbc == 64 && k0 == 0) ? inert + 1 : 0;
                const int nwin = wtot < 64u ? (int)wtot : 64;
                if (take) {
                    for (uint32_t t = 0; t < cnt && pos + t < 64u; ++t) {
                        m.win.q[pos + t] = bs + skip + t;
                        m.win.ent[pos + t] = (uint16_t)l;
                        m.win.off[pos + t] = (uint16_t)(skip + t);
                    }
                }
                poly_sync<C>();
                const bool wa = l < nwin;
                const uint32_t wq = wa ? m.win.q[l] : 0u;
                wko = wa ? ((uint32_t)m.win.ent[l] | ((uint32_t)m.win.off[l] << 8)) : 0u;
                wj = 0xffffffffu;
                wd0 = wd1 = wd2 = wdsq = 0.0;
                if (wa) {
                    const double4 q = A.spos[wq];
                    wj = (uint32_t)__double_as_longlong(q.w);
                    wd0 = q.x - gx; wd1 = q.y - gy; wd2 = q.z - gz;
                    wdsq = wd0 * wd0 + wd1 * wd1 + wd2 * wd2;   // tessellation.rs:447-456
                }
                wvm = ballot(wa);
                wcm = (pre && kPretestMinDiv * nwin >= P.V)
                          ? pretest<C>(m, P, wa && wj != li, wd0, wd1, wd2, gx, gy, gz, 0, P.V) : ~0ull;
                last = -1;
                win = true;
                poly_sync<C>();
            }
            // next candidate of the window that passes the distance test (tessellation.rs:458)
            const double lim = 4.0 * r2;
            const uint64_t above = last < 0 ? ~0ull : (last >= 63 ? 0ull : ~((2ull << last) - 1ull));
            const uint64_t pm = wvm & wcm & above & ballot(wj != li) & ballot(!(wdsq > lim));
            if (pm == 0) {
                // window exhausted: find where the walk continues
                win = false;
                bool walk_end = false;
                if (wtot > 64u) {
                    const uint32_t ko63 = rlu(wko, 63);
                    const int k63 = (int)(ko63 & 0xffu), off63 = (int)(ko63 >> 8);
                    if ((wvm >> 63) & 1ull) {
                        if ((uint32_t)(off63 + 1) < rlu(bc, k63)) { k0 = k63; o0 = off63 + 1; }
                        else { k0 = k63 + 1; o0 = 0; }
                    } else if (k63 >= kbc) {
                        walk_end = true;
                    } else {
                        k0 = k63 + 1; o0 = 0;
                    }
                } else if (kbc < 64) {
                    walk_end = true;
                } else {
                    k0 = 64;
                }
                if (walk_end) {
                    // the walk stops at entry kbc: a real break, or the end of the shell prefix
                    if (rl((int)bend, kbc)) {
                        const uint32_t e = e0 + (uint32_t)kbc;
                        if (!(e == nshell && shell_next > lim)) status = kCellShell;
                    }
                    break;
                }
                if (k0 >= 64) { e0 += 64; k0 = 0; o0 = 0; load = true; }
                continue;
            }
            cl = ffs64(pm);
            const double d0 = rld(wd0, cl), d1 = rld(wd1, cl), d2 = rld(wd2, cl);
            pl = Plane{gx + d0 * 0.5, gy + d1 * 0.5, gz + d2 * 0.5, d0, d1, d2};
            nid = (int32_t)rlu(wj, cl);
            withgen = true;
            last = cl;
        }

        double nr2 = 0.0;
        const int rc = clip_any<C>(m, P, pl, nid, withgen, gx, gy, gz, nr2);   // tessellation.rs:462-463
        if (rc < 0) {
            status = (rc == kFace255) ? kCellFace255 : kCellCapacity;
            break;
        }
        if (rc == kNoChange) {
            if (++noc >= kPretestAfter) pre = true;   // takes effect at the next window
            continue;
        }
        noc = 0;
        // a cut (tessellation.rs:462-469)
        if (P.V == 0) nr2 = 0.0;
        const uint32_t koc = rlu(wko, cl);
        const int kc = (int)(koc & 0xffu), offc = (int)(koc >> 8);
        r2 = nr2;
        if (nr2 > r2b) {
            // the radius grew: entries dropped from this window may now pass, gather again
            if ((uint32_t)(offc + 1) < rlu(bc, kc)) { k0 = kc; o0 = offc + 1; }
            else { k0 = kc + 1; o0 = 0; }
            if (k0 >= 64) { e0 += 64; k0 = 0; o0 = 0; load = true; }
            win = false;
            continue;
        }
        // the radius shrank or stayed: re-filter the rest of the window with it
        if (pre && wcm != ~0ull) {
            // candidates cleared by the pre-test are tested against the new vertices only
            const uint64_t todo = wvm & ~wcm & (cl >= 63 ? 0ull : ~((2ull << cl) - 1ull));
            if (todo) wcm |= pretest<C>(m, P, (todo >> l) & 1ull, wd0, wd1, wd2, gx, gy, gz, P.nk, P.V);
        }
        const double lim = 4.0 * r2;
        kbc = ffs64(ballot(l > kc) & (ballot(bend) | ballot(bd2 > lim)));
        const int wk = (int)(wko & 0xffu);
        const double wbd = lane_get_d(bdist, wk);   // bin distance of the candidate's entry
        wvm &= ballot(l <= cl) | ballot(wk == kc) | (ballot(wk < kbc) & ballot(wbd <= lim));
    }
    poly_sync<C>();

    if constexpr (SEAL) {
        // calculate_sealed (tessellation.rs:266-315): for every neighbour of the previous pass, in
        // face order, and every curved wall both cells touched, the wall's plane for the
        // neighbour's generator clips this cell when this generator is on its valid side
        const uint32_t mi = A.seal_mask[i];
        uint32_t nlog = 0;
        if (status == kCellOk && mi) {
            const uint32_t q0 = A.seal_off[i], q1 = A.seal_off[i + 1];
            int32_t *lg = A.seal_log + 3 * (size_t)A.seal_slots * q0;
            for (uint32_t q = q0; q < q1 && status == kCellOk; ++q) {
                const uint32_t j = A.seal_nb[q];
                const size_t gj = D.gen_off + j;
                const uint32_t mj = A.seal_mask[gj];
                const double nx = A.gen[3 * gj], ny = A.gen[3 * gj + 1], nz = A.gen[3 * gj + 2];
                uint32_t tk = 0;
                for (int w = 0; w < nwalls; ++w) {
                    const WallDev &W = D.walls[w];
                    const bool tube = W.kind == kWallTube;
                    const bool use = !wall_planar(W) && ((mi >> w) & 1u) && ((mj >> w) & 1u);
                    // every plane the wall's cut emits for the neighbour's generator, in order
                    // (one for the built-in curved walls)
                    for (uint32_t wp = 0; use; ++wp) {
                    WallPlane q{};
                    bool has = false;
                    if (tube) {
                        const double *tp = A.tube_planes + 7 * (gj * A.ntube + tk);
                        has = wp == 0 && tp[6] != 0.0;
                        q = WallPlane{tp[0], tp[1], tp[2], tp[3], tp[4], tp[5]};
                    } else if (W.kind == kWallHost) {
                        const uint64_t *ho = A.host_off + gj * A.nhost + W.npts;
                        const uint64_t pq = ho[0] + wp;
                        has = pq < ho[1];
                        if (has) {
                            const double *hp = A.host_planes + 6 * pq;
                            q = WallPlane{hp[0], hp[1], hp[2], hp[3], hp[4], hp[5]};
                        }
                    } else {
                        has = wall_plane(W, wp, nx, ny, nz, q);
                    }
                    if (!has) break;
                    double dot = 0.0;
                    dot += (gx - q.px) * q.nx;
                    dot += (gy - q.py) * q.ny;
                    dot += (gz - q.pz) * q.nz;
                    if (!(dot <= 1e-9)) continue;
                    double wr2 = 0.0;
                    const int rc = clip_any<C>(m, P, Plane{q.px, q.py, q.pz, q.nx, q.ny, q.nz}, W.id,
                                               false, gx, gy, gz, wr2);
                    if (rc < 0) {
                        status = (rc == kFace255) ? kCellFace255 : kCellCapacity;
                        break;
                    }
                    if (rc == kModified) {
                        if (l == 0) {
                            lg[3 * nlog] = (int32_t)li;
                            lg[3 * nlog + 1] = (int32_t)j;
                            lg[3 * nlog + 2] = W.id;
                        }
                        ++nlog;
                    }
                    }
                    if (tube) ++tk;
                    if (status != kCellOk) break;
                }
            }
        }
        if (l == 0) A.seal_logcnt[i] = nlog;
    }

    // outputs; a cell that walked past its shell prefix leaves its radius r^2 in the volume slot,
    // so the prefix can be extended to the length its rerun needs
    const int lo = lane_id_here();   // once-per-cell code: see lane_id_here
    if (lo == 0) A.status[i] = status;
    if (status != kCellOk) {
        if (status == kCellShell && lo == 0) A.vol[i] = r2;
        poly_sync<C>();
        return;
    }
    double vol, cen[3];
    poly_reduce<C>(m, P, vol, cen);
    if (lo == 0) {
        A.vol[i] = vol;
        A.cen[3 * (size_t)i] = cen[0];
        A.cen[3 * (size_t)i + 1] = cen[1];
        A.cen[3 * (size_t)i + 2] = cen[2];
        A.nface[i] = (uint32_t)P.F;
        A.nvert[i] = (uint32_t)P.V;
    }
    if (A.vtx) {
        const uint64_t ov = arena_take(A, ac, 0, (uint64_t)P.V);
        const uint64_t of = arena_take(A, ac, 1, (uint64_t)P.F);
        const uint64_t oi = arena_take(A, ac, 2, (uint64_t)P.S);
        if (ov + P.V > A.arena_cap[0] || of + P.F > A.arena_cap[1] || oi + P.S > A.arena_cap[2]) {
            if (lo == 0) A.status[i] = kCellArena;
        } else {
            const int c = cur_buf<C>(P);
            for (int v = lo; v < P.V; v += 64) {
                A.vtx[3 * (ov + v)] = m.vx[c][v];
                A.vtx[3 * (ov + v) + 1] = m.vy[c][v];
                A.vtx[3 * (ov + v) + 2] = m.vz[c][v];
            }
            for (int f = lo; f < P.F; f += 64) {
                A.fcnt[of + f] = (uint8_t)(m.fstart[c][f + 1] - m.fstart[c][f]);
                A.fnb[of + f] = m.fnb[c][f];
            }
            for (int s = lo; s < P.S; s += 64) A.fidx[oi + s] = (uint16_t)(m.sedge[c][s] & 0xffffu);
            if (lo == 0) {
                A.cell_off[3 * (size_t)i] = ov;
                A.cell_off[3 * (size_t)i + 1] = of;
                A.cell_off[3 * (size_t)i + 2] = oi;
            }
        }
    } else if (A.fnb) {
        // reduced mode keeps the neighbour ids (face order) only
        const uint64_t of = arena_take(A, ac, 1, (uint64_t)P.F);
        if (of + P.F > A.arena_cap[1]) {
            if (lo == 0) A.status[i] = kCellArena;
        } else {
            for (int f = lo; f < P.F; f += 64) A.fnb[of + f] = m.fnb[cur_buf<C>(P)][f];
            if (lo == 0) A.cell_off[3 * (size_t)i + 1] = of;
        }
    }
    poly_sync<C>();
}

template <class C, bool SEAL, bool PROBE>
__device__ inline void cells_loop(Mem<C> &m, const CellArgs &A) {
    ArenaChunks ac;
    for (;;) {
        // A.qgrain consecutive list entries per queue atomic
        uint32_t q = 0;
        if (lane_id() == 0) q = atomicAdd(A.queue, A.qgrain);
        q = first_lane_u32(q);
        if (q >= A.ncells) break;
        const uint32_t qe = A.ncells - q < A.qgrain ? A.ncells : q + A.qgrain;
        for (; q < qe; ++q) process_cell<C, SEAL, PROBE>(m, A, A.cells[q], q, ac);
    }
    // entries the wave's cells occupy (the chunks' unused tails excluded)
    if (lane_id() == 0)
        for (int k = 0; k < 3; ++k)
            if (ac.used[k]) atomicAdd((unsigned long long *)&A.arena_top[3 + k], (unsigned long long)ac.used[k]);
}

template <class C, bool SEAL, bool PROBE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(C::WAVES, 8)))
void k_cells(CellArgs A) {
    __shared__ Mem<C> m;
    cells_loop<C, SEAL, PROBE>(m, A);
}

// The global-memory tier: wave b builds its cells in slot b of A.gmem (the launch has at most as
// many waves as slots).
template <class C, bool SEAL>
__global__ __launch_bounds__(64) void k_cells_global(CellArgs A) {
    Mem<C> &m = reinterpret_cast<Mem<C> *>(A.gmem)[blockIdx.x];
    cells_loop<C, SEAL, true>(m, A);
}

// Tier 0: register fast path (one vertex per lane), small LDS (many waves per CU), compiled for
// 5, 6 and 7 resident waves per SIMD (the register budget each allows; the launch picks one).
// Tier 1: the same path with two vertices per lane (128 vertices). Tiers 2-3: generic path for
// cells that outgrow the tier before, with 256 and 1024 vertices (the last one wave per CU).
template <int W>
using Tier0W = Cap<64, 192, 48, 64, true, W, false, (W >= 8)>;
using Tier0 = Tier0W<kTier0WavesDefault>;   // sizes (tier_max_vertices, tier_lds_bytes)
using Tier1 = Cap<128, 384, 64, 64, true, 4>;
using Tier2 = Cap<256, 768, 128, 384, false, 1>;
using Tier3 = Cap<1024, 2560, 320, 1280, false, 1>;
// Tier 4: the generic path with its polytope in global memory, for cells beyond LDS. Face-slot
// offsets are 16-bit in every tier and a crossing's vertex id shares 16 bits with a flag in
// clip_generic, so it holds up to 32,767 vertices and faces, 65,535 face slots and 32,767
// crossings per clip (the reference's own face indices are u16, cell_3d_faces.rs:29).
using TierG = Cap<32767, 65535, 32767, 32767, false, 1, true>;

// Work-list entries whose cell stopped at the end of its shell prefix and whose walk covers about
// min_entries bins or more: 4/3 pi (2r)^3 over the bin volume (a size estimate only; the lists
// change no result). Their limit: the first prefix entry with key > L = 4 r^2 (1 + kListMargin)
// (keys do not decrease), and the key their list ends at. A walk whose limit reaches L at the end
// of its list reports it ran past it, as at the end of a prefix, and reruns on a longer list.
__global__ void k_far_pick(const uint32_t *list, uint32_t n, const int32_t *status, const double *vol,
                           const uint32_t *gen_tess, const TessDesc *desc, double min_entries, uint32_t *flags,
                           uint32_t *limit, double *next) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const uint32_t i = list[q];
    uint32_t f = 0u, lim = 0u;
    double nk = 0.0;
    if (status[i] == kCellShell) {
        const TessDesc &D = desc[gen_tess[i]];
        const GridGeom &G = D.grid;
        const double r2 = vol[i];
        f = 33.510321638291124 * r2 * sqrt(r2) >= min_entries * G.cs[0] * G.cs[1] * G.cs[2];
        if (f) {
            const double l4 = 4.0 * r2 * (1.0 + kListMargin);
            uint32_t lo = 0u, hi = D.nshell;
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo) / 2u;
                int dx, dy, dz;
                shell_unpack(D.shell[mid], D.pack, dx, dy, dz);
                if (shell_key(dx, dy, dz, G.cs) > l4) hi = mid;
                else lo = mid + 1u;
            }
            lim = lo;
            nk = lo < D.nshell ? l4 : D.shell_next_d2;
        }
    }
    flags[q] = f;
    limit[q] = lim;
    next[q] = nk;
}

// Per-cell shell lists: block b takes kListChunk entries of one selected cell, 4 consecutive
// entries per thread, and keeps those below the cell's limit whose bin lies in the grid and holds
// a generator. An empty or out-of-grid entry gives the walk no candidate and keys do not decrease
// along the prefix, so a walk over the kept entries visits the same candidates in the same order
// and ends where the walk over the prefix ends: at the first kept entry past its limit, or at the
// end of the list, whose next key (L for a list cut below the prefix's end, else the prefix's next
// key) decides as at the end of a prefix whether the list was long enough.
__global__ __launch_bounds__(256) void k_far_lists(const uint32_t *cells, const uint32_t *limit,
                                                   const uint32_t *cp, uint32_t nsel, const TessDesc *desc,
                                                   const uint32_t *gen_tess, const double *gen,
                                                   const uint32_t *bin_start, uint32_t *counts,
                                                   const uint32_t *offs, ShellCode *out) {
    __shared__ uint32_t wsum[4];
    const uint32_t b = blockIdx.x;
    uint32_t lo = 0u, hi = nsel;   // the selected cell s with cp[s] <= b < cp[s + 1]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) / 2u;
        if (cp[mid] <= b) lo = mid;
        else hi = mid;
    }
    const uint32_t s = lo, chunk = b - cp[s], end = limit[s];
    const int t = (int)threadIdx.x, w = t >> 6, l = t & 63;
    const uint32_t i = cells[s];
    const TessDesc &D = desc[gen_tess[i]];
    const GridGeom &G = D.grid;
    const double gx = gen[3 * (size_t)i], gy = gen[3 * (size_t)i + 1], gz = gen[3 * (size_t)i + 2];
    const int ix = (int)axis_bin(gx, G.lo[0], G.s[0], G.lim[0]);
    const int iy = (int)axis_bin(gy, G.lo[1], G.s[1], G.lim[1]);
    const int iz = (int)axis_bin(gz, G.lo[2], G.s[2], G.lim[2]);
    const uint32_t *bs = bin_start + D.bin_off;
    ShellCode code[4];
    uint32_t keep = 0u;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t e = chunk * kListChunk + 4u * (uint32_t)t + (uint32_t)j;
        code[j] = 0u;
        if (e < end) {
            code[j] = D.shell[e];
            int dx, dy, dz;
            shell_unpack(code[j], D.pack, dx, dy, dz);
            const int bx = ix + dx, by = iy + dy, bz = iz + dz;
            if (bx >= 0 && bx < G.n[0] && by >= 0 && by < G.n[1] && bz >= 0 && bz < G.n[2]) {
                const uint32_t bb = (uint32_t)bx + (uint32_t)by * (uint32_t)G.n[0] +
                                    (uint32_t)bz * (uint32_t)G.n[0] * (uint32_t)G.n[1];
                if (bs[bb + 1] > bs[bb]) keep |= 1u << j;
            }
        }
    }
    const uint32_t cnt = (uint32_t)__popc(keep);
    const uint32_t inc = wave_incl_scan(cnt);
    if (l == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t before = inc - cnt;
    for (int k = 0; k < w; ++k) before += wsum[k];
    if (!out) {
        if (t == 0) counts[b] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        return;
    }
    uint64_t o = (uint64_t)offs[b] + before;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if ((keep >> j) & 1u) out[o++] = code[j];
}

}  // namespace

hipError_t launch_far_pick(const uint32_t *list, uint32_t n, const int32_t *status, const double *vol,
                           const uint32_t *gen_tess, const TessDesc *desc, double min_entries,
                           uint32_t *flags, uint32_t *limit, double *next, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_far_pick<<<(n + 255) / 256, 256, 0, st>>>(list, n, status, vol, gen_tess, desc, min_entries, flags, limit,
                                                next);
    return hipGetLastError();
}

hipError_t launch_far_lists(const uint32_t *cells, const uint32_t *limit, const uint32_t *cp, uint32_t nsel,
                            uint32_t blocks, const TessDesc *desc, const uint32_t *gen_tess, const double *gen,
                            const uint32_t *bin_start, uint32_t *counts, const uint32_t *offs, ShellCode *out,
                            hipStream_t st) {
    if (nsel == 0 || blocks == 0) return hipSuccess;
    static_assert(kListChunk == 4 * 256, "one block covers a chunk, 4 entries per thread");
    k_far_lists<<<blocks, 256, 0, st>>>(cells, limit, cp, nsel, desc, gen_tess, gen, bin_start, counts, offs, out);
    return hipGetLastError();
}

int tier_count() { return 5; }

int tier_max_vertices(int tier) {
    return tier == 0 ? Tier0::VM : tier == 1 ? Tier1::VM : tier == 2 ? Tier2::VM : tier == 3 ? Tier3::VM : TierG::VM;
}

size_t tier_lds_bytes(int tier) {
    return tier == 0 ? sizeof(Mem<Tier0>) : tier == 1 ? sizeof(Mem<Tier1>)
         : tier == 2 ? sizeof(Mem<Tier2>) : tier == 3 ? sizeof(Mem<Tier3>) : 0;
}

size_t tier_global_bytes() { return sizeof(Mem<TierG>); }

template <bool SEAL, bool PROBE>
void launch_tier(int tier, const CellArgs &a, int grid_waves, hipStream_t st, int t0_waves) {
    if (tier == 0) {
        if (t0_waves == 5) k_cells<Tier0W<5>, SEAL, PROBE><<<grid_waves, 64, 0, st>>>(a);
        else if (t0_waves == 6) k_cells<Tier0W<6>, SEAL, PROBE><<<grid_waves, 64, 0, st>>>(a);
        else if (t0_waves == 8) k_cells<Tier0W<8>, SEAL, PROBE><<<grid_waves, 64, 0, st>>>(a);
        else k_cells<Tier0W<7>, SEAL, PROBE><<<grid_waves, 64, 0, st>>>(a);
    }
    else if (tier == 1) k_cells<Tier1, SEAL, PROBE><<<grid_waves, 64, 0, st>>>(a);
    else if (tier == 2) k_cells<Tier2, SEAL, true><<<grid_waves, 64, 0, st>>>(a);
    else if (tier == 3) k_cells<Tier3, SEAL, true><<<grid_waves, 64, 0, st>>>(a);
    else k_cells_global<TierG, SEAL><<<grid_waves, 64, 0, st>>>(a);
}

bool tier0_waves_valid(int w) { return w >= 5 && w <= 8; }

hipError_t launch_cells(int tier, const CellArgs &a, int grid_waves, hipStream_t st, bool seal, bool probe,
                        int t0_waves) {
    if (a.ncells == 0) return hipSuccess;
    if (tier == 4 && !a.gmem) return hipErrorInvalidValue;
    if (a.qgrain == 0 || !tier0_waves_valid(t0_waves)) return hipErrorInvalidValue;
    if (seal) {
        if (probe) launch_tier<true, true>(tier, a, grid_waves, st, t0_waves);
        else launch_tier<true, false>(tier, a, grid_waves, st, t0_waves);
    } else {
        if (probe) launch_tier<false, true>(tier, a, grid_waves, st, t0_waves);
        else launch_tier<false, false>(tier, a, grid_waves, st, t0_waves);
    }
    return hipGetLastError();
}

hipError_t cells_occupancy(int tier, int *blocks_per_cu, int t0_waves) {
    if (tier == 0) {
        if (t0_waves == 5)
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_cells<Tier0W<5>, false, false>, 64, 0);
        if (t0_waves == 6)
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_cells<Tier0W<6>, false, false>, 64, 0);
        if (t0_waves == 8)
            return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_cells<Tier0W<8>, false, false>, 64, 0);
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_cells<Tier0W<7>, false, false>, 64, 0);
    }
    if (tier == 1)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_cells<Tier1, false, false>, 64, 0);
    if (tier == 2)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_cells<Tier2, false, true>, 64, 0);
    if (tier == 3)
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_cells<Tier3, false, true>, 64, 0);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_cells_global<TierG, false>, 64, 0);
}

}  // namespace vd
