/* vd.h -- C ABI of libvoronoid_hip.so, the MI355X (gfx950) implementation of voronoid's
 * Tessellation<3, Cell3DFaces, Algorithm3DGrid> hot path.
 *
 * Each entry point replaces one call of the reference crate (mdt-re/vorothree, Rust crate
 * `voronoid`); the reference interface is cited as file:line next to it. Plain pointers and sizes
 * only. Conventions:
 *   - every int-returning call returns VD_OK (0) or a negative VD_E* code; vd_last_error() gives
 *     the message of the calling thread's last failure;
 *   - output buffers are caller-owned host memory; the library owns all device memory;
 *   - one handle is not safe for concurrent use (the reference takes &mut self); separate handles
 *     may be used from separate threads and run on their own HIP streams (reruns of cells that
 *     reach across empty space also use one of a few side streams per device that the library
 *     shares between handles);
 *   - where the reference panics (wall id > WALL_ID_MAX) the call returns VD_EINVAL; where it
 *     returns None (get_cell out of range) the call returns VD_ERANGE.
 */
#ifndef VD_H
#define VD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VD_OK 0
#define VD_EINVAL (-1)    /* invalid argument (reference: panic) */
#define VD_ERANGE (-2)    /* index out of range (reference: None) */
#define VD_ENOMEM (-3)    /* device or host allocation failed */
#define VD_EDEVICE (-4)   /* HIP runtime error */
#define VD_ECAPACITY (-5) /* a cell exceeded every polytope capacity tier (32,767 vertices, 65,535
                             face slots), has a face of 256 or more vertices (the reference's u8
                             face count would wrap), or reruns did not finish */
#define VD_ESTATE (-6)    /* call out of order (e.g. results before calculate) */

#define VD_WALL_ID_MAX (-1000) /* src/wall/mod.rs:8 */

/* A wall (Wall::new(id, geometry), src/wall/mod.rs:24-31) for vd_tess_add_wall, batch members and
 * server jobs. id <= VD_WALL_ID_MAX. Per kind (src/wall/wall_3d.rs):
 *   VD_WALL_PLANE      PlaneGeometry::new(a, b)                     :23-27  (b normalised)
 *   VD_WALL_SPHERE     SphereGeometry::new(a, r)                    :67-69
 *   VD_WALL_CYLINDER   CylinderGeometry::new(a, b, r)               :124-128 (b normalised)
 *   VD_WALL_CONE       ConeGeometry::new(a, b, r = half-angle)      :194-198 (b normalised)
 *   VD_WALL_TORUS      TorusGeometry::new(a, b, r = major, r2 = minor) :285-289 (b normalised)
 *   VD_WALL_POLYHEDRON ConvexPolyhedronGeometry { planes }          :462-491; pts = npts planes,
 *                      6 doubles each (point xyz, outward normal xyz), used as given
 *   VD_WALL_TUBE       the tube around a sampled curve shared by TrefoilKnotGeometry (:370-456,
 *                      always closed), CubicBezierGeometry (:718-815) and CatmullRomGeometry
 *                      (:821-961): pts = npts samples (xyz), r = tube radius, closed = 1 when the
 *                      last sample joins the first. The samples are the geometry's own `samples`.
 * pts is read during the call that receives the wall; the library keeps a copy. */
#define VD_WALL_PLANE 0
#define VD_WALL_SPHERE 1
#define VD_WALL_CYLINDER 2
#define VD_WALL_CONE 3
#define VD_WALL_TORUS 4
#define VD_WALL_POLYHEDRON 5
#define VD_WALL_TUBE 6
#define VD_WALL_MAX_POINTS 65536 /* planes of a polyhedron or samples of a tube */
typedef struct vd_wall {
    int32_t id;
    int32_t kind;
    double a[3];
    double b[3];
    double r;
    double r2;
    const double *pts;
    uint32_t npts;
    int32_t closed;
} vd_wall;

typedef struct vd_tess vd_tess;

/* Tessellation::new(bounds, Algorithm3DGrid::new(nx, ny, nz, &bounds))
 *   src/tessellation.rs:25-36, src/algorithm/algo_3d_grid.rs:47-95, src/bounds/mod.rs:8-12.
 * Binds the calling thread's current HIP device and creates the handle's stream. */
int vd_tess_create(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny,
                   uint32_t nz, vd_tess **out);

/* Drop of Tessellation: frees all device memory of the handle. */
void vd_tess_destroy(vd_tess *t);

/* add_wall(Wall::new(id, Box::new(PlaneGeometry::new(point, normal))))
 *   src/tessellation.rs:197-200 (prune + re-bin), src/wall/mod.rs:24-32 (id check),
 *   src/wall/wall_3d.rs:23-27 (normal normalised; zero normal becomes (0,0,1)). */
int vd_tess_add_wall_plane(vd_tess *t, int32_t id, const double point[3], const double normal[3]);

/* add_wall(Wall::new(id, Box::new(SphereGeometry::new(center, radius))))
 *   src/tessellation.rs:197-200, src/wall/wall_3d.rs:66-100. */
int vd_tess_add_wall_sphere(vd_tess *t, int32_t id, const double center[3], double radius);
/* add_wall for any wall kind of vd_wall (the same pruning as the two calls above). */
int vd_tess_add_wall(vd_tess *t, const vd_wall *wall);

/* A wall whose geometry stays with the caller: Wall::new(id, Box<dyn WallGeometry>) for any
 * geometry (src/wall/mod.rs:24-31, trait :53-71), evaluated on the host through these callbacks.
 *   contains(ctx, p): WallGeometry::contains (nonzero: p is in the valid region);
 *   cut(ctx, g, planes, cap): WallGeometry::cut for generator g: writes up to cap planes, 6
 *     doubles each (a point on the plane, the normal pointing out of the valid region), in the
 *     order the geometry emits them, and returns how many it emits; a count above cap makes the
 *     library call again with room for all (at most VD_WALL_MAX_POINTS per generator);
 *   planar: WallGeometry::is_planar (calculate_sealed and prune_boundaries treat non-planar walls
 *     as curved).
 * The callbacks run on the calling thread during vd_tess_add_wall_planes_per_generator,
 * vd_tess_set_generators(_device), vd_tess_random_generators and vd_tess_relax (contains for every
 * input point or candidate, cut for every kept generator) and vd_tess_set_generator (both for the
 * moved generator). ctx must stay valid while the handle holds the wall. */
typedef struct vd_wall_host {
    void *ctx;
    int (*contains)(void *ctx, const double p[3]);
    int (*cut)(void *ctx, const double g[3], double *planes, int cap);
    int planar;
} vd_wall_host;

/* add_wall(Wall::new(id, Box::new(geometry))) for a geometry of the caller's: the prune of
 *   src/tessellation.rs:175-200 with wall->contains, and at every calculate() the planes of
 *   wall->cut for each generator clip its cell in wall order (tessellation.rs:429-436). */
int vd_tess_add_wall_planes_per_generator(vd_tess *t, int32_t id, const vd_wall_host *wall);

/* clear_walls(), src/tessellation.rs:203-205 (generators are kept as they are). */
int vd_tess_clear_walls(vd_tess *t);

/* set_generators(&[f64]) with a host buffer of n points, xyz interleaved.
 *   src/tessellation.rs:43-67 (wall filter in input order) + algo_3d_grid.rs:111-130 (binning). */
int vd_tess_set_generators(vd_tess *t, const double *xyz, size_t n);

/* The same with the points already in device memory (HBM); d_xyz is not retained. */
int vd_tess_set_generators_device(vd_tess *t, const double *d_xyz, size_t n);

/* random_generators(count), src/tessellation.rs:96-122: candidates uniform in the bounding box,
 * kept when every wall contains them, until `count` are kept or count * 1000 candidates were drawn;
 * the kept points become the generators (then as set_generators). The reference seeds its
 * generator from thread_rng, so its points are not reproducible; here candidate i is drawn from
 * outputs 3i, 3i + 1, 3i + 2 of the SplitMix64 stream of `seed` (u = (x >> 11) * 2^-53 per axis,
 * then min + u * (max - min) as in the reference), so equal seeds give equal generators. The
 * candidates are drawn and filtered on the device in batches; the caller's walls (add_wall_planes_
 * per_generator) are asked on the host for every candidate of a batch. */
int vd_tess_random_generators(vd_tess *t, size_t count, uint64_t seed);

/* set_generator(index, &generator), src/tessellation.rs:71-92: moves generator `index` to xyz if
 * every wall contains xyz (otherwise nothing changes and the call returns VD_OK, as the reference
 * ignores it); the bins follow update_generator (src/algorithm/algo_3d_grid.rs:132-143: swap_remove
 * from the old bin, push to the end of the new one), which changes the visiting order of later
 * calculate() calls exactly as in the reference. An index past the generators returns VD_ERANGE
 * (the reference ignores it). As after set_generators, count_cells() is 0 until the next
 * calculate(). The first call after a binning reads the generators back once and holds bins and
 * positions on the host (O(n)); each call then changes only those (O(size of the two bins)), with
 * no device work unless the caller's walls (add_wall_planes_per_generator) are present, whose
 * planes for the moved generator go up at once; the next calculate() uploads bins and positions
 * in one pass (O(n + bins)). */
int vd_tess_set_generator(vd_tess *t, size_t index, const double xyz[3]);

/* calculate(), src/tessellation.rs:213-230: every cell, results kept on the device. */
int vd_tess_calculate(vd_tess *t);

/* relax(), src/tessellation.rs:482-495: each generator moves to its cell's centroid (an empty
 * cell keeps its position), then set_generators. Requires calculate() first. */
int vd_tess_relax(vd_tess *t);

/* calculate_sealed(), src/tessellation.rs:234-318: calculate(), then every cell with faces on a
 * curved wall (not plane or polyhedron) is clipped, for each of its neighbours in face order and
 * each curved wall both touch, by that wall's plane for the neighbour's generator, when this
 * generator is on the plane's valid side ((g - p).n <= 1e-9). The seal log holds one
 * (cell, neighbour, wall id) triple per clip that changed the cell, in cell order. */
int vd_tess_calculate_sealed(vd_tess *t);

/* prune_boundaries(), src/tessellation.rs:324-385 (needs full cells): for each cell, neighbour
 * (first occurrence in face order) and curved wall both touch, the vertices of the cell's faces on
 * that wall that also lie on its face shared with the neighbour (Cell3DFaces::shared_vertices,
 * src/cell/cell_3d_faces.rs:536-568). The prune log holds (cell, neighbour, wall id, vertex count)
 * per non-empty entry, the position log their vertices (xyz). calculate() clears both logs. */
int vd_tess_prune_boundaries(vd_tess *t);

/* seal_log: up to cap ints (3 per entry); *count receives the total number of ints. */
int vd_tess_get_seal_log(const vd_tess *t, int32_t *out, size_t cap, size_t *count);
/* prune_log and prune_pos_log: up to icap ints and pcap doubles; counts[0], counts[1] receive
 * their totals. */
int vd_tess_get_prune_log(const vd_tess *t, int32_t *ints, size_t icap, double *pos, size_t pcap,
                          size_t counts[2]);

/* count_generators() / count_cells(), src/tessellation.rs:498-505. */
size_t vd_tess_count_generators(const vd_tess *t);
size_t vd_tess_count_cells(const vd_tess *t);

/* generators(), src/tessellation.rs:519-521: n = count_generators() points to out (3n doubles). */
int vd_tess_get_generators(const vd_tess *t, double *out, size_t n);

/* Per-cell reductions for all n = count_cells() cells, index order:
 *   get_cell(i).volume()       src/cell/cell_3d_faces.rs:106-143
 *   get_cell(i).centroid()     src/cell/cell_3d_faces.rs:145-205 (3 doubles per cell)
 *   get_cell(i).face_counts().len()  src/cell/cell_3d_faces.rs:87-89 (number of faces) */
int vd_tess_get_volumes(const vd_tess *t, double *out, size_t n);
int vd_tess_get_centroids(const vd_tess *t, double *out3, size_t n);
int vd_tess_get_face_counts(const vd_tess *t, uint32_t *out, size_t n);

/* get_cell(i).face_neighbors() of every cell as CSR in face order, src/cell/cell_3d_faces.rs:97-99.
 * offsets has n + 1 entries (n = count_cells()); ids has room for cap entries. With ids == NULL
 * only the offsets are written (size query). */
int vd_tess_get_neighbors(const vd_tess *t, uint64_t *offsets, size_t n, int32_t *ids, size_t cap);

/* get_cell(i), src/tessellation.rs:514-516, as the four arrays of Cell3DFaces
 * (src/cell/cell_3d_faces.rs:23-34, accessors :82-99):
 *   sizes[0] vertices, sizes[1] faces, sizes[2] face index entries;
 *   vertices 3*sizes[0] doubles, face_counts sizes[1], face_indices sizes[2],
 *   face_neighbors sizes[1]. Pointers may be NULL to query sizes only.
 * Returns VD_ERANGE when i >= count_cells(), VD_ESTATE when cells are not kept. */
int vd_tess_get_cell(const vd_tess *t, size_t i, uint64_t sizes[3], double *vertices,
                     uint32_t *face_counts, uint32_t *face_indices, int32_t *face_neighbors);

/* Options (no counterpart in the reference; they change what is kept, never the results). */
#define VD_OPT_KEEP_CELLS 1 /* 1 (default): keep every full cell for get_cell; 0: reductions and
                               neighbour ids only, the output of Tessellation::map */
#define VD_OPT_GRAPHS 2     /* 1 (default): when calculate() repeats the launch sequence of the
                               previous call (same sizes, grids, walls and buffers; the generators
                               may differ), its first round (binning, the fast-tier cell launch,
                               the retry collect) is captured once into a hipGraph and replayed;
                               0: always launch eagerly. Results are identical either way. */
#define VD_OPT_ARENA 3      /* initial full-cell arena per cell, in vertices (default 40; 3/5 as many
                               faces and 3 times as many face indices): a launch that fills it
                               doubles it and runs again, with identical results */
#define VD_OPT_FAR_LIST 4   /* cells that reach across empty space and would walk about this many
                               shell entries or more (default 262144; 0: never) rerun over a list
                               of only the entries whose bins hold generators, in shell order:
                               the same candidates in the same order, so identical results */
#define VD_OPT_ATTEMPTS 5   /* launches, each with twice the arena of the one before, after which
                               calculate() gives up with VD_ECAPACITY and keeps no results
                               (default 8, at most 62) */
/* Launch shape of the cell kernels; every choice gives identical results. */
#define VD_OPT_T0_WAVES 6   /* resident waves per SIMD of the small-LDS tier's kernel: 5, 6, 7 or
                               8 (default; a layout under 5 KiB of LDS per wave), each a separate
                               compiled instantiation */
#define VD_OPT_QUEUE_GRAIN 7 /* cells a wave takes from the work queue per atomic, 1 (default) to 64 */
#define VD_OPT_FIRST_TIER 8 /* tier of the first launch over every cell: 0 (default, 64 vertices)
                               or 1 (128 vertices); cells that outgrow it move up as usual */
#define VD_OPT_FIRST_PROBE 9 /* 1: the first launch uses the instantiation that skips runs of
                               empty shell batches in bulk (cells in sparse regions walk far);
                               0 (default): the plain walk, fewer registers */
int vd_tess_set_option(vd_tess *t, int key, int64_t value);
/* Times calculate() under the launch shapes (coordinate descent over VD_OPT_FIRST_PROBE {0, 1},
 * VD_OPT_T0_WAVES {5, 6, 7, 8}, VD_OPT_QUEUE_GRAIN {1, 2, 4, 8}, VD_OPT_FIRST_TIER {0, 1}, from the
 * current shape), keeps the fastest and writes it to picks[4] (waves, grain, first tier,
 * skip-ahead). Generators must be set; about a dozen calculates run, and the results of the last
 * one are kept (identical under every shape). No counterpart in the reference. */
int vd_tess_autotune_launch(vd_tess *t, int picks[4]);

/* Diagnostics of the last calculate(): [0] cells of the first launch (its tier: VD_OPT_FIRST_TIER), [1] cell runs
 * in the larger tiers, [2] shell-prefix extensions, [3] entries of the shell prefixes,
 * [4] time in microseconds of the first cell launch (of the whole replayed first round when a
 * graph ran), [5] total calculate time (us), [6] vertices and [7] face-index entries written to
 * the cell arena (full-cell mode, else 0), [8] graph replays and [9] graph captures so far. */
int vd_tess_get_stats(const vd_tess *t, double out[10]);

/* Host-only (no device needed): the stable bin shell order of Algorithm3DGrid restricted to the
 * entries whose key is <= ((kmax - 1) * max cell size)^2 (algo_3d_grid.rs:47-95 with a stable
 * sort). Writes up to cap entries (dx, dy, dz triples and keys); *count receives the prefix
 * length, *next_key the key of the first entry after it (+inf when the prefix is the whole
 * table). */
int vd_shell_prefix(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny,
                    uint32_t nz, uint32_t kmax, int32_t *dxyz, double *keys, size_t cap,
                    size_t *count, double *next_key);

/* The same prefix built the way calculate() builds it: on the device (key per offset, stable
 * radix sort by key). No counterpart in the reference; a diagnostic that lets tests compare the
 * device construction with vd_shell_prefix entry by entry. */
int vd_shell_prefix_device(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny,
                           uint32_t nz, uint32_t kmax, int32_t *dxyz, double *keys, size_t cap,
                           size_t *count, double *next_key);

/* ---- Batched tessellations (no counterpart in the reference: many independent
 * Tessellation::<3, Cell3DFaces, Algorithm3DGrid> computed in one persistent launch) ---------- */

typedef struct vd_batch vd_batch;

int vd_batch_create(vd_batch **out);
void vd_batch_destroy(vd_batch *b);
/* Removes every member. */
int vd_batch_clear(vd_batch *b);
/* Adds one tessellation: Tessellation::new(bounds, Algorithm3DGrid::new(nx, ny, nz, &bounds)),
 * its walls (add_wall in order) and set_generators(xyz[0..3n]); *index receives its member
 * number. nx = ny = nz = 0 picks the grid with vd_grid_for at the batch's generators-per-bin
 * target (VD_BOPT_PER_BIN). Points are staged on the host and uploaded by the next
 * vd_batch_upload / vd_batch_calculate. */
int vd_batch_add(vd_batch *b, const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny,
                 uint32_t nz, const vd_wall *walls, int nwalls, const double *xyz, size_t n,
                 uint32_t *index);
/* Uploads the staged points of every member to device memory (done once; calculate reuses it). */
int vd_batch_upload(vd_batch *b);
/* Every member's calculate() in one launch sequence: wall filter and binning of all members, then
 * one persistent cell launch over all their cells. */
int vd_batch_calculate(vd_batch *b);
/* Every member's relax() (src/tessellation.rs:482-495); requires vd_batch_calculate first. */
int vd_batch_relax(vd_batch *b);
size_t vd_batch_count(const vd_batch *b);
/* info[0] cells (generators kept), [1..3] grid nx ny nz, [4] first output slot of the member. */
int vd_batch_member_info(const vd_batch *b, uint32_t t, uint64_t info[5]);
int vd_batch_get_volumes(const vd_batch *b, uint32_t t, double *out, size_t n);
int vd_batch_get_centroids(const vd_batch *b, uint32_t t, double *out3, size_t n);
int vd_batch_get_face_counts(const vd_batch *b, uint32_t t, uint32_t *out, size_t n);
int vd_batch_get_generators(const vd_batch *b, uint32_t t, double *out, size_t n);
int vd_batch_get_neighbors(const vd_batch *b, uint32_t t, uint64_t *offsets, size_t n, int32_t *ids,
                           size_t cap);
int vd_batch_get_cell(const vd_batch *b, uint32_t t, size_t i, uint64_t sizes[3], double *vertices,
                      uint32_t *face_counts, uint32_t *face_indices, int32_t *face_neighbors);
#define VD_BOPT_KEEP_CELLS 1 /* as VD_OPT_KEEP_CELLS; default 0 for batches */
#define VD_BOPT_PER_BIN 2    /* generators-per-bin target of vd_batch_add's grid choice, x1000 (default 350) */
#define VD_BOPT_GRAPHS 3     /* as VD_OPT_GRAPHS (default 1) */
#define VD_BOPT_FAR_LIST 4   /* as VD_OPT_FAR_LIST (default 262144) */
#define VD_BOPT_T0_WAVES 5   /* as VD_OPT_T0_WAVES */
#define VD_BOPT_QUEUE_GRAIN 6 /* as VD_OPT_QUEUE_GRAIN */
#define VD_BOPT_FIRST_TIER 7 /* as VD_OPT_FIRST_TIER */
#define VD_BOPT_FIRST_PROBE 8 /* as VD_OPT_FIRST_PROBE */
int vd_batch_set_option(vd_batch *b, int key, int64_t value);
int vd_batch_get_stats(const vd_batch *b, double out[10]);
/* Times vd_batch_calculate for each candidate generators-per-bin target (grids re-derived with
 * vd_grid_for for the members added with a zero grid), keeps the fastest, and returns it. */
int vd_batch_autotune(vd_batch *b, const double *per_bin, int ncand, double *best);
/* vd_tess_autotune_launch for a batch (its grids as they are: run vd_batch_autotune first). */
int vd_batch_autotune_launch(vd_batch *b, int picks[4]);

/* The grid rule of the autotuner: bins per axis proportional to the box edges with about
 * per_bin generators per bin, at least 1 per axis. */
int vd_grid_for(const double bmin[3], const double bmax[3], size_t n, double per_bin, uint32_t out[3]);

/* The grid rule for uneven input (the server's choice for jobs submitted without a grid): as
 * vd_grid_for, with per_bin divided by r^(2/3), where r is how much denser the average
 * generator's neighbourhood is than the mean (sum of c (c - 1) over a coarse histogram of the
 * points), clamped to [1, 64]; when over a quarter of the coarse bins are empty, per_bin times 3
 * instead (cells next to empty space walk across it). Uniform input gets about the vd_grid_for
 * grid; clustered input with a background a finer one; clusters in empty space a coarser one. */
int vd_grid_for_points(const double bmin[3], const double bmax[3], const double *xyz, size_t n,
                       double per_bin, uint32_t out[3]);

/* ---- Serving (no counterpart in the reference): continuous batching of tessellation jobs and
 * Lloyd relaxation jobs on one device ---------------------------------------------------------- */

typedef struct vd_server vd_server;

/* hbm_budget: device bytes the admitted jobs may hold together; max_slots: jobs in flight. */
int vd_server_create(size_t hbm_budget, int max_slots, vd_server **out);
void vd_server_destroy(vd_server *s);
/* Queues one job: a tessellation as vd_batch_add describes it, run for `iterations` calculate()
 * calls with a relax() between consecutive ones (iterations = 1: a plain tessellation). The points
 * are copied. A job larger than the whole budget is refused with VD_ENOMEM (its id is still
 * returned and reports state 3). */
int vd_server_submit(vd_server *s, const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny,
                     uint32_t nz, const vd_wall *walls, int nwalls, const double *xyz, size_t n,
                     uint32_t iterations, uint64_t *job);
/* One step: admit queued jobs in arrival order while slots and budget allow; collect the large-job
 * lane's step if it has finished and start its next one (one batched calculate over its jobs, on a
 * worker thread); then run one batched calculate over the small-job lane (VD_SOPT_SMALL_JOB). A
 * lane's step finishes the jobs whose last iteration it was and relaxes the others on the device.
 * *finished receives the number of jobs finished. Call it in a loop while jobs are pending or
 * running; with no small-job work it waits at most a few milliseconds for the large lane. */
int vd_server_step(vd_server *s, uint32_t *finished);
/* counts[0] queued jobs, [1] admitted jobs, [2] budget bytes in use. */
int vd_server_pending(const vd_server *s, uint64_t counts[3]);
/* state: 0 queued, 1 running, 2 finished, 3 refused; info: cells of the last calculate,
 * calculate() runs done, cells over all runs; times: submit, admit, finish (seconds, steady clock). */
int vd_server_job_info(const vd_server *s, uint64_t job, int32_t *state, uint64_t info[3], double times[3]);
/* The grid a job runs on (its own, or the one vd_grid_for_points picked for it at submission). */
int vd_server_job_grid(const vd_server *s, uint64_t job, uint32_t out[3]);
/* Results of a finished job's last calculate(): volumes, centroids (3 per cell) and the generators
 * it ran on (3 per cell); NULL pointers are skipped. */
int vd_server_result(const vd_server *s, uint64_t job, double *volumes, double *centroids,
                     double *generators, size_t n);
/* Forgets a finished or refused job. */
int vd_server_release(vd_server *s, uint64_t job);
/* Device memory of the server, counted at every allocation and free it makes (lanes, jobs, spare
 * buffers): [0] bytes held now, [1] the most it has held at once, [2] its fixed reserve (the
 * lanes' global-memory tier slots), [3] the budget. Admission keeps [1] within [3]: each job is
 * charged an upper bound of what it can hold while it runs (its points, its share of a lane's
 * buffers, its grid's bins and its shell table at the largest with the table's construction
 * scratch), a lane that kept buffers from earlier steps counts with what it holds, and a budget
 * at or below the reserve is refused by vd_server_create. */
int vd_server_memory(const vd_server *s, uint64_t out[4]);
/* [0] steps, [1] cells computed, [2] budget bytes in use, [3] budget. */
int vd_server_stats(const vd_server *s, double out[4]);
#define VD_SOPT_SMALL_JOB 2 /* jobs with at most this many points (default 65536; 0: none) run in a
                               lane of their own, stepped by vd_server_step itself, while the steps
                               of larger jobs run on a worker thread and leave 16 CUs to it: small
                               jobs do not wait for large ones. Applies to later submits. */
#define VD_SOPT_PER_BIN 1   /* generators-per-bin target of vd_grid_for_points for jobs submitted
                               without a grid, x1000 (default 350); applies to later submits. A
                               Lloyd job (iterations > 1) whose input has large empty regions gets
                               the plain target instead of three times it: relaxation fills them. */
int vd_server_set_option(vd_server *s, int key, int64_t value);

/* ---- Without device allocation (no counterpart in the reference) --------------------------
 * For callers that cannot allocate device memory: code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc. One tessellation, as vd_tess_create + vd_tess_add_wall (walls in
 * order) + vd_tess_set_generators_device + vd_tess_calculate compute it, with the reduced outputs:
 * every intermediate buffer is carved from a static device array of VD_STATIC_SCRATCH_BYTES that
 * each GPU the library runs on holds until the process ends. No hipMalloc, hipFree, pinned memory,
 * event, stream or graph is created. d_xyz (3n doubles), d_vol (n), d_cen (3n) and d_nface (n) are
 * device memory of the caller; *ncells receives the cells kept by the walls, whose results fill
 * the first *ncells entries. stream: a hipStream_t of the caller (NULL: the null stream); the call
 * returns after its work on it has finished.
 * Limits: at most VD_STATIC_MAX_POINTS points and VD_STATIC_MAX_BINS bins (VD_EINVAL beyond);
 * inputs that need more scratch than the array holds during the computation (a shell prefix grown
 * for a very sparse grid, many reruns) return VD_ENOMEM. The array's contents are left from the
 * previous call and cleared where they must be; calls that overlap on one GPU (from several
 * threads, or on several streams) share the array and race: serialise them. */
#define VD_STATIC_SCRATCH_BYTES (96u << 20)
#define VD_STATIC_MAX_POINTS 262144u
#define VD_STATIC_MAX_BINS 2097152u
int vd_calculate_static(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny, uint32_t nz,
                        const vd_wall *walls, int nwalls, const double *d_xyz, size_t n, double *d_vol,
                        double *d_cen, uint32_t *d_nface, size_t *ncells, void *stream);

/* Message of the calling thread's last failure ("" if none). */
const char *vd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VD_H */
