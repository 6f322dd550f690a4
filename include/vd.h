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
 *     may be used from separate threads and run on their own HIP streams;
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
#define VD_ECAPACITY (-5) /* a cell exceeded every polytope capacity tier */
#define VD_ESTATE (-6)    /* call out of order (e.g. results before calculate) */

#define VD_WALL_ID_MAX (-1000) /* src/wall/mod.rs:8 */

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

/* clear_walls(), src/tessellation.rs:203-205 (generators are kept as they are). */
int vd_tess_clear_walls(vd_tess *t);

/* set_generators(&[f64]) with a host buffer of n points, xyz interleaved.
 *   src/tessellation.rs:43-67 (wall filter in input order) + algo_3d_grid.rs:111-130 (binning). */
int vd_tess_set_generators(vd_tess *t, const double *xyz, size_t n);

/* The same with the points already in device memory (HBM); d_xyz is not retained. */
int vd_tess_set_generators_device(vd_tess *t, const double *d_xyz, size_t n);

/* calculate(), src/tessellation.rs:213-230: every cell, results kept on the device. */
int vd_tess_calculate(vd_tess *t);

/* relax(), src/tessellation.rs:482-495: each generator moves to its cell's centroid (an empty
 * cell keeps its position), then set_generators. Requires calculate() first. */
int vd_tess_relax(vd_tess *t);

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
int vd_tess_set_option(vd_tess *t, int key, int64_t value);

/* Diagnostics of the last calculate(): [0] cells computed in the small-LDS tier, [1] cells
 * retried in the large tier, [2] shell-prefix extensions, [3] entries of the shell prefix,
 * [4] kernel time in microseconds of the first cell launch, [5] total calculate time (us). */
int vd_tess_get_stats(const vd_tess *t, double out[8]);

/* Host-only (no device needed): the stable bin shell order of Algorithm3DGrid restricted to the
 * entries whose key is <= ((kmax - 1) * max cell size)^2 (algo_3d_grid.rs:47-95 with a stable
 * sort). Writes up to cap entries (dx, dy, dz triples and keys); *count receives the prefix
 * length, *next_key the key of the first entry after it (+inf when the prefix is the whole
 * table). */
int vd_shell_prefix(const double bmin[3], const double bmax[3], uint32_t nx, uint32_t ny,
                    uint32_t nz, uint32_t kmax, int32_t *dxyz, double *keys, size_t cap,
                    size_t *count, double *next_key);

/* Message of the calling thread's last failure ("" if none). */
const char *vd_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* VD_H */
