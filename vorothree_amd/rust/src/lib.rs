//! # voronoid-hip
//!
//! `Tessellation::<3, Cell3DFaces, Algorithm3DGrid>` of the `voronoid` crate (mdt-re/vorothree)
//! computed on an AMD MI355X by libvoronoid_hip.so. The calls keep the reference's names,
//! arguments and results (src/tessellation.rs:24-530); a program switches by importing
//! `Tessellation`, `Cell3DFaces` and `Algorithm3DGrid` from this crate instead of `voronoid`.
//! `BoundingBox`, `Wall`, `WallGeometry` and every wall geometry are the reference's own types,
//! re-exported.
//!
//! Walls: `add_wall` takes any `Wall<3>`. The library prunes the generators with the wall's
//! `contains` and clips every cell with the planes of its `cut`, both evaluated here on the host
//! through the wall's public methods (src/wall/mod.rs:34-48) at every `set_generators`, `add_wall`
//! and `relax` (and for the moved generator at `set_generator`); `is_planar` decides what
//! `calculate_sealed` and `prune_boundaries` treat as curved. `add_wall_native` takes the
//! built-in geometries in the device's own form instead (`sys::vd_wall`), which skips the host
//! round trip.
//!
//! Differences from the reference, all documented in include/vd.h:
//! - cells are computed on the device; `get_cell` copies one cell back, `volumes`,
//!   `centroids` and `face_counts_all` read every cell's reductions at once;
//! - after `set_generators`, `add_wall`, `clear_walls` or `relax`, `count_cells` is 0 until the
//!   next `calculate` (the reference keeps the previous cells until then);
//! - device failures, which the reference has no counterpart for, panic with the library's
//!   message, as the reference panics on an invalid wall id.

pub mod sys;

use std::ffi::CStr;
use std::marker::PhantomData;
use std::os::raw::{c_int, c_void};
use std::panic::{AssertUnwindSafe, catch_unwind};

pub use voronoid::{BoundingBox, WALL_ID_MAX, Wall, WallGeometry, box_side, wall_3d};

fn last_error() -> String {
    unsafe {
        let p = sys::vd_last_error();
        if p.is_null() { String::new() } else { CStr::from_ptr(p).to_string_lossy().into_owned() }
    }
}

/// Panics with the library's message unless rc is VD_OK (the reference has no error path here).
fn check(rc: c_int, what: &str) {
    if rc != sys::VD_OK {
        panic!("{what}: vd error {rc}: {}", last_error());
    }
}

/// The grid of `voronoid::Algorithm3DGrid::new(nx, ny, nz, &bounds)` (src/algorithm/algo_3d_grid.rs:47-95).
/// The library builds the bin shell order on the device as far as the walks need it, so the
/// constructor only records the resolution.
pub struct Algorithm3DGrid {
    pub grid_res_x: usize,
    pub grid_res_y: usize,
    pub grid_res_z: usize,
}

impl Algorithm3DGrid {
    pub fn new(nx: usize, ny: usize, nz: usize, _bounds: &BoundingBox<3>) -> Self {
        Self { grid_res_x: nx, grid_res_y: ny, grid_res_z: nz }
    }
}

/// A computed cell: the four arrays of `voronoid::Cell3DFaces` (src/cell/cell_3d_faces.rs:23-34)
/// and its accessors (:77-99); `volume` and `centroid` restate :106-205 (fan triangulation from
/// each face's first vertex, left folds, the |6V| < 1e-9 guard), so they return the values the
/// library computed on the device, bit for bit.
#[derive(Clone, Debug)]
pub struct Cell3DFaces {
    id: usize,
    vertices: Vec<f64>,
    face_counts: Vec<u32>,
    face_indices: Vec<u32>,
    face_neighbors: Vec<i32>,
}

impl Cell3DFaces {
    /// The ID of the generator associated with this cell.
    pub fn id(&self) -> usize {
        self.id
    }

    /// Flat array of vertices [x, y, z, x, y, z, ...].
    pub fn vertices(&self) -> Vec<f64> {
        self.vertices.clone()
    }

    /// Number of vertices for each face.
    pub fn face_counts(&self) -> Vec<u32> {
        self.face_counts.clone()
    }

    /// Flattened indices for all faces.
    pub fn face_indices(&self) -> Vec<u32> {
        self.face_indices.clone()
    }

    /// Neighbor ID for each face. Negative values indicate walls/boundaries.
    pub fn face_neighbors(&self) -> Vec<i32> {
        self.face_neighbors.clone()
    }

    /// The neighbour ids (Cell::neighbors, cell_3d_faces.rs:532-534).
    pub fn neighbors(&self) -> &[i32] {
        &self.face_neighbors
    }

    /// A cell with no vertices (cell_3d_faces.rs:528-530).
    pub fn is_empty(&self) -> bool {
        self.vertices.is_empty()
    }

    fn vertex(&self, i: u32) -> [f64; 3] {
        let k = 3 * i as usize;
        [self.vertices[k], self.vertices[k + 1], self.vertices[k + 2]]
    }

    /// Volume from the fan triangulation of every face, origin-relative (cell_3d_faces.rs:106-143).
    pub fn volume(&self) -> f64 {
        let mut volume: f64 = 0.0;
        let mut off: usize = 0;
        for &count in &self.face_counts {
            let count = count as usize;
            if count >= 3 {
                let a = self.vertex(self.face_indices[off]);
                for i in 1..count - 1 {
                    let b = self.vertex(self.face_indices[off + i]);
                    let c = self.vertex(self.face_indices[off + i + 1]);
                    volume += a[0] * (b[1] * c[2] - b[2] * c[1])
                        + a[1] * (b[2] * c[0] - b[0] * c[2])
                        + a[2] * (b[0] * c[1] - b[1] * c[0]);
                }
            }
            off += count;
        }
        (volume / 6.0).abs()
    }

    /// Area of face `face_index` (cell_3d_faces.rs:207-254): half the cross-product norm of each
    /// fan triangle from the face's first vertex, summed in fan order; 0.0 past the last face or
    /// for a face of fewer than 3 vertices.
    pub fn face_area(&self, face_index: usize) -> f64 {
        if face_index >= self.face_counts.len() {
            return 0.0;
        }
        let off: usize = self.face_counts[..face_index].iter().map(|&c| c as usize).sum();
        let count = self.face_counts[face_index] as usize;
        if count < 3 {
            return 0.0;
        }
        let p0 = self.vertex(self.face_indices[off]);
        let mut area = 0.0;
        for i in 1..count - 1 {
            let p1 = self.vertex(self.face_indices[off + i]);
            let p2 = self.vertex(self.face_indices[off + i + 1]);
            let u = [p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]];
            let w = [p2[0] - p0[0], p2[1] - p0[1], p2[2] - p0[2]];
            let n = [u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]];
            area += 0.5 * (n[0] * n[0] + n[1] * n[1] + n[2] * n[2]).sqrt();
        }
        area
    }

    /// The vertex indices of every face, in face order (cell_3d_faces.rs:256-269).
    pub fn faces(&self) -> Vec<Vec<usize>> {
        let mut out = Vec::with_capacity(self.face_counts.len());
        let mut off: usize = 0;
        for &count in &self.face_counts {
            let end = off + count as usize;
            out.push(self.face_indices[off..end].iter().map(|&i| i as usize).collect());
            off = end;
        }
        out
    }

    /// Centroid from the same triangulation (cell_3d_faces.rs:145-205): [0, 0, 0] when the sum of
    /// the determinants (6V) is below 1e-9 in magnitude.
    pub fn centroid(&self) -> [f64; 3] {
        let (mut cx, mut cy, mut cz, mut total) = (0.0f64, 0.0f64, 0.0f64, 0.0f64);
        let mut off: usize = 0;
        for &count in &self.face_counts {
            let count = count as usize;
            if count >= 3 {
                let a = self.vertex(self.face_indices[off]);
                for i in 1..count - 1 {
                    let b = self.vertex(self.face_indices[off + i]);
                    let c = self.vertex(self.face_indices[off + i + 1]);
                    let x = b[1] * c[2] - b[2] * c[1];
                    let y = b[2] * c[0] - b[0] * c[2];
                    let z = b[0] * c[1] - b[1] * c[0];
                    let det = a[0] * x + a[1] * y + a[2] * z;
                    total += det;
                    cx += det * (a[0] + b[0] + c[0]);
                    cy += det * (a[1] + b[1] + c[1]);
                    cz += det * (a[2] + b[2] + c[2]);
                }
            }
            off += count;
        }
        if total.abs() < 1e-9 {
            return [0.0, 0.0, 0.0];
        }
        let f = 1.0 / (4.0 * total);
        [cx * f, cy * f, cz * f]
    }
}

// Callbacks of vd_wall_host over a reference Wall<3> (src/wall/mod.rs:34-48). A panic inside the
// geometry must not unwind into the library: it counts as "outside" / "no plane".
unsafe extern "C" fn wall_contains(ctx: *mut c_void, p: *const f64) -> c_int {
    let wall = unsafe { &*(ctx as *const Wall<3>) };
    let p = unsafe { &*(p as *const [f64; 3]) };
    catch_unwind(AssertUnwindSafe(|| wall.contains(p) as c_int)).unwrap_or(0)
}

unsafe extern "C" fn wall_cut(ctx: *mut c_void, g: *const f64, planes: *mut f64, cap: c_int) -> c_int {
    let wall = unsafe { &*(ctx as *const Wall<3>) };
    let g = unsafe { &*(g as *const [f64; 3]) };
    catch_unwind(AssertUnwindSafe(|| {
        let mut n: c_int = 0;
        wall.cut(g, &mut |point: [f64; 3], normal: [f64; 3]| {
            if n < cap {
                let o = unsafe { std::slice::from_raw_parts_mut(planes.add(6 * n as usize), 6) };
                o[..3].copy_from_slice(&point);
                o[3..].copy_from_slice(&normal);
            }
            n += 1;
        });
        n
    }))
    .unwrap_or(0)
}

/// `voronoid::Tessellation` for the specialisation the library implements,
/// `Tessellation<3, Cell3DFaces, Algorithm3DGrid>` (src/tessellation.rs:13-22).
pub struct Tessellation<const D: usize, C, A> {
    raw: *mut sys::vd_tess,
    bounds: BoundingBox<D>,
    // boxed, so the callbacks' context pointers stay where the library was given them
    walls: Vec<Box<Wall<D>>>,
    _types: PhantomData<(C, A)>,
}

// One handle is used from one thread at a time (&mut self for every change), like the reference.
unsafe impl<const D: usize, C, A> Send for Tessellation<D, C, A> {}

impl Tessellation<3, Cell3DFaces, Algorithm3DGrid> {
    /// Tessellation::new (src/tessellation.rs:25-36).
    pub fn new(bounds: BoundingBox<3>, algorithm: Algorithm3DGrid) -> Self {
        let mut raw: *mut sys::vd_tess = std::ptr::null_mut();
        let rc = unsafe {
            sys::vd_tess_create(bounds.min.as_ptr(), bounds.max.as_ptr(), algorithm.grid_res_x as u32,
                                algorithm.grid_res_y as u32, algorithm.grid_res_z as u32, &mut raw)
        };
        check(rc, "Tessellation::new");
        Self { raw, bounds, walls: Vec::new(), _types: PhantomData }
    }

    /// set_generators (src/tessellation.rs:43-67): keeps the points every wall contains.
    pub fn set_generators(&mut self, generators: &[f64]) {
        let rc = unsafe { sys::vd_tess_set_generators(self.raw, generators.as_ptr(), generators.len() / 3) };
        check(rc, "set_generators");
    }

    /// random_generators (src/tessellation.rs:96-122): `count` points uniform in the bounding box
    /// that every wall contains. Seeded per call from the standard library's per-process random
    /// hasher keys, as the reference seeds from thread_rng; `random_generators_seeded` repeats.
    pub fn random_generators(&mut self, count: usize) {
        use std::hash::{BuildHasher, Hasher};
        let mut h = std::collections::hash_map::RandomState::new().build_hasher();
        h.write_usize(count);
        self.random_generators_seeded(count, h.finish());
    }

    /// random_generators with the candidates drawn from the SplitMix64 stream of `seed`.
    pub fn random_generators_seeded(&mut self, count: usize, seed: u64) {
        let rc = unsafe { sys::vd_tess_random_generators(self.raw, count, seed) };
        check(rc, "random_generators");
    }

    /// set_generator (src/tessellation.rs:71-92): moves one generator when every wall contains the
    /// new position; the grid bins follow update_generator (src/algorithm/algo_3d_grid.rs:132-143,
    /// swap_remove from the old bin, push to the new one). An index past the generators is ignored,
    /// as in the reference.
    pub fn set_generator(&mut self, index: usize, generator: &[f64; 3]) {
        let rc = unsafe { sys::vd_tess_set_generator(self.raw, index, generator.as_ptr()) };
        if rc != sys::VD_ERANGE {
            check(rc, "set_generator");
        }
    }

    /// read_generators (src/tessellation.rs:153-171): `id x y z` lines, the id ignored, missing or
    /// unparsable coordinates 0.0.
    pub fn read_generators(&mut self, input: &str) {
        let mut raw = Vec::new();
        for line in input.lines() {
            if line.trim().is_empty() {
                continue;
            }
            let mut parts = line.split_whitespace();
            if parts.next().is_none() {
                continue;
            }
            for _ in 0..3 {
                raw.push(parts.next().and_then(|s| s.parse::<f64>().ok()).unwrap_or(0.0));
            }
        }
        self.set_generators(&raw);
    }

    /// import_generators (src/tessellation.rs:126-150).
    pub fn import_generators<P: AsRef<std::path::Path>>(&mut self, path: P) -> std::io::Result<()> {
        let text = std::fs::read_to_string(path)?;
        self.read_generators(&text);
        Ok(())
    }

    /// add_wall (src/tessellation.rs:197-200) for any wall: contains and cut are evaluated on the
    /// host through the wall's public methods.
    pub fn add_wall(&mut self, wall: Wall<3>) {
        let boxed = Box::new(wall);
        let host = sys::vd_wall_host {
            ctx: &*boxed as *const Wall<3> as *mut c_void,
            contains: Some(wall_contains),
            cut: Some(wall_cut),
            planar: boxed.is_planar() as c_int,
        };
        let rc = unsafe { sys::vd_tess_add_wall_planes_per_generator(self.raw, boxed.id(), &host) };
        if rc == sys::VD_EINVAL && boxed.id() > WALL_ID_MAX {
            panic!("Wall ID must be <= {}", WALL_ID_MAX);   // src/wall/mod.rs:25-27
        }
        check(rc, "add_wall");
        self.walls.push(boxed);
    }

    /// add_wall for a built-in geometry in the device's form (evaluated on the device).
    pub fn add_wall_native(&mut self, wall: &sys::vd_wall) {
        check(unsafe { sys::vd_tess_add_wall(self.raw, wall) }, "add_wall_native");
    }

    /// clear_walls (src/tessellation.rs:203-205).
    pub fn clear_walls(&mut self) {
        check(unsafe { sys::vd_tess_clear_walls(self.raw) }, "clear_walls");
        self.walls.clear();
    }

    /// calculate (src/tessellation.rs:213-230).
    pub fn calculate(&mut self) {
        check(unsafe { sys::vd_tess_calculate(self.raw) }, "calculate");
    }

    /// calculate_sealed (src/tessellation.rs:234-318).
    pub fn calculate_sealed(&mut self) {
        check(unsafe { sys::vd_tess_calculate_sealed(self.raw) }, "calculate_sealed");
    }

    /// prune_boundaries (src/tessellation.rs:323-385).
    pub fn prune_boundaries(&mut self) {
        check(unsafe { sys::vd_tess_prune_boundaries(self.raw) }, "prune_boundaries");
    }

    /// relax (src/tessellation.rs:482-495).
    pub fn relax(&mut self) {
        check(unsafe { sys::vd_tess_relax(self.raw) }, "relax");
    }

    /// A library option (`sys::VD_OPT_*`; no counterpart in the reference). Options change what is
    /// kept or how the kernels are launched, never the results.
    pub fn set_option(&mut self, key: c_int, value: i64) {
        check(unsafe { sys::vd_tess_set_option(self.raw, key, value) }, "set_option");
    }

    /// Times calculate() under the launch shapes and keeps the fastest (vd_tess_autotune_launch):
    /// [tier-0 waves per SIMD, cells per queue atomic, first tier, first-launch skip-ahead].
    /// Generators must be set; the results of the last calculate are kept.
    pub fn autotune_launch(&mut self) -> [i32; 4] {
        let mut picks = [0 as c_int; 4];
        check(unsafe { sys::vd_tess_autotune_launch(self.raw, picks.as_mut_ptr()) }, "autotune_launch");
        picks
    }

    /// map (src/tessellation.rs:391-412): computes every cell and applies f to each. The cells are
    /// computed by one calculate on the device, whose results the handle keeps.
    pub fn map<F, T>(&self, f: F) -> Vec<T>
    where
        F: Fn(Cell3DFaces) -> T + Sync + Send,
        T: Send,
    {
        check(unsafe { sys::vd_tess_calculate(self.raw) }, "map");
        (0..self.count_cells()).filter_map(|i| self.get_cell(i)).map(f).collect()
    }

    pub fn count_generators(&self) -> usize {
        unsafe { sys::vd_tess_count_generators(self.raw) }
    }

    pub fn count_cells(&self) -> usize {
        unsafe { sys::vd_tess_count_cells(self.raw) }
    }

    pub fn generators(&self) -> Vec<f64> {
        let n = self.count_generators();
        let mut out = vec![0.0; 3 * n];
        check(unsafe { sys::vd_tess_get_generators(self.raw, out.as_mut_ptr(), n) }, "generators");
        out
    }

    pub fn get_generator(&self, index: usize) -> [f64; 3] {
        let g = self.generators();
        [g[3 * index], g[3 * index + 1], g[3 * index + 2]]
    }

    /// get_cell (src/tessellation.rs:514-516): None past the last cell.
    pub fn get_cell(&self, index: usize) -> Option<Cell3DFaces> {
        let mut s = [0u64; 3];
        let null = std::ptr::null_mut();
        let rc = unsafe { sys::vd_tess_get_cell(self.raw, index, s.as_mut_ptr(), null, null as *mut u32,
                                                null as *mut u32, null as *mut i32) };
        if rc == sys::VD_ERANGE {
            return None;
        }
        check(rc, "get_cell");
        let (nv, nf, ni) = (s[0] as usize, s[1] as usize, s[2] as usize);
        let mut c = Cell3DFaces { id: index, vertices: vec![0.0; 3 * nv], face_counts: vec![0; nf],
                                  face_indices: vec![0; ni], face_neighbors: vec![0; nf] };
        let rc = unsafe {
            sys::vd_tess_get_cell(self.raw, index, s.as_mut_ptr(), c.vertices.as_mut_ptr(), c.face_counts.as_mut_ptr(),
                                  c.face_indices.as_mut_ptr(), c.face_neighbors.as_mut_ptr())
        };
        check(rc, "get_cell");
        Some(c)
    }

    pub fn cells(&self) -> Vec<Cell3DFaces> {
        (0..self.count_cells()).filter_map(|i| self.get_cell(i)).collect()
    }

    /// Every cell's volume (get_cell(i).volume(), bit for bit), read at once.
    pub fn volumes(&self) -> Vec<f64> {
        let mut v = vec![0.0; self.count_cells()];
        check(unsafe { sys::vd_tess_get_volumes(self.raw, v.as_mut_ptr(), v.len()) }, "volumes");
        v
    }

    /// Every cell's centroid (get_cell(i).centroid()), read at once.
    pub fn centroids(&self) -> Vec<[f64; 3]> {
        let n = self.count_cells();
        let mut flat = vec![0.0; 3 * n];
        check(unsafe { sys::vd_tess_get_centroids(self.raw, flat.as_mut_ptr(), n) }, "centroids");
        flat.chunks_exact(3).map(|c| [c[0], c[1], c[2]]).collect()
    }

    /// Every cell's number of faces (get_cell(i).face_counts().len()), read at once.
    pub fn face_counts_all(&self) -> Vec<u32> {
        let mut v = vec![0u32; self.count_cells()];
        check(unsafe { sys::vd_tess_get_face_counts(self.raw, v.as_mut_ptr(), v.len()) }, "face_counts_all");
        v
    }

    /// The (cell, neighbour, wall id) triples of the last calculate_sealed.
    pub fn seal_log(&self) -> Vec<i32> {
        let mut n = 0usize;
        check(unsafe { sys::vd_tess_get_seal_log(self.raw, std::ptr::null_mut(), 0, &mut n) }, "seal_log");
        let mut out = vec![0i32; n];
        check(unsafe { sys::vd_tess_get_seal_log(self.raw, out.as_mut_ptr(), n, &mut n) }, "seal_log");
        out
    }

    pub fn bounds(&self) -> &BoundingBox<3> {
        &self.bounds
    }
}

impl<const D: usize, C, A> Drop for Tessellation<D, C, A> {
    fn drop(&mut self) {
        unsafe { sys::vd_tess_destroy(self.raw) };
    }
}
