//! Raw declarations of include/vd.h (the C ABI of libvoronoid_hip.so) used by this crate.
#![allow(non_camel_case_types)]

use std::os::raw::{c_char, c_int, c_void};

pub const VD_OK: c_int = 0;
pub const VD_EINVAL: c_int = -1;
pub const VD_ERANGE: c_int = -2;
pub const VD_ENOMEM: c_int = -3;
pub const VD_EDEVICE: c_int = -4;
pub const VD_ECAPACITY: c_int = -5;
pub const VD_ESTATE: c_int = -6;

pub const VD_OPT_KEEP_CELLS: c_int = 1;
pub const VD_OPT_GRAPHS: c_int = 2;
pub const VD_OPT_ARENA: c_int = 3;
pub const VD_OPT_FAR_LIST: c_int = 4;
pub const VD_OPT_ATTEMPTS: c_int = 5;
pub const VD_OPT_T0_WAVES: c_int = 6;
pub const VD_OPT_QUEUE_GRAIN: c_int = 7;
pub const VD_OPT_FIRST_TIER: c_int = 8;
pub const VD_OPT_FIRST_PROBE: c_int = 9;

pub const VD_WALL_PLANE: i32 = 0;
pub const VD_WALL_SPHERE: i32 = 1;
pub const VD_WALL_CYLINDER: i32 = 2;
pub const VD_WALL_CONE: i32 = 3;
pub const VD_WALL_TORUS: i32 = 4;
pub const VD_WALL_POLYHEDRON: i32 = 5;
pub const VD_WALL_TUBE: i32 = 6;

#[repr(C)]
pub struct vd_tess {
    _opaque: [u8; 0],
}

/// struct vd_wall: a built-in geometry evaluated on the device.
#[repr(C)]
#[derive(Clone, Copy)]
pub struct vd_wall {
    pub id: i32,
    pub kind: i32,
    pub a: [f64; 3],
    pub b: [f64; 3],
    pub r: f64,
    pub r2: f64,
    pub pts: *const f64,
    pub npts: u32,
    pub closed: i32,
}

pub type vd_contains_fn = unsafe extern "C" fn(ctx: *mut c_void, p: *const f64) -> c_int;
pub type vd_cut_fn = unsafe extern "C" fn(ctx: *mut c_void, g: *const f64, planes: *mut f64, cap: c_int) -> c_int;

/// struct vd_wall_host: a geometry evaluated on the host through callbacks.
#[repr(C)]
pub struct vd_wall_host {
    pub ctx: *mut c_void,
    pub contains: Option<vd_contains_fn>,
    pub cut: Option<vd_cut_fn>,
    pub planar: c_int,
}

unsafe extern "C" {
    pub fn vd_tess_create(bmin: *const f64, bmax: *const f64, nx: u32, ny: u32, nz: u32,
                          out: *mut *mut vd_tess) -> c_int;
    pub fn vd_tess_destroy(t: *mut vd_tess);
    pub fn vd_tess_add_wall_plane(t: *mut vd_tess, id: i32, point: *const f64, normal: *const f64) -> c_int;
    pub fn vd_tess_add_wall_sphere(t: *mut vd_tess, id: i32, center: *const f64, radius: f64) -> c_int;
    pub fn vd_tess_add_wall(t: *mut vd_tess, wall: *const vd_wall) -> c_int;
    pub fn vd_tess_add_wall_planes_per_generator(t: *mut vd_tess, id: i32, wall: *const vd_wall_host) -> c_int;
    pub fn vd_tess_clear_walls(t: *mut vd_tess) -> c_int;
    pub fn vd_tess_set_generators(t: *mut vd_tess, xyz: *const f64, n: usize) -> c_int;
    pub fn vd_tess_set_generators_device(t: *mut vd_tess, d_xyz: *const f64, n: usize) -> c_int;
    pub fn vd_tess_set_generator(t: *mut vd_tess, index: usize, xyz: *const f64) -> c_int;
    pub fn vd_tess_random_generators(t: *mut vd_tess, count: usize, seed: u64) -> c_int;
    pub fn vd_tess_calculate(t: *mut vd_tess) -> c_int;
    pub fn vd_tess_relax(t: *mut vd_tess) -> c_int;
    pub fn vd_tess_calculate_sealed(t: *mut vd_tess) -> c_int;
    pub fn vd_tess_prune_boundaries(t: *mut vd_tess) -> c_int;
    pub fn vd_tess_get_seal_log(t: *const vd_tess, out: *mut i32, cap: usize, count: *mut usize) -> c_int;
    pub fn vd_tess_get_prune_log(t: *const vd_tess, ints: *mut i32, icap: usize, pos: *mut f64, pcap: usize,
                                 counts: *mut usize) -> c_int;
    pub fn vd_tess_count_generators(t: *const vd_tess) -> usize;
    pub fn vd_tess_count_cells(t: *const vd_tess) -> usize;
    pub fn vd_tess_get_generators(t: *const vd_tess, out: *mut f64, n: usize) -> c_int;
    pub fn vd_tess_get_volumes(t: *const vd_tess, out: *mut f64, n: usize) -> c_int;
    pub fn vd_tess_get_centroids(t: *const vd_tess, out3: *mut f64, n: usize) -> c_int;
    pub fn vd_tess_get_face_counts(t: *const vd_tess, out: *mut u32, n: usize) -> c_int;
    pub fn vd_tess_get_neighbors(t: *const vd_tess, offsets: *mut u64, n: usize, ids: *mut i32, cap: usize) -> c_int;
    pub fn vd_tess_get_cell(t: *const vd_tess, i: usize, sizes: *mut u64, vertices: *mut f64, face_counts: *mut u32,
                            face_indices: *mut u32, face_neighbors: *mut i32) -> c_int;
    pub fn vd_tess_set_option(t: *mut vd_tess, key: c_int, value: i64) -> c_int;
    pub fn vd_tess_autotune_launch(t: *mut vd_tess, picks: *mut c_int) -> c_int;
    pub fn vd_tess_get_stats(t: *const vd_tess, out: *mut f64) -> c_int;
    pub fn vd_last_error() -> *const c_char;
}
