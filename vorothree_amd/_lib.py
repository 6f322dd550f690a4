"""ctypes binding of libvoronoid_hip.so (the C ABI declared in include/vd.h).

The library is the only compute path: if it is missing or cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvoronoid_hip.so")

VD_OK = 0
VD_EINVAL = -1
VD_ERANGE = -2
VD_ENOMEM = -3
VD_EDEVICE = -4
VD_ECAPACITY = -5
VD_ESTATE = -6
VD_OPT_KEEP_CELLS = 1
VD_OPT_GRAPHS = 2
VD_OPT_ARENA = 3
VD_OPT_FAR_LIST = 4
VD_OPT_ATTEMPTS = 5
VD_OPT_T0_WAVES = 6
VD_OPT_QUEUE_GRAIN = 7
VD_OPT_FIRST_TIER = 8
VD_OPT_FIRST_PROBE = 9
VD_STATIC_SCRATCH_BYTES = 96 << 20
VD_STATIC_MAX_POINTS = 262144
VD_STATIC_MAX_BINS = 2097152
VD_BOPT_KEEP_CELLS = 1
VD_BOPT_PER_BIN = 2
VD_BOPT_GRAPHS = 3
VD_BOPT_FAR_LIST = 4
VD_BOPT_T0_WAVES = 5
VD_BOPT_QUEUE_GRAIN = 6
VD_BOPT_FIRST_TIER = 7
VD_BOPT_FIRST_PROBE = 8
VD_SOPT_PER_BIN = 1
VD_SOPT_SMALL_JOB = 2


class VdWall(ctypes.Structure):
    """struct vd_wall of include/vd.h."""
    _fields_ = [("id", ctypes.c_int32), ("kind", ctypes.c_int32), ("a", ctypes.c_double * 3),
                ("b", ctypes.c_double * 3), ("r", ctypes.c_double), ("r2", ctypes.c_double),
                ("pts", ctypes.c_void_p), ("npts", ctypes.c_uint32), ("closed", ctypes.c_int32)]


# struct vd_wall_host of include/vd.h: a wall whose geometry is evaluated by the caller
HOST_CONTAINS = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double))
HOST_CUT = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double), ctypes.c_int)


class VdWallHost(ctypes.Structure):
    _fields_ = [("ctx", ctypes.c_void_p), ("contains", HOST_CONTAINS), ("cut", HOST_CUT),
                ("planar", ctypes.c_int)]


VD_WALL_PLANE, VD_WALL_SPHERE, VD_WALL_CYLINDER, VD_WALL_CONE, VD_WALL_TORUS = 0, 1, 2, 3, 4
VD_WALL_POLYHEDRON, VD_WALL_TUBE = 5, 6

_vp = ctypes.c_void_p
_sz = ctypes.c_size_t
_u32 = ctypes.c_uint32
_i32 = ctypes.c_int32

# name -> (restype, argtypes); the list of exported entry points of include/vd.h
SIGNATURES = {
    "vd_tess_create": (ctypes.c_int, [_vp, _vp, _u32, _u32, _u32, ctypes.POINTER(_vp)]),
    "vd_tess_destroy": (None, [_vp]),
    "vd_tess_add_wall_plane": (ctypes.c_int, [_vp, _i32, _vp, _vp]),
    "vd_tess_add_wall_sphere": (ctypes.c_int, [_vp, _i32, _vp, ctypes.c_double]),
    "vd_tess_add_wall": (ctypes.c_int, [_vp, ctypes.POINTER(VdWall)]),
    "vd_tess_clear_walls": (ctypes.c_int, [_vp]),
    "vd_tess_set_generators": (ctypes.c_int, [_vp, _vp, _sz]),
    "vd_tess_set_generators_device": (ctypes.c_int, [_vp, _vp, _sz]),
    "vd_tess_set_generator": (ctypes.c_int, [_vp, _sz, _vp]),
    "vd_tess_random_generators": (ctypes.c_int, [_vp, _sz, ctypes.c_uint64]),
    "vd_tess_calculate": (ctypes.c_int, [_vp]),
    "vd_tess_relax": (ctypes.c_int, [_vp]),
    "vd_tess_calculate_sealed": (ctypes.c_int, [_vp]),
    "vd_tess_prune_boundaries": (ctypes.c_int, [_vp]),
    "vd_tess_get_seal_log": (ctypes.c_int, [_vp, _vp, _sz, _vp]),
    "vd_tess_get_prune_log": (ctypes.c_int, [_vp, _vp, _sz, _vp, _sz, _vp]),
    "vd_tess_count_generators": (_sz, [_vp]),
    "vd_tess_count_cells": (_sz, [_vp]),
    "vd_tess_get_generators": (ctypes.c_int, [_vp, _vp, _sz]),
    "vd_tess_get_volumes": (ctypes.c_int, [_vp, _vp, _sz]),
    "vd_tess_get_centroids": (ctypes.c_int, [_vp, _vp, _sz]),
    "vd_tess_get_face_counts": (ctypes.c_int, [_vp, _vp, _sz]),
    "vd_tess_get_neighbors": (ctypes.c_int, [_vp, _vp, _sz, _vp, _sz]),
    "vd_tess_get_cell": (ctypes.c_int, [_vp, _sz, _vp, _vp, _vp, _vp, _vp]),
    "vd_tess_set_option": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64]),
    "vd_tess_autotune_launch": (ctypes.c_int, [_vp, _vp]),
    "vd_tess_get_stats": (ctypes.c_int, [_vp, _vp]),
    "vd_shell_prefix": (ctypes.c_int, [_vp, _vp, _u32, _u32, _u32, _u32, _vp, _vp, _sz, _vp, _vp]),
    "vd_shell_prefix_device": (ctypes.c_int, [_vp, _vp, _u32, _u32, _u32, _u32, _vp, _vp, _sz, _vp, _vp]),
    "vd_batch_create": (ctypes.c_int, [ctypes.POINTER(_vp)]),
    "vd_batch_destroy": (None, [_vp]),
    "vd_batch_clear": (ctypes.c_int, [_vp]),
    "vd_batch_add": (ctypes.c_int, [_vp, _vp, _vp, _u32, _u32, _u32, _vp, ctypes.c_int, _vp, _sz, _vp]),
    "vd_batch_upload": (ctypes.c_int, [_vp]),
    "vd_batch_calculate": (ctypes.c_int, [_vp]),
    "vd_batch_relax": (ctypes.c_int, [_vp]),
    "vd_batch_count": (_sz, [_vp]),
    "vd_batch_member_info": (ctypes.c_int, [_vp, _u32, _vp]),
    "vd_batch_get_volumes": (ctypes.c_int, [_vp, _u32, _vp, _sz]),
    "vd_batch_get_centroids": (ctypes.c_int, [_vp, _u32, _vp, _sz]),
    "vd_batch_get_face_counts": (ctypes.c_int, [_vp, _u32, _vp, _sz]),
    "vd_batch_get_generators": (ctypes.c_int, [_vp, _u32, _vp, _sz]),
    "vd_batch_get_neighbors": (ctypes.c_int, [_vp, _u32, _vp, _sz, _vp, _sz]),
    "vd_batch_get_cell": (ctypes.c_int, [_vp, _u32, _sz, _vp, _vp, _vp, _vp, _vp]),
    "vd_batch_set_option": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64]),
    "vd_batch_get_stats": (ctypes.c_int, [_vp, _vp]),
    "vd_batch_autotune": (ctypes.c_int, [_vp, _vp, ctypes.c_int, _vp]),
    "vd_batch_autotune_launch": (ctypes.c_int, [_vp, _vp]),
    "vd_grid_for": (ctypes.c_int, [_vp, _vp, _sz, ctypes.c_double, _vp]),
    "vd_tess_add_wall_planes_per_generator": (ctypes.c_int, [_vp, _i32, ctypes.POINTER(VdWallHost)]),
    "vd_calculate_static": (ctypes.c_int, [_vp, _vp, _u32, _u32, _u32, _vp, ctypes.c_int, _vp, _sz, _vp, _vp,
                                           _vp, ctypes.POINTER(_sz), _vp]),
    "vd_grid_for_points": (ctypes.c_int, [_vp, _vp, _vp, _sz, ctypes.c_double, _vp]),
    "vd_server_create": (ctypes.c_int, [_sz, ctypes.c_int, ctypes.POINTER(_vp)]),
    "vd_server_destroy": (None, [_vp]),
    "vd_server_submit": (ctypes.c_int, [_vp, _vp, _vp, _u32, _u32, _u32, _vp, ctypes.c_int, _vp, _sz, _u32, _vp]),
    "vd_server_step": (ctypes.c_int, [_vp, _vp]),
    "vd_server_pending": (ctypes.c_int, [_vp, _vp]),
    "vd_server_job_info": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp, _vp, _vp]),
    "vd_server_result": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp, _vp, _vp, _sz]),
    "vd_server_release": (ctypes.c_int, [_vp, ctypes.c_uint64]),
    "vd_server_stats": (ctypes.c_int, [_vp, _vp]),
    "vd_server_memory": (ctypes.c_int, [_vp, _vp]),
    "vd_server_job_grid": (ctypes.c_int, [_vp, ctypes.c_uint64, _vp]),
    "vd_server_set_option": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int64]),
    "vd_last_error": (ctypes.c_char_p, []),
}

_lib = None


class VdError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vd error {code}: {msg}")
        self.code = code


def lib():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m vorothree_amd.build` "
                "(the tessellation has no CPU fallback)")
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int) -> None:
    if rc != VD_OK:
        msg = lib().vd_last_error()
        raise VdError(rc, msg.decode() if msg else "")
