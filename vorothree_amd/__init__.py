"""MI355X (gfx950) implementation of voronoid's Tessellation<3, Cell3DFaces, Algorithm3DGrid>.

The compute path is libvoronoid_hip.so (HIP kernels behind the C ABI of include/vd.h); this
package mirrors the reference's Rust API on top of it.
"""
from .tessellation import (  # noqa: F401
    WALL_ID_MAX,
    Algorithm3DGrid,
    Batch,
    BoundingBox,
    CatmullRomGeometry,
    Cell3DFaces,
    ConeGeometry,
    ConvexPolyhedronGeometry,
    CubicBezierGeometry,
    CylinderGeometry,
    PlaneGeometry,
    Server,
    SphereGeometry,
    Tessellation,
    TorusGeometry,
    TrefoilKnotGeometry,
    Wall,
    box_side,
    calculate_static,
    grid_for,
    grid_for_points,
    parse_generators,
    shell_prefix,
)
