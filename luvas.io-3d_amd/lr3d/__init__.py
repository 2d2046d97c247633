"""lr3d: MI355X (gfx950) PatchMatch MVS + TSDF integration behind the C ABI of include/lr3d.h.

Host-side mirror of the reference's stage API (frontend/src/app/vision/mvs/dense-stereo.ts,
frontend/src/app/vision/fusion/tsdf-volume.ts, frontend/src/app/vision/mvs/methods/):
``HipDenseStereo``, ``HipTsdfVolume`` and ``BrickGrid`` (the ``STEREO_METHODS`` registry lives in
the Node host, luvas.io-3d_amd/node/lr3d.js), plus the adjoints (``HipDenseStereo.ncc_backward``,
``HipTsdfVolume.backward``) and the gradient refinement driver ``refine``.
"""
from ._native import Lr3dError, MAP_FILTERED, MAP_RAW, MAX_SOURCES
from .camera import Intrinsics, Pose, StereoView, look_at, relative_pose, scale_intrinsics
from .runtime import Context
from .stereo import HipDenseStereo, SolveOptions
from .tsdf import BRICK, BrickGrid, HipTsdfVolume, IntegrationView, dense_grid
from . import refine
from .mesh import HipMesh
from .fusion import fuse_depth_maps
from .packing import upsample_depth

__all__ = [
    "Lr3dError", "MAP_FILTERED", "MAP_RAW", "MAX_SOURCES", "Intrinsics", "Pose", "StereoView",
    "look_at", "relative_pose", "scale_intrinsics", "Context", "HipDenseStereo", "SolveOptions",
    "BRICK", "BrickGrid", "HipTsdfVolume", "IntegrationView", "dense_grid", "refine", "HipMesh", "fuse_depth_maps",
    "upsample_depth",
]
