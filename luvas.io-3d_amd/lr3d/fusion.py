"""Dense-cloud fusion on the GPU: the reference's fuseDepthMaps (frontend/src/app/vision/fusion/
depth-fusion.ts:23-74) with its PointAccumulator (point-accumulator.ts), over lr3d_fuse_depth_maps."""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import numpy as np

from . import _native as N


class Thumb(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgb", C.c_void_p)]


def fuse_depth_maps(stereo, layers: Sequence[int], thumbs: Sequence[np.ndarray], cell: float, cell_budget: int,
                    map: int = N.MAP_FILTERED, min_support: int = 1):
    """Returns ({positions (P, 3) f32, normals (P, 3) f32, colors (P, 3) u8, support (P,) u16}, final cell).
    thumbs: per layer an (h, w, 3) u8 RGB image of any size.  Points come out in cell-key order."""
    lib = N.load()
    keep = [np.ascontiguousarray(t, dtype=np.uint8) for t in thumbs]
    arr = (Thumb * len(layers))(*[Thumb(t.shape[1], t.shape[0], t.ctypes.data) for t in keep])
    ls = (C.c_int32 * len(layers))(*layers)
    handle = C.c_void_p()
    final = C.c_double()
    N.check(lib.lr3d_fuse_depth_maps(stereo.handle, ls, map, len(layers), arr, float(cell), int(cell_budget),
                                     int(min_support), C.byref(handle), C.byref(final)))
    try:
        n = C.c_int64()
        N.check(lib.lr3d_cloud_size(handle, C.byref(n)))
        p = n.value
        out = {"positions": np.empty((p, 3), np.float32), "normals": np.empty((p, 3), np.float32),
               "colors": np.empty((p, 3), np.uint8), "support": np.empty(p, np.uint16)}
        N.check(lib.lr3d_cloud_read(handle, N.ptr(out["positions"]), N.ptr(out["normals"]), N.ptr(out["colors"]),
                                    N.ptr(out["support"])))
    finally:
        lib.lr3d_cloud_destroy(handle)
    return out, final.value
