"""BrickGrid and HipTsdfVolume: the reference's TSDF volume
(frontend/src/app/vision/fusion/tsdf-volume.ts, TV:13-291) on the lr3d C ABI.

The pool layout is the reference's: 8^3 bricks in ascending grid order, voxel (x, y, z) of pool
brick b at b*512 + x + 8y + 64z; ``read()`` returns the flat tsdf / weight pools that
surface-nets (frontend/src/app/vision/fusion/surface-nets.ts:4-12) consumes with ``lookup``.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import _native as N
from .camera import Intrinsics, Pose
from .runtime import Context

BRICK = 8
BRICK_VOXELS = BRICK ** 3


class BrickGrid:
    """TV:13-65: sparse set of 8^3 voxel bricks on a bounded grid."""

    def __init__(self, origin, voxel: float, extent):
        self.origin = tuple(float(x) for x in origin)
        self.voxel = float(voxel)
        side = voxel * BRICK
        self.dims = tuple(max(1, math.ceil(e / side)) for e in extent)
        self.marks = np.zeros(self.dims[0] * self.dims[1] * self.dims[2], dtype=np.uint8)

    @property
    def cells(self) -> int:
        return int(self.marks.size)

    def mark(self, x: float, y: float, z: float, radius: float) -> None:
        """TV:36-51: marks every brick touched by the cube of half-size `radius`."""
        side = self.voxel * BRICK
        ox, oy, oz = self.origin
        dx, dy, dz = self.dims
        x0 = max(0, math.floor((x - radius - ox) / side))
        x1 = min(dx - 1, math.floor((x + radius - ox) / side))
        y0 = max(0, math.floor((y - radius - oy) / side))
        y1 = min(dy - 1, math.floor((y + radius - oy) / side))
        z0 = max(0, math.floor((z - radius - oz) / side))
        z1 = min(dz - 1, math.floor((z + radius - oz) / side))
        for bz in range(z0, z1 + 1):
            for by in range(y0, y1 + 1):
                self.marks[(bz * dy + by) * dx + x0:(bz * dy + by) * dx + x1 + 1] = 1

    def mark_all(self) -> None:
        """A dense volume: every cell allocated."""
        self.marks[:] = 1

    def build(self):
        """TV:54-64: (coords int32 n*4, lookup int32 cells, count) in ascending grid order."""
        dx, dy = self.dims[0], self.dims[1]
        idx = np.flatnonzero(self.marks).astype(np.int64)
        lookup = np.full(self.marks.size, -1, dtype=np.int32)
        lookup[idx] = np.arange(idx.size, dtype=np.int32)
        coords = np.zeros((idx.size, 4), dtype=np.int32)
        coords[:, 0] = idx % dx
        coords[:, 1] = (idx // dx) % dy
        coords[:, 2] = idx // (dx * dy)
        return coords.reshape(-1), lookup, int(idx.size)


@dataclass
class IntegrationView:
    """TV:67-79 plus the map size."""
    intrinsics: Intrinsics
    pose: Pose
    max_depth: float
    width: int
    height: int
    prior: Optional[np.ndarray] = None
    confidence: float = 0.0


def _iv_struct(v: IntegrationView):
    s = N.IntegrationView()
    s.width, s.height = int(v.width), int(v.height)
    k = v.intrinsics
    s.fx, s.fy, s.cx, s.cy = k.fx, k.fy, k.cx, k.cy
    s.R[:] = [float(x) for x in np.asarray(v.pose.R, dtype=np.float64).reshape(9)]
    s.t[:] = [float(x) for x in np.asarray(v.pose.t, dtype=np.float64).reshape(3)]
    s.max_depth = float(v.max_depth)
    prior = None
    if v.prior is not None:
        prior = np.ascontiguousarray(v.prior, dtype=np.float32).reshape(-1)
    s.prior = N.float_ptr(prior)
    s.confidence = float(v.confidence)
    return s, prior


class HipTsdfVolume:
    """TV:157-291: truncated signed distance volume on the GPU (tsdf := 1, weight := 0)."""

    def __init__(self, ctx: Context, grid: BrickGrid, layout, truncation_voxels: float = 4.0):
        self._lib = N.load()
        self.ctx = ctx
        self.grid = grid
        coords, lookup, count = layout
        self.layout = (np.ascontiguousarray(coords, dtype=np.int32), lookup, count)
        self.truncation = grid.voxel * truncation_voxels  # TV:179
        origin = (C.c_double * 3)(*grid.origin)
        dims = (C.c_int32 * 3)(*grid.dims)
        handle = C.c_void_p()
        N.check(self._lib.lr3d_tsdf_create(ctx.handle, origin, grid.voxel, dims,
                                           N.ptr(self.layout[0]), count, float(truncation_voxels),
                                           C.byref(handle)))
        self.handle = handle

    @property
    def count(self) -> int:
        return self.layout[2]

    def integrate(self, stereo, layer: int, view: IntegrationView, map: int = N.MAP_RAW) -> None:
        """TV:240-268 for one depth map of a stereo layer."""
        s, keep = _iv_struct(view)
        N.check(self._lib.lr3d_tsdf_integrate(self.handle, stereo.handle, layer, map, C.byref(s)))

    def integrate_host(self, packed: np.ndarray, view: IntegrationView) -> None:
        p = np.ascontiguousarray(packed, dtype=np.uint32).reshape(-1)
        s, keep = _iv_struct(view)
        N.check(self._lib.lr3d_tsdf_integrate_host(self.handle, N.ptr(p), C.byref(s)))

    def integrate_batched(self, stereo, layers: Sequence[int], views: Sequence[IntegrationView],
                          map: int = N.MAP_RAW) -> None:
        """Every view in one pass; per voxel the views apply in list order."""
        structs, keep = [], []
        for v in views:
            s, k = _iv_struct(v)
            structs.append(s)
            keep.append(k)
        arr = (N.IntegrationView * len(structs))(*structs)
        ls = (C.c_int32 * len(layers))(*layers)
        N.check(self._lib.lr3d_tsdf_integrate_batched(self.handle, stereo.handle, ls, map, arr,
                                                      len(structs)))

    def autotune(self, stereo, layers: Sequence[int], views: Sequence[IntegrationView], map: int = N.MAP_RAW,
                 repeats: int = 3) -> int:
        """Times integrate passes with 1, 2 and 4 bricks per block, each with several chunks of
        consecutive blocks per XCD, and keeps the fastest pair; returns the bricks per block (the
        chunk is `xcd_chunk`)."""
        structs, keep = [], []
        for v in views:
            s, k = _iv_struct(v)
            structs.append(s)
            keep.append(k)
        arr = (N.IntegrationView * len(structs))(*structs)
        ls = (C.c_int32 * len(layers))(*layers)
        N.check(self._lib.lr3d_tsdf_autotune(self.handle, stereo.handle, ls, map, arr, len(structs), int(repeats)))
        return self.bricks_per_block

    @property
    def bricks_per_block(self) -> int:
        v = C.c_int32()
        N.check(self._lib.lr3d_tsdf_get_bricks_per_block(self.handle, C.byref(v)))
        return v.value

    @bricks_per_block.setter
    def bricks_per_block(self, bricks: int) -> None:
        N.check(self._lib.lr3d_tsdf_set_bricks_per_block(self.handle, int(bricks)))

    @property
    def xcd_chunk(self) -> int:
        """Consecutive blocks of the integration pass given to one XCD (1: round robin)."""
        v = C.c_int32()
        N.check(self._lib.lr3d_tsdf_get_xcd_chunk(self.handle, C.byref(v)))
        return v.value

    @xcd_chunk.setter
    def xcd_chunk(self, chunk: int) -> None:
        N.check(self._lib.lr3d_tsdf_set_xcd_chunk(self.handle, int(chunk)))

    def set_depth_culling(self, on: bool) -> None:
        """Drop (brick, view) pairs that cannot update a voxel (on by default; results unchanged)."""
        N.check(self._lib.lr3d_tsdf_set_depth_culling(self.handle, int(bool(on))))

    def backward(self, stereo, layers: Sequence[int], views: Sequence[IntegrationView], voxels, targets,
                 map: int = N.MAP_RAW, grad_depth: bool = True) -> dict:
        """Replay of integrate_batched into a fresh volume at the listed voxels (pool indices) and the
        adjoint of L = sum (tsdf - target)^2: dict(tsdf, grad_depth (list per view, H x W), grad_pose
        (n, 12): dL/dR row-major and dL/dt per view, loss).  grad_depth=False leaves the per-pixel
        depth gradients on the device (not copied to the host; "grad_depth" is then None)."""
        structs, keep = [], []
        for v in views:
            s, k = _iv_struct(v)
            structs.append(s)
            keep.append(k)
        arr = (N.IntegrationView * len(structs))(*structs)
        ls = (C.c_int32 * len(layers))(*layers)
        vox = np.ascontiguousarray(voxels, dtype=np.uint32).reshape(-1)
        tgt = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1)
        sizes = [v.width * v.height for v in views]
        tsdf = np.empty(vox.size, dtype=np.float32)
        gd = np.empty(sum(sizes), dtype=np.float32) if grad_depth else None
        gp = np.empty((len(views), 12), dtype=np.float64)
        loss = np.zeros(1, dtype=np.float64)
        N.check(self._lib.lr3d_tsdf_backward(self.handle, stereo.handle, ls, map, arr, len(structs), N.ptr(vox),
                                             N.ptr(tgt), vox.size, N.ptr(tsdf), N.float_ptr(gd), N.ptr(gp),
                                             N.ptr(loss)))
        parts = np.split(gd, np.cumsum(sizes)[:-1]) if grad_depth else None
        return {"tsdf": tsdf,
                "grad_depth": [p.reshape(v.height, v.width) for p, v in zip(parts, views)] if grad_depth else None,
                "grad_pose": gp, "loss": float(loss[0])}

    def write(self, tsdf: np.ndarray, weight: np.ndarray) -> None:
        """Overwrite the tsdf and weight pools (count * 512 floats each)."""
        n = self.count * BRICK_VOXELS
        t = np.ascontiguousarray(tsdf, dtype=np.float32).reshape(n)
        w = np.ascontiguousarray(weight, dtype=np.float32).reshape(n)
        N.check(self._lib.lr3d_tsdf_write(self.handle, N.ptr(t), N.ptr(w)))

    def extract_surface(self, min_weight: float = 0.6):
        """surface-nets.ts extractSurface (with computeNormals) on the GPU; returns a HipMesh."""
        from .mesh import HipMesh
        out = C.c_void_p()
        N.check(self._lib.lr3d_tsdf_extract_surface(self.handle, float(min_weight), C.byref(out)))
        return HipMesh(out)

    def read(self) -> tuple[np.ndarray, np.ndarray]:
        """TV:270-277"""
        n = self.count * BRICK_VOXELS
        tsdf = np.empty(n, dtype=np.float32)
        weight = np.empty(n, dtype=np.float32)
        N.check(self._lib.lr3d_tsdf_read(self.handle, N.ptr(tsdf), N.ptr(weight)))
        return tsdf, weight

    def set_counting(self, enable: bool) -> None:
        N.check(self._lib.lr3d_tsdf_set_counting(self.handle, 1 if enable else 0))

    def update_count(self) -> int:
        v = C.c_int64()
        N.check(self._lib.lr3d_tsdf_update_count(self.handle, C.byref(v)))
        return v.value

    def pass_counts(self) -> dict:
        """Totals of the counting passes: updates, (voxel, view) pairs evaluated, bricks read,
        voxels written."""
        v = (C.c_int64 * 4)()
        N.check(self._lib.lr3d_tsdf_pass_counts(self.handle, v))
        return {"updates": v[0], "pairs": v[1], "bricks_read": v[2], "voxels_written": v[3]}

    def destroy(self) -> None:
        if self.handle:
            self._lib.lr3d_tsdf_destroy(self.handle)
            self.handle = None


def dense_grid(origin, voxel: float, voxels: int) -> tuple[BrickGrid, tuple]:
    """A dense cube of voxels^3 (voxels a multiple of 8) at `origin`, every brick allocated."""
    assert voxels % BRICK == 0
    extent = (voxels * voxel,) * 3
    grid = BrickGrid(origin, voxel, extent)
    # guard against ceil() rounding one brick up when extent/side is not exact in floating point
    grid.dims = (voxels // BRICK,) * 3
    grid.marks = np.ones(grid.dims[0] ** 3, dtype=np.uint8)
    return grid, grid.build()
