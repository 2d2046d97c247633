"""Meshes extracted on the GPU: the reference's extractSurface + computeNormals
(frontend/src/app/vision/fusion/surface-nets.ts:47-166, mesh.ts:9-32) and removeSmallComponents
(mesh.ts:38-90) over lr3d_tsdf_extract_surface / lr3d_mesh_*."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N


class HipMesh:
    """A mesh held on the device (positions, normals, uint32 triangle indices)."""

    def __init__(self, handle):
        self._lib = N.load()
        self.handle = handle

    @property
    def size(self) -> tuple[int, int]:
        v, i = C.c_int64(), C.c_int64()
        N.check(self._lib.lr3d_mesh_size(self.handle, C.byref(v), C.byref(i)))
        return v.value, i.value

    def read(self) -> dict:
        """{positions (V, 3) f32, normals (V, 3) f32, indices (3T,) u32} (the reference's Mesh)."""
        v, i = self.size
        pos = np.empty((v, 3), dtype=np.float32)
        nrm = np.empty((v, 3), dtype=np.float32)
        idx = np.empty(i, dtype=np.uint32)
        N.check(self._lib.lr3d_mesh_read(self.handle, N.ptr(pos), N.ptr(nrm), N.ptr(idx)))
        return {"positions": pos, "normals": nrm, "indices": idx}

    def remove_small_components(self, fraction: float, min_faces: int) -> "HipMesh":
        """mesh.ts:38-90: a new mesh without the components smaller than max(min_faces, fraction x faces)."""
        out = C.c_void_p()
        N.check(self._lib.lr3d_mesh_remove_small_components(self.handle, float(fraction), int(min_faces), C.byref(out)))
        return HipMesh(out)

    def destroy(self) -> None:
        if self.handle:
            self._lib.lr3d_mesh_destroy(self.handle)
            self.handle = None
