"""Pinhole and pose conventions of the reference (frontend/src/app/vision/geometry/camera.ts).

Poses are world -> camera in computer-vision axes (x right, y down, z forward):
``x_cam = R @ X + t`` with R row-major (camera.ts:19-26).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Intrinsics:
    fx: float
    fy: float
    cx: float
    cy: float


@dataclass
class Pose:
    R: np.ndarray = field(default_factory=lambda: np.eye(3))
    t: np.ndarray = field(default_factory=lambda: np.zeros(3))

    def __post_init__(self):
        self.R = np.asarray(self.R, dtype=np.float64).reshape(3, 3)
        self.t = np.asarray(self.t, dtype=np.float64).reshape(3)

    @property
    def center(self) -> np.ndarray:
        """Camera centre in world coordinates (camera.ts:42-45)."""
        return -(self.R.T @ self.t)


@dataclass
class StereoView:
    """A stereo layer: intrinsics and pose at its own image size (dense-stereo.ts:57-60)."""
    intrinsics: Intrinsics
    pose: Pose
    width: int
    height: int


def scale_intrinsics(k: Intrinsics, sx: float, sy: float | None = None) -> Intrinsics:
    """camera.ts:28-36: pixel centres stay aligned, (c + 0.5) * s - 0.5."""
    sy = sx if sy is None else sy
    return Intrinsics(k.fx * sx, k.fy * sy, (k.cx + 0.5) * sx - 0.5, (k.cy + 0.5) * sy - 0.5)


def relative_pose(a: Pose, b: Pose) -> Pose:
    """camera.ts:73-77: transform taking camera-a coordinates to camera-b coordinates."""
    R = b.R @ a.R.T
    return Pose(R, b.t - R @ a.t)


def project(k: Intrinsics, pose: Pose, x) -> np.ndarray:
    """camera.ts:62-65: [u, v, depth] of a world point."""
    p = pose.R @ np.asarray(x, dtype=np.float64) + pose.t
    return np.array([k.fx * p[0] / p[2] + k.cx, k.fy * p[1] / p[2] + k.cy, p[2]])


def look_at(center, target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0)) -> Pose:
    """Camera at `center` looking at `target`, world +y up (frontend/tools/scene-renderer.ts:11-20)."""
    c = np.asarray(center, dtype=np.float64)
    z = np.asarray(target, dtype=np.float64) - c
    z /= np.linalg.norm(z)
    x = np.cross(z, np.asarray(up, dtype=np.float64))
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    R = np.stack([x, y, z])
    return Pose(R, -R @ c)
