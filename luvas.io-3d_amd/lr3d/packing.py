"""Host decode of the packed per-pixel depth format (frontend/src/app/vision/mvs/depth-maps.ts).

u32 = f16 depth (bits 0-15) | octahedral normal q.x, q.y (6 bits each, bits 16-27) |
min(15, floor(15 * cost)) (bits 28-31); 0 = no depth.
"""
from __future__ import annotations

import numpy as np


def half_to_float(bits) -> np.ndarray:
    """depth-maps.ts:44-60 (IEEE half -> float; numpy's float16 has the same table)."""
    return np.asarray(bits, dtype=np.uint32).astype(np.uint16).view(np.float16).astype(np.float32)


def unpack_depth(packed) -> np.ndarray:
    return half_to_float(np.asarray(packed, dtype=np.uint32) & 0xFFFF)


def unpack_cost(packed) -> np.ndarray:
    return np.asarray(packed, dtype=np.uint32) >> 28


def unpack_normal(packed) -> np.ndarray:
    """depth-maps.ts:66-85 in float64: octahedral unfold, then normalise."""
    p = np.asarray(packed, dtype=np.uint32)
    ex = ((p >> 16) & 63).astype(np.float64) / 63 * 2 - 1
    ey = ((p >> 22) & 63).astype(np.float64) / 63 * 2 - 1
    nz = 1 - np.abs(ex) - np.abs(ey)
    fold = nz < 0
    nx = np.where(fold, (1 - np.abs(ey)) * np.where(ex >= 0, 1.0, -1.0), ex)
    ny = np.where(fold, (1 - np.abs(ex)) * np.where(ey >= 0, 1.0, -1.0), ey)
    n = np.stack([nx, ny, nz], axis=-1)
    length = np.linalg.norm(n, axis=-1, keepdims=True)
    return n / np.where(length > 0, length, 1)


def unpack_depth_sample(value: int):
    """depth-maps.ts:66-85: (depth, normal) or None for an empty pixel."""
    if value == 0:
        return None
    d = float(unpack_depth(value))
    if not d > 0:
        return None
    return d, unpack_normal(value)
