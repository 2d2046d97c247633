"""Host decode of the packed per-pixel depth format and the phone-depth prior
(frontend/src/app/vision/mvs/depth-maps.ts).

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


def upsample_depth(meters, grid_width: int, grid_height: int, width: int, height: int) -> np.ndarray:
    """depth-maps.ts:9-41 upsampleDepth: bilinear upsampling of a phone depth grid (0 = no data) to
    width x height, ignoring cells without data; a pixel is 0 unless the weights of the cells with
    data sum to more than 0.5.  The prior input of a solve (float32, row-major).

    Computed in float64 with the JavaScript's operations in its order (the per-pixel sums add the
    four samples in order, skipping those without data) and rounded to float32 like the
    Float32Array it fills."""
    g = np.asarray(meters, dtype=np.float32).reshape(grid_height, grid_width).astype(np.float64)
    gw, gh = int(grid_width), int(grid_height)

    def axis(n, gn):
        c = np.minimum(gn - 1, np.maximum(0.0, ((np.arange(n) + 0.5) / n) * gn - 0.5))
        c0 = np.floor(c)
        c1 = np.minimum(gn - 1, c0 + 1)
        return c0.astype(np.int64), c1.astype(np.int64), c - c0

    y0, y1, wy = axis(int(height), gh)
    x0, x1, wx = axis(int(width), gw)
    wy, wx = wy[:, None], wx[None, :]
    samples = [(g[y0][:, x0], (1 - wx) * (1 - wy)), (g[y0][:, x1], wx * (1 - wy)),
               (g[y1][:, x0], (1 - wx) * wy), (g[y1][:, x1], wx * wy)]
    total = np.zeros((int(height), int(width)))
    weight = np.zeros((int(height), int(width)))
    with np.errstate(invalid="ignore"):
        for d, w in samples:
            w = np.broadcast_to(w, total.shape)
            ok = d > 0
            total = np.where(ok, total + d * w, total)
            weight = np.where(ok, weight + w, weight)
        out = np.where(weight > 0.5, total / np.where(weight > 0.5, weight, 1.0), 0.0)
    return out.astype(np.float32).reshape(-1)
