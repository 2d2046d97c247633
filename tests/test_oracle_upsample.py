"""The phone-depth prior (frontend/src/app/vision/mvs/depth-maps.ts:9-41 upsampleDepth): the host
function lr3d.upsample_depth (numpy) against the oracle's C restatement, bit for bit, and both
against closed forms: a constant grid stays constant, a grid linear in x and y is reproduced by the
bilinear weights away from the clamped border, a cell without data is skipped and a pixel whose
cells with data weigh at most 0.5 is 0."""
import numpy as np
import pytest

import oracle as O
from lr3d import upsample_depth


def grid(gw, gh, seed, holes=0.2):
    rng = np.random.default_rng(seed)
    m = rng.uniform(0.3, 5.0, gw * gh).astype(np.float32)
    m[rng.random(m.size) < holes] = 0.0
    m[rng.random(m.size) < 0.02] = -1.0
    m[rng.random(m.size) < 0.02] = np.nan
    return m


@pytest.mark.parametrize("gw,gh,w,h", [(256, 192, 1280, 720), (64, 48, 320, 240), (40, 30, 33, 17),
                                       (1, 1, 8, 6), (7, 1, 50, 3), (256, 192, 256, 192)])
def test_host_equals_oracle(gw, gh, w, h):
    m = grid(gw, gh, gw * 1000 + h)
    a = upsample_depth(m, gw, gh, w, h)
    b = O.upsample_depth(m, gw, gh, w, h)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_constant_grid_stays_constant():
    m = np.full(32 * 24, 1.7, dtype=np.float32)
    out = upsample_depth(m, 32, 24, 320, 240)
    assert np.all(np.abs(out - np.float32(1.7)) <= np.spacing(np.float32(1.7)))


def test_linear_grid_is_reproduced():
    gw, gh, w, h = 16, 12, 160, 120
    gx, gy = np.meshgrid(np.arange(gw), np.arange(gh))
    m = (1.0 + 0.05 * gx + 0.03 * gy).astype(np.float64)
    out = upsample_depth(m.astype(np.float32), gw, gh, w, h).reshape(h, w)
    x, y = np.meshgrid(np.arange(w), np.arange(h))
    cx = np.clip((x + 0.5) / w * gw - 0.5, 0, gw - 1)
    cy = np.clip((y + 0.5) / h * gh - 0.5, 0, gh - 1)
    expect = 1.0 + 0.05 * cx + 0.03 * cy
    assert np.abs(out - expect).max() < 1e-6


def test_missing_cells_are_skipped_and_weak_support_is_zero():
    # 2 x 2 grid, cell (1, 0) without data, upsampled 2x: pixel centres map to cell coordinates
    # -0.25, 0.25, 0.75, 1.25 (clamped to [0, 1])
    m = np.array([2.0, 0.0, 4.0, 6.0], dtype=np.float32)
    out = upsample_depth(m, 2, 2, 4, 4).reshape(4, 4)
    # pixel (0, 0) sits on cell (0, 0): weight 1 on a cell with data
    assert out[0, 0] == np.float32(2.0)
    # pixel (3, 0) sits on cell (1, 0), which has no data: the weight with data is 0
    assert out[0, 3] == 0.0
    # pixel (2, 0): gx = 0.75, gy = 0: weights 0.25 on (0, 0) (data), 0.75 on (1, 0) (none) -> 0
    assert out[0, 2] == 0.0
    # pixel (1, 1): gx = gy = 0.25: data weights 0.5625 (0,0) + 0.1875 (0,1) + 0.0625 (1,1) = 0.8125
    expect = (2.0 * 0.5625 + 4.0 * 0.1875 + 6.0 * 0.0625) / 0.8125
    assert out[1, 1] == np.float32(expect)
