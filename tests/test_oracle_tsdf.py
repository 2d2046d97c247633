"""Pins of the oracle's TSDF integration (TV:103-154 kernel, TV:220-234 culling, TV:54-64 layout)
by closed forms on a fronto-parallel view of a constant-depth map."""
import numpy as np
import pytest

import oracle as O
from lr3d import BrickGrid

F, W, H = 100.0, 80, 60


def packed_map(depth, normal=(0.0, 0.0, -1.0), cost=0.0):
    p = O.pack_sample(depth, normal, cost)
    return np.full(W * H, p, dtype=np.uint32)


def setup(voxel=0.01, bricks=(2, 2, 4), origin=(-0.08, -0.08, 0.5), trunc=4.0):
    dims = list(bricks)
    coords = np.zeros(dims[0] * dims[1] * dims[2] * 4, dtype=np.int32)
    i = 0
    for bz in range(dims[2]):
        for by in range(dims[1]):
            for bx in range(dims[0]):
                coords[i * 4:i * 4 + 4] = (bx, by, bz, 0)
                i += 1
    g = O.Grid(origin, voxel, dims, coords, voxel * trunc)
    view = O.integration_view(W, H, F, F, (W - 1) / 2, (H - 1) / 2, np.eye(3), np.zeros(3), 5.0)
    return g, view, coords


def voxel_centres(g, coords, origin, voxel):
    out = []
    for b in range(g.n_bricks):
        base = coords[b * 4:b * 4 + 3] * 8
        for local in range(512):
            v = np.array([local % 8, (local // 8) % 8, local // 64])
            out.append(np.asarray(origin) + (base + v + 0.5) * voxel)
    return np.array(out)


def test_dense_layout_matches_brickgrid_build():
    dims = (3, 2, 2)
    grid = BrickGrid((0, 0, 0), 0.01, (0.24, 0.16, 0.16))
    assert grid.dims == dims
    grid.mark_all()
    coords, lookup, count = grid.build()
    assert count == 12 and np.array_equal(lookup, np.arange(12))
    # pool order = ascending grid index (bz * dy + by) * dx + bx
    c = coords.reshape(-1, 4)
    assert np.array_equal((c[:, 2] * 2 + c[:, 1]) * 3 + c[:, 0], np.arange(12))


def test_brickgrid_mark_touches_the_cube_around_a_point():
    grid = BrickGrid((0, 0, 0), 0.01, (0.8, 0.8, 0.8))  # 10^3 bricks of 8 cm
    grid.mark(0.40, 0.40, 0.40, 0.05)  # cube [0.35, 0.45]^3 spans bricks 4..5 on each axis
    coords, lookup, count = grid.build()
    assert count == 8
    assert set(map(tuple, coords.reshape(-1, 4)[:, :3])) == {(x, y, z) for x in (4, 5) for y in (4, 5) for z in (4, 5)}


def test_fronto_parallel_closed_form():
    origin, voxel = (-0.08, -0.08, 0.5), 0.01
    g, view, coords = setup(voxel=voxel, origin=origin)
    d = np.float32(0.7)
    tsdf = np.ones(g.n_bricks * 512, dtype=np.float32)
    weight = np.zeros_like(tsdf)
    updates = O.tsdf_integrate(g, view, packed_map(float(d)), None, tsdf, weight)
    centres = voxel_centres(g, coords, origin, voxel)
    mu = voxel * 4
    z = centres[:, 2]
    depth = float(O.unpack_depth(int(packed_map(float(d))[0])))
    sdf = depth - z
    upd = sdf >= -mu
    assert updates == int(upd.sum())
    want = np.minimum(1.0, sdf / mu)
    assert np.allclose(tsdf[upd], want[upd], atol=1e-5)
    assert np.all(tsdf[~upd] == 1.0) and np.all(weight[~upd] == 0.0)
    # weight = max(0.2, |n . normalize(cam)|) with n = (0, 0, -1)
    wexp = np.maximum(0.2, np.abs(z / np.linalg.norm(centres, axis=1)))
    assert np.allclose(weight[upd], wexp[upd], atol=1e-6)


def test_running_mean_and_weight_cap():
    origin, voxel = (-0.08, -0.08, 0.5), 0.01
    g, view, coords = setup(voxel=voxel, origin=origin)
    tsdf = np.ones(g.n_bricks * 512, dtype=np.float32)
    weight = np.zeros_like(tsdf)
    O.tsdf_integrate(g, view, packed_map(0.70), None, tsdf, weight)
    t1, w1 = tsdf.copy(), weight.copy()
    O.tsdf_integrate(g, view, packed_map(0.74), None, tsdf, weight)
    t2 = np.minimum(1.0, (float(O.unpack_depth(int(packed_map(0.74)[0]))) -
                          voxel_centres(g, coords, origin, voxel)[:, 2]) / (4 * voxel))
    both = (w1 > 0) & (weight > w1)
    assert np.allclose(tsdf[both], (t1[both] * w1[both] + t2[both] * w1[both]) / (2 * w1[both]), atol=1e-5)
    for _ in range(400):
        O.tsdf_integrate(g, view, packed_map(0.70), None, tsdf, weight)
    assert weight.max() == 96.0


def test_prior_only_pixels_use_the_prior_weight():
    origin, voxel = (-0.08, -0.08, 0.5), 0.01
    g, view, coords = setup(voxel=voxel, origin=origin)
    tsdf = np.ones(g.n_bricks * 512, dtype=np.float32)
    weight = np.zeros_like(tsdf)
    prior = np.full(W * H, 0.7, dtype=np.float32)
    O.tsdf_integrate(g, view, np.zeros(W * H, np.uint32), prior, tsdf, weight)
    assert set(np.unique(weight)) == {np.float32(0.0), np.float32(0.12)}


def test_confidence_weighting():
    origin, voxel = (-0.08, -0.08, 0.5), 0.01
    g, _, coords = setup(voxel=voxel, origin=origin)
    view = O.integration_view(W, H, F, F, (W - 1) / 2, (H - 1) / 2, np.eye(3), np.zeros(3), 5.0, confidence=1.4)
    tsdf = np.ones(g.n_bricks * 512, dtype=np.float32)
    weight = np.zeros_like(tsdf)
    O.tsdf_integrate(g, view, packed_map(0.7, cost=0.4), None, tsdf, weight)
    centres = voxel_centres(g, coords, origin, voxel)
    base = np.maximum(0.2, np.abs(centres[:, 2] / np.linalg.norm(centres, axis=1)))
    c = O.unpack_cost(int(packed_map(0.7, cost=0.4)[0])) / 8.0
    r = np.clip(1.4 / O.unpack_depth(int(packed_map(0.7)[0])), 0.5, 2.0)
    want = base * max(0.15, (1 - c) ** 2) * r * r
    upd = weight > 0
    assert np.allclose(weight[upd], want[upd], rtol=1e-5)


def test_culling_keeps_every_brick_that_receives_an_update():
    origin, voxel = (-1.2, -1.2, 0.2), 0.02  # 15 x 15 x 6 bricks of 16 cm, wider than the frustum
    g, view, coords = setup(voxel=voxel, bricks=(15, 15, 6), origin=origin)
    flags = O.visible_bricks(g, view)
    tsdf = np.ones(g.n_bricks * 512, dtype=np.float32)
    weight = np.zeros_like(tsdf)
    O.tsdf_integrate(g, view, packed_map(0.9), None, tsdf, weight)
    touched = (weight.reshape(-1, 512) > 0).any(axis=1)
    assert np.all(flags[touched] == 1)
    assert flags.sum() < g.n_bricks  # some bricks are culled
