"""Pins of the oracle's dense-cloud fusion (oracle/fusion_oracle.c).

The reference's own known answers for the accumulator (fusion.spec.ts:85-106): samples average per
cell, cells merge when coarsened, and the table grows past its initial capacity; a closed form for
what the growth's rebuild does (it re-bins every cell by its Float32 mean); and closed forms
for the whole fusion: a fronto-parallel constant-depth map seen by an identity camera lands on one
plane of cells with the depth, the camera-facing normal and the thumbnail colour."""
import ctypes as C

import numpy as np

import oracle as O
from lr3d import Intrinsics, Pose, StereoView


def _acc():
    lib = O.load()
    if not hasattr(lib, "_acc"):
        lib.orc_acc_create.restype = C.c_void_p
        lib.orc_acc_create.argtypes = [C.c_double]
        lib.orc_acc_add.restype = None
        lib.orc_acc_add.argtypes = [C.c_void_p, C.POINTER(C.c_double)]
        lib.orc_acc_coarsen.restype = None
        lib.orc_acc_coarsen.argtypes = [C.c_void_p, C.c_double]
        lib.orc_acc_size.restype = C.c_int64
        lib.orc_acc_size.argtypes = [C.c_void_p]
        lib.orc_acc_cloud.restype = C.c_int64
        lib.orc_acc_cloud.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
        lib.orc_acc_destroy.restype = None
        lib.orc_acc_destroy.argtypes = [C.c_void_p]
        lib._acc = True
    return lib


def test_accumulator_averages_and_coarsens():
    """fusion.spec.ts:86-99"""
    lib = _acc()
    a = lib.orc_acc_create(0.01)
    for v in ((0.001, 0.001, 0.001, 0, 0, 1, 100, 0, 0), (0.003, 0.003, 0.003, 0, 0, 1, 200, 0, 0),
              (0.015, 0.001, 0.001, 1, 0, 0, 0, 50, 0)):
        lib.orc_acc_add(a, (C.c_double * 9)(*v))
    pos = np.zeros((8, 3), np.float32)
    col = np.zeros((8, 3), np.uint8)
    sup = np.zeros(8, np.uint16)
    n = lib.orc_acc_cloud(a, 1, pos.ctypes.data, col.ctypes.data, sup.ctypes.data)
    assert n == 2
    first = int(np.flatnonzero(sup[:n] == 2)[0])
    assert abs(pos[first, 0] - 0.002) < 5e-7
    assert col[first, 0] == 150
    lib.orc_acc_coarsen(a, 2.0)
    n = lib.orc_acc_cloud(a, 1, pos.ctypes.data, col.ctypes.data, sup.ctypes.data)
    assert n == 1 and sup[0] == 3
    lib.orc_acc_destroy(a)


def test_accumulator_grows_past_its_capacity():
    """fusion.spec.ts:101-105"""
    lib = _acc()
    a = lib.orc_acc_create(0.001)
    for i in range(100000):
        lib.orc_acc_add(a, (C.c_double * 9)(i * 0.002, 0, 0, 0, 1, 0, 1, 2, 3))
    assert lib.orc_acc_size(a) == 100000
    lib.orc_acc_destroy(a)


def test_fronto_parallel_map_fuses_onto_one_plane():
    W, H, f, depth = 64, 48, 80.0, 0.75
    k = Intrinsics(f, f, (W - 1) / 2, (H - 1) / 2)
    view = StereoView(k, Pose(np.eye(3), np.zeros(3)), W, H)
    packed = np.full(W * H, O.pack_sample(depth, (0.0, 0.0, -1.0), 0.0), np.uint32)
    thumb = np.zeros((12, 16, 3), np.uint8)
    thumb[..., 0] = 40
    thumb[..., 1] = 90
    thumb[..., 2] = 200
    cloud, cell = O.fuse_depth_maps([view], [packed], [thumb], 0.005, 1 << 20)
    d16 = float(np.float16(depth))
    assert cell == 0.005
    assert cloud["support"].sum() == W * H
    assert np.allclose(cloud["positions"][:, 2], d16, atol=1e-6)
    assert np.allclose(cloud["normals"], [0, 0, -1], atol=1e-6)
    assert (cloud["colors"] == [40, 90, 200]).all()
    # one point per 5 mm cell the pixels fall in
    xs = ((np.arange(W) - k.cx) / f) * d16
    ys = ((np.arange(H) - k.cy) / f) * d16
    cells = {(int(np.floor(x / 0.005)), int(np.floor(y / 0.005))) for x in xs for y in ys}
    assert len(cloud["support"]) == len(cells)


def test_cell_budget_coarsens_the_grid():
    W, H, f = 64, 48, 80.0
    k = Intrinsics(f, f, (W - 1) / 2, (H - 1) / 2)
    view = StereoView(k, Pose(np.eye(3), np.zeros(3)), W, H)
    packed = np.full(W * H, O.pack_sample(0.75, (0.0, 0.0, -1.0), 0.0), np.uint32)
    thumb = np.full((12, 16, 3), 7, np.uint8)
    full, _ = O.fuse_depth_maps([view], [packed], [thumb], 0.002, 1 << 20)
    small, cell = O.fuse_depth_maps([view], [packed], [thumb], 0.002, 200)
    assert len(full["support"]) > 200 >= len(small["support"])
    assert cell > 0.002 and small["support"].sum() == W * H


def test_table_growth_rebins_a_cell_whose_float_mean_leaves_it():
    """PA:67,87-110: the table doubles (65536 -> 131072 slots) before the first sample after more than
    60 % of its slots are used (39322 cells), and the rebuild re-bins every cell by its mean position
    at the same cell size.  A sample at x = 1000.99999 lies in cell 1000 (1 m cells), but its Float32
    sum rounds to 1001 (float spacing there is 2^-14), so the growth moves its cell to 1001; a later
    sample at x = 1000.5 then opens a new cell 1000 instead of joining it."""
    assert np.float32(1000.99999) == np.float32(1001.0)  # the closed form the case rests on
    lib = _acc()
    a = lib.orc_acc_create(1.0)
    lib.orc_acc_add(a, (C.c_double * 9)(1000.99999, 0.5, 0.5, 0, 0, 1, 10, 20, 30))
    for i in range(1, 39322):  # 39321 more cells: 39322 in use, more than 60 % of 65536 slots
        lib.orc_acc_add(a, (C.c_double * 9)(5.5, 2.0 * i + 0.5, 0.5, 0, 0, 1, 0, 0, 0))
    assert lib.orc_acc_size(a) == 39322
    lib.orc_acc_add(a, (C.c_double * 9)(1000.5, 0.5, 0.5, 0, 0, 1, 40, 50, 60))  # growth, then cell 1000
    assert lib.orc_acc_size(a) == 39323
    pos = np.zeros((39323, 3), np.float32)
    col = np.zeros((39323, 3), np.uint8)
    sup = np.zeros(39323, np.uint16)
    n = lib.orc_acc_cloud(a, 1, pos.ctypes.data, col.ctypes.data, sup.ctypes.data)
    assert n == 39323
    moved = np.flatnonzero(pos[:n, 0] == np.float32(1001.0))
    assert len(moved) == 1 and sup[moved[0]] == 1 and (col[moved[0]] == [10, 20, 30]).all()
    fresh = np.flatnonzero(pos[:n, 0] == np.float32(1000.5))
    assert len(fresh) == 1 and sup[fresh[0]] == 1 and (col[fresh[0]] == [40, 50, 60]).all()
    lib.orc_acc_destroy(a)
