"""Pins of the oracle's forward-backward reprojection (PM:166-190), its consistency filter
(PM:436-469), the geometric term of the hypothesis cost (PM:207-215) and the coarse-to-fine start
(PM:296-333), against closed forms and an FP64 model of the geometry.

* A rig of three parallel cameras 5 cm apart (f = 200 px) sees a fronto-parallel plane at 2 m: every
  depth is exactly 2, the disparity is exactly 5 px, so the filter's kept/rejected columns, its
  mean depth and its thresholds (1 % depth, cost code 7, `needed` clamped to [1, n]) have closed
  forms.
* Two cameras at general poses, of different sizes and with fx != fy, see a tilted plane whose
  packed map holds the FP64 ray-plane depths: the oracle's reprojection must agree with an FP64
  evaluation of the same construction (nearest source pixel, its f16 depth, back into the
  reference) within 1e-3 px and 1e-5 in depth, pixel by pixel.
* The geometric term adds 0.2 * min(distance, 3) per view (0.6 when the reprojection fails).
* init_coarse with an empty coarse level is init; with the true ground plane on the coarse level
  it transfers depths that equal the ray-plane intersection along each fine ray; a plane on one
  coarse corner pixel reaches exactly the fine pixels whose clamped 2 x 2 neighbourhood holds it.
"""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes
from lr3d.camera import Intrinsics, Pose, StereoView, look_at, scale_intrinsics

UP = np.array([0.0, 1.0, 0.0])


def rig(b=0.05, W=64, H=48, f=200.0):
    """Three parallel cameras with centres at x = -b, 0, +b, looking down +z."""
    k = Intrinsics(f, f, (W - 1) / 2, (H - 1) / 2)
    return [StereoView(k, Pose(np.eye(3), -np.array([c, 0.0, 0.0])), W, H) for c in (-b, 0.0, b)]


def packed_maps(views, depth_scale=(1.0, 1.0, 1.0), cost=(0.0, 0.0, 0.0), zero=()):
    """Map pool of the rig looking at the plane z = 2: depth 2 * scale, normal (0, 0, -1)."""
    pool = O.map_pool(views)
    for l, v in enumerate(views):
        word = O.pack_sample(2.0 * depth_scale[l], (0.0, 0.0, -1.0), cost[l])
        m = O.layer_map(views, pool, l)
        m[:] = word
        if l in zero:
            m[:] = 0
    return pool


def filtered(views, pool, checks, needed, ref=1):
    return O.filter_map(views, ref, checks, needed, pool).reshape(views[ref].height, views[ref].width)


def f16(x):
    return float(np.float16(np.float32(x)))


def test_rig_plane_keeps_exactly_the_columns_both_sources_see():
    views = rig()
    pool = packed_maps(views)
    word = int(O.layer_map(views, pool, 1)[0])
    assert O.unpack_depth(word) == 2.0
    out = filtered(views, pool, [0, 2], 2)
    # source 0 (centre -b) sees reference column u at u + 5, source 2 at u - 5 (f b / z = 5 px)
    cols = np.arange(64)
    both = (cols >= 5) & (cols <= 58)
    assert np.all(out[:, both] == word), "kept depths: mean of three equal depths, same packed word"
    assert np.all(out[:, ~both] == 0)
    # needed = 1: every column has at least one agreeing source; needed is clamped to [1, n]
    assert np.all(filtered(views, pool, [0, 2], 1) == word)
    assert np.array_equal(filtered(views, pool, [0, 2], 0), filtered(views, pool, [0, 2], 1))
    assert np.array_equal(filtered(views, pool, [0, 2], 5), out)
    one = filtered(views, pool, [0], 3)  # needed 3 of one source -> 1
    assert np.all(one[:, cols <= 58] == word) and np.all(one[:, cols > 58] == 0)


@pytest.mark.parametrize("scale,agrees", [(1.008, True), (1.005, True), (0.992, True), (1.012, False),
                                          (0.988, False), (1.02, False)])
def test_depth_threshold_is_one_percent_of_the_reference_depth(scale, agrees):
    views = rig()
    pool = packed_maps(views, depth_scale=(scale, 1.0, 1.0))
    ds = O.unpack_depth(int(O.layer_map(views, pool, 0)[0]))
    assert (abs(ds - 2.0) < 0.01 * 2.0) == agrees  # the f16 depth keeps the side of the threshold
    out = filtered(views, pool, [0, 2], 2)
    inner = out[:, 5:59]
    if agrees:
        mean = np.float32(np.float32(np.float32(2.0) + np.float32(ds)) + np.float32(2.0)) / np.float32(3.0)
        assert np.all((inner & 0xFFFF) == np.uint32(np.float16(mean).view(np.uint16)))
        assert np.all((inner >> 16) == (int(O.layer_map(views, pool, 1)[0]) >> 16))
    else:
        assert np.all(out == 0)
        # with needed = 1 the other source keeps the depth where it sees it, at the mean of two
        one = filtered(views, pool, [0, 2], 1)
        assert np.all(one[:, 5:] == int(O.layer_map(views, pool, 1)[0])) and np.all(one[:, :5] == 0)


def test_cost_code_and_invalid_depths():
    views = rig()
    # cost 0.5 -> code 7 (kept), 0.55 -> code 8 (rejected), on the source and on the reference
    for c_src, ok in ((0.5, True), (0.55, False)):
        pool = packed_maps(views, cost=(c_src, 0.0, 0.0))
        assert O.unpack_cost(int(O.layer_map(views, pool, 0)[0])) == (7 if ok else 8)
        out = filtered(views, pool, [0, 2], 2)
        assert np.all(out[:, 5:59] != 0) == ok and np.all(out[:, 5:59] == 0) == (not ok)
    for c_ref, ok in ((0.5, True), (0.55, False)):
        pool = packed_maps(views, cost=(0.0, c_ref, 0.0))
        assert (np.count_nonzero(filtered(views, pool, [0, 2], 2)) > 0) == ok
    assert np.all(filtered(views, packed_maps(views, zero=(1,)), [0, 2], 1) == 0)
    assert np.all(filtered(views, packed_maps(views, zero=(0, 2)), [0, 2], 1) == 0)


def test_reprojection_closed_forms():
    views = rig()
    pool = packed_maps(views)
    p = O.make_params(views, 1, [0, 2], 0.01, 100.0)
    for x, y in ((10, 7), (31, 23), (50, 40)):
        ray = np.array([(x - 31.5) / 200.0, (y - 23.5) / 200.0, 1.0], np.float32)
        r = O.reproject(p, pool, 0, float(x), float(y), ray * np.float32(2.0))
        assert r[3] == 1.0 and r[1] == 2.0 and r[0] < 1e-4
    # behind the source camera: the source turned around (R = diag(-1, 1, -1))
    back = [views[0], views[1], StereoView(views[2].intrinsics, Pose(np.diag([-1.0, 1.0, -1.0]), [0, 0, 0]), 64, 48)]
    pb = O.make_params(back, 1, [2], 0.01, 100.0)
    assert np.all(O.reproject(pb, packed_maps(back), 0, 31.0, 23.0, np.array([0.0, 0.0, 2.0], np.float32)) == -1.0)
    # outside the source image, and onto a pixel without depth
    assert np.all(O.reproject(p, pool, 0, 60.0, 20.0, np.array([(60 - 31.5) / 100, 0, 2.0], np.float32)) == -1.0)
    empty = packed_maps(views, zero=(0,))
    assert np.all(O.reproject(p, empty, 0, 31.0, 23.0, np.array([0.0, 0.0, 2.0], np.float32)) == -1.0)


def general_pair():
    """Reference 96x72 and source 80x60 at general poses (fx != fy) and a tilted plane."""
    ref = StereoView(Intrinsics(110.0, 104.0, 47.3, 35.8), look_at((0.3, 0.9, -1.6), (0.05, -0.1, 0.0)), 96, 72)
    src = StereoView(Intrinsics(92.0, 97.0, 40.1, 29.6), look_at((-0.45, 1.05, -1.45), (0.0, -0.05, 0.1)), 80, 60)
    n = np.array([0.2, 1.0, 0.3])
    n /= np.linalg.norm(n)
    return [ref, src], n, -0.1  # plane n . X = off (world)


def plane_depth(view, n_w, off, u, v):
    """FP64 depth (camera z) of the plane along the ray of pixel (u, v); normal in camera axes."""
    k, R, t = view.intrinsics, view.pose.R, view.pose.t
    ray = np.stack([(u - k.cx) / k.fx, (v - k.cy) / k.fy, np.ones_like(u, dtype=np.float64)], -1)
    a = (R.T @ ray.reshape(-1, 3).T).T.reshape(ray.shape) @ n_w
    d = (off + n_w @ (R.T @ t)) / a
    return d, R @ n_w


def test_reprojection_matches_an_fp64_model_pixel_by_pixel():
    views, n_w, off = general_pair()
    ref, src = views
    pool = O.map_pool(views)
    vs, us = np.mgrid[0:src.height, 0:src.width].astype(np.float64)
    ds, nc = plane_depth(src, n_w, off, us, vs)
    nc = -nc if nc[2] > 0 else nc
    m = O.layer_map(views, pool, 1)
    m[:] = [O.pack_sample(float(d), tuple(nc), 0.1) for d in ds.reshape(-1)]
    p = O.make_params(views, 0, [1], 0.01, 100.0)
    Rs = np.array(p.src[0].R[:], np.float64).reshape(3, 3)
    ts = np.array(p.src[0].t[:], np.float64)
    ks = np.array(p.src[0].k[:], np.float64)
    kr = np.array(p.k[:], np.float64)
    src_depth = np.array([O.unpack_depth(int(w)) for w in m], np.float64).reshape(src.height, src.width)
    checked = valid = 0
    for y in range(0, ref.height, 3):
        for x in range(0, ref.width, 3):
            d, _ = plane_depth(ref, n_w, off, np.float64(x), np.float64(y))
            ray = np.array([(np.float32(x) - np.float32(kr[2])) / np.float32(kr[0]),
                            (np.float32(y) - np.float32(kr[3])) / np.float32(kr[1]), 1.0], np.float32)
            X = ray * np.float32(d)
            got = O.reproject(p, pool, 0, float(x), float(y), X)
            # FP64 model of the same construction
            xs = Rs @ X.astype(np.float64) + ts
            proj = np.array([ks[0] * xs[0] / xs[2] + ks[2], ks[1] * xs[1] / xs[2] + ks[3]])
            if np.any(np.abs(proj - np.floor(proj) - 0.5) < 1e-3):
                continue  # a rounding tie: FP32 and FP64 may pick different pixels
            px = np.rint(proj).astype(int)
            if xs[2] <= 1e-6 or px[0] < 0 or px[1] < 0 or px[0] >= src.width or px[1] >= src.height:
                assert np.all(got == -1.0), (x, y)
                checked += 1
                continue
            dsrc = src_depth[px[1], px[0]]
            back_s = np.array([(px[0] - ks[2]) / ks[0], (px[1] - ks[3]) / ks[1], 1.0]) * dsrc
            back = Rs.T @ (back_s - ts)
            q = np.array([kr[0] * back[0] / back[2] + kr[2], kr[1] * back[1] / back[2] + kr[3]])
            dist = np.hypot(q[0] - x, q[1] - y)
            assert got[3] == 1.0, (x, y)
            assert abs(got[0] - dist) < 1e-3, (x, y, got[0], dist)
            assert abs(got[1] - back[2]) < 1e-5 * back[2], (x, y, got[1], back[2])
            checked += 1
            valid += 1
    assert checked > 600 and valid > 300


def test_geometric_term_adds_a_capped_reprojection_distance():
    views, n_w, off = general_pair()
    ref, src = views
    rng = np.random.default_rng(5)
    images = O.Images([rng.integers(0, 256, (v.height, v.width), dtype=np.uint8) for v in views])
    vs, us = np.mgrid[0:src.height, 0:src.width].astype(np.float64)
    ds, nc_src = plane_depth(src, n_w, off, us, vs)
    nc_src = -nc_src if nc_src[2] > 0 else nc_src
    for scale, capped in ((1.0, False), (3.0, True)):
        pool = O.map_pool(views)
        O.layer_map(views, pool, 1)[:] = [O.pack_sample(float(d) * scale, tuple(nc_src), 0.1)
                                          for d in ds.reshape(-1)]
        p0 = O.make_params(views, 0, [1], 0.3, 6.0, flags=0)
        p1 = O.make_params(views, 0, [1], 0.3, 6.0, flags=1)
        checked = 0
        for y in range(2, ref.height - 2, 7):
            for x in range(2, ref.width - 2, 7):
                d, nr = plane_depth(ref, n_w, off, np.float64(x), np.float64(y))
                nr = -nr if nr[2] > 0 else nr
                kr = ref.intrinsics
                ray64 = np.array([(x - kr.cx) / kr.fx, (y - kr.cy) / kr.fy, 1.0])
                if not (0.31 < d < 5.9) or nr @ ray64 > -2e-3:
                    continue  # an invalid hypothesis returns INVALID before any view term (PM:193-200)
                win = O.window(p0, images, x, y)
                c0 = O.hypothesis_cost(p0, images, x, y, win, nr, float(d), depths=pool)
                c1 = O.hypothesis_cost(p1, images, x, y, win, nr, float(d), depths=pool)
                ray = np.array([(np.float32(x) - np.float32(p0.k[2])) / np.float32(p0.k[0]),
                                (np.float32(y) - np.float32(p0.k[3])) / np.float32(p0.k[1]), 1.0], np.float32)
                r = O.reproject(p1, pool, 0, float(x), float(y), ray * np.float32(d))
                e = min(float(r[0]), 3.0) if r[3] > 0 else 3.0
                assert abs((c1 - c0) - 0.2 * e) < 2e-6, (x, y, c0, c1, e)
                if capped:
                    assert e == 3.0
                checked += 1
        assert checked > 60


def ground_set(W=160, H=120, f=150.0, n=5, seed=11):
    """A reference view and n - 1 sources looking down at a textured ground plane y = -0.5 that
    fills every frame.  The sources sit a few centimetres beside the reference with a wider field
    of view (f * 0.6), so every window of the reference, corners included, lies inside them."""
    scene = scenes.SceneSpec([scenes.Primitive(scenes.SG_RECT_Y, (-5.0, -0.5, -5.0), (5.0, 0.0, 5.0))],
                             seed=seed, freq=0.6)
    views = []
    for i in range(n):
        fi = f if i == 0 else 0.6 * f
        k = Intrinsics(fi, fi, (W - 1) / 2, (H - 1) / 2)
        a = 0.0 if i == 0 else 0.05 * ((i + 1) // 2) * (1 if i % 2 else -1)
        # rolled (up vector off the vertical) so the ground normal has an x component in camera axes
        pose = look_at((0.8 * np.sin(a), 1.0, 0.8 * np.cos(a)), (0.0, -0.5, 0.0), up=(0.4, 1.0, 0.1))
        views.append(StereoView(k, pose, W, H))
    images = [scene.render(v)[0].reshape(H, W) for v in views]
    return views, images


def coarse_ground(view, cw, ch, only=None):
    """Coarse level (cw x ch) of `view` holding the true ground plane (n, depth) per pixel; only: a
    list of (x, y) coarse pixels to keep (the rest w = 0)."""
    ck = scale_intrinsics(view.intrinsics, cw / view.width, ch / view.height)
    cv = StereoView(ck, view.pose, cw, ch)
    vs, us = np.mgrid[0:ch, 0:cw].astype(np.float64)
    d, nc = plane_depth(cv, UP, -0.5, us, vs)
    nc = -nc if nc[2] > 0 else nc
    hyp = np.zeros((ch * cw, 4), np.float32)
    hyp[:, :3] = nc.astype(np.float32)
    hyp[:, 3] = d.reshape(-1).astype(np.float32)
    if only is not None:
        keep = np.zeros(ch * cw, bool)
        for (x, y) in only:
            keep[y * cw + x] = True
        hyp[~keep, 3] = 0.0
    ckf = [np.float32(ck.fx), np.float32(ck.fy), np.float32(ck.cx), np.float32(ck.cy)]
    return ckf, hyp, nc.astype(np.float32)


def test_init_coarse_with_an_empty_coarse_level_is_init():
    views, images = ground_set()
    imgs = O.Images(images)
    p = O.make_params(views, 0, [1, 2, 3, 4], 0.3, 4.0)
    ck, hyp, _ = coarse_ground(views[0], 80, 60, only=[])
    h1, c1 = O.init_coarse(p, imgs, ck, 80, 60, hyp)
    h0, c0 = O.init(p, imgs)
    assert np.array_equal(h1.view(np.uint32), h0.view(np.uint32))
    assert np.array_equal(c1.view(np.uint32), c0.view(np.uint32))


def test_init_coarse_transfers_the_true_plane_along_each_fine_ray():
    views, images = ground_set()
    imgs = O.Images(images)
    p = O.make_params(views, 0, [1, 2, 3, 4], 0.3, 4.0)
    ck, hyp, nc = coarse_ground(views[0], 80, 60)
    h, c = O.init_coarse(p, imgs, ck, 80, 60, hyp)
    took = np.all(h[:, :3].view(np.uint32) == nc.view(np.uint32), axis=1)
    vs, us = np.mgrid[0:120, 0:160].astype(np.float64)
    d, _ = plane_depth(views[0], UP, -0.5, us, vs)
    rel = np.abs(h[:, 3] - d.reshape(-1)) / d.reshape(-1)
    # the true plane wins wherever the sources see the pixel's window (the random plane keeps the
    # rest, e.g. windows that leave every source image); what it transfers is the exact depth
    assert took.mean() > 0.75, took.mean()
    assert rel[took].max() < 1e-5, rel[took].max()


@pytest.mark.parametrize("corner", ["top-left", "bottom-right"])
def test_init_coarse_reaches_the_clamped_two_by_two_neighbourhood(corner):
    views, images = ground_set()
    imgs = O.Images(images)
    p = O.make_params(views, 0, [1, 2, 3, 4], 0.3, 4.0)
    cw, ch = 80, 60
    cpix = (0, 0) if corner == "top-left" else (cw - 1, ch - 1)
    ck, hyp, nc = coarse_ground(views[0], cw, ch, only=[cpix])
    h, c = O.init_coarse(p, imgs, ck, cw, ch, hyp)
    h0, c0 = O.init(p, imgs)
    changed = np.any(h.view(np.uint32) != h0.view(np.uint32), axis=1).reshape(120, 160)
    # fine x reaches coarse column 0 iff floor((x + 0.5) / 2 - 0.5) <= 0, i.e. x <= 2; column
    # cw - 1 iff floor(...) + 1 >= cw - 1, i.e. x >= 2 cw - 3 (the same for rows)
    ys, xs = np.nonzero(changed)
    if corner == "top-left":
        reach = (xs <= 2) & (ys <= 2)
    else:
        reach = (xs >= 2 * cw - 3) & (ys >= 2 * ch - 3)
    assert reach.all() and changed.sum() >= 5, (changed.sum(), list(zip(xs, ys)))
    hc = h.reshape(120, 160, 4)[changed]
    assert np.all(hc[:, :3].view(np.uint32) == nc.view(np.uint32))
    assert np.all(c.reshape(120, 160)[changed] < c0.reshape(120, 160)[changed])
