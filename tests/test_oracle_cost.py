"""Pins of the oracle's cost function: the bilateral window (PM:85-110), the plane homography
(PM:112-128), the bilateral NCC of one view (PM:130-159) and the aggregated hypothesis cost
(PM:192-233) — by special cases with known answers and by an FP64 brute force in the reference's
own units (r8unorm values, sigma_colour 0.06, variance floor 1e-5)."""
import math
import os

import numpy as np
import pytest

import oracle as O

W, H = 64, 48
K = (60.0, 60.0, 31.5, 23.5)


class V:
    """A camera for the oracle: width/height/intrinsics/pose."""

    def __init__(self, R=np.eye(3), t=np.zeros(3), w=W, h=H):
        self.width, self.height = w, h
        self.intrinsics = type("K", (), dict(fx=K[0], fy=K[1], cx=K[2], cy=K[3]))
        self.pose = type("P", (), dict(R=np.asarray(R, float), t=np.asarray(t, float)))


def texture(seed=0, w=W, h=H):
    rng = np.random.default_rng(seed)
    coarse = rng.uniform(0, 255, (h // 4 + 2, w // 4 + 2))
    ys, xs = np.mgrid[0:h, 0:w] / 4.0
    y0, x0 = np.floor(ys).astype(int), np.floor(xs).astype(int)
    fy, fx = ys - y0, xs - x0
    img = (coarse[y0, x0] * (1 - fx) * (1 - fy) + coarse[y0, x0 + 1] * fx * (1 - fy) +
           coarse[y0 + 1, x0] * (1 - fx) * fy + coarse[y0 + 1, x0 + 1] * fx * fy)
    return np.clip(np.round(img), 0, 255).astype(np.uint8)


def rot(axis, deg):
    a = np.radians(deg)
    c, s = np.cos(a), np.sin(a)
    if axis == "y":
        return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def params(views, sources, topk=0, lo=0.5, hi=3.0):
    return O.make_params(views, 0, sources, lo, hi, topk=topk)


def ray(x, y):
    return np.array([(x - K[2]) / K[0], (y - K[3]) / K[1], 1.0])


def test_window_weights_match_fp64_formula():
    img = texture(1)
    views = [V(), V()]
    p = params(views, [1])
    imgs = O.Images([img, img])
    for (x, y) in [(20, 20), (0, 0), (63, 47), (33, 5)]:
        values, weights, stats = O.window(p, imgs, x, y)
        c = float(img[y, x]) / 255
        w64, v64 = [], []
        for i in range(36):
            ox, oy = (i % 6) * 2 - 5, (i // 6) * 2 - 5
            v = float(img[min(max(y + oy, 0), H - 1), min(max(x + ox, 0), W - 1)]) / 255
            w64.append(math.exp(-(v - c) ** 2 / (2 * 0.06 ** 2) - (ox * ox + oy * oy) / (2 * 25.0)))
            v64.append(v)
        w64, v64 = np.array(w64), np.array(v64)
        assert np.allclose(weights, w64, rtol=2e-6, atol=1e-30)
        assert np.allclose(values / 255, v64, rtol=0, atol=1e-7)
        sw = w64.sum()
        mean = (w64 * v64).sum() / sw
        var = max((w64 * v64 * v64).sum() / sw - mean * mean, 0)
        assert stats[0] == pytest.approx(sw, rel=1e-5)
        assert stats[1] / 255 == pytest.approx(mean, rel=1e-5)
        assert stats[2] / 255 ** 2 == pytest.approx(var, rel=1e-3, abs=1e-9)


def test_homography_maps_points_of_the_plane():
    views = [V(), V(rot("y", 8) @ rot("x", -3), [0.12, -0.02, 0.03])]
    p = params(views, [1])
    n = np.array([0.3, -0.4, -0.866])
    n /= np.linalg.norm(n)
    d = 1.4
    X = ray(30, 22) * d
    rho = float(n @ X)
    Hm = O.homography(p, 0, n.astype(np.float32), np.float32(rho)).astype(np.float64).reshape(3, 3)
    src = p.src[0]
    R = np.array(list(src.R)).reshape(3, 3)
    t = np.array(list(src.t))
    for (u, v) in [(25, 17), (35, 27), (30, 22), (10, 40)]:
        Xp = ray(u, v) * (rho / (n @ ray(u, v)))
        Xs = R @ Xp + t
        truth = np.array([src.k[0] * Xs[0] / Xs[2] + src.k[2], src.k[1] * Xs[1] / Xs[2] + src.k[3]])
        q = Hm @ np.array([u, v, 1.0])
        assert np.allclose(q[:2] / q[2], truth, atol=2e-3)


def _view_cost(img_ref, img_src, x, y, views=None, n=(0.0, 0.0, -1.0), d=1.0):
    views = views or [V(), V()]
    p = params(views, [1])
    imgs = O.Images([img_ref, img_src])
    win = O.window(p, imgs, x, y)
    X = ray(x, y) * d
    nn = np.array(n, dtype=np.float32)
    Hm = O.homography(p, 0, nn, np.float32(float(nn @ X)))
    return O.view_cost(p, imgs, 0, Hm, x, y, win)


# FP32 bound: 1 - cov/den cancels to ~1e-5 when the patches are (almost) identical.
FP32_NCC = 1e-4


def test_identical_views_cost_zero():
    img = texture(2)
    for x, y in [(20, 20), (40, 30), (10, 10)]:
        assert _view_cost(img, img, x, y) < FP32_NCC


def test_affine_intensity_maps_give_zero_and_two():
    img = (texture(3) // 2 * 2).astype(np.uint8)
    brighter = (img // 2 + 40).astype(np.uint8)
    inverted = (255 - img).astype(np.uint8)
    for x, y in [(20, 20), (40, 30)]:
        assert _view_cost(img, brighter, x, y) < FP32_NCC
        assert _view_cost(img, inverted, x, y) > 2 - FP32_NCC


def test_constant_source_is_degenerate():
    img = texture(4)
    flat = np.full_like(img, 128)
    assert _view_cost(img, flat, 20, 20) == 1.0


def test_window_leaving_the_source_is_invalid():
    img = texture(5)
    # identity geometry: the 11x11 window of x = 3 reaches x = -2
    assert _view_cost(img, img, 3, 20) == 2.0
    assert _view_cost(img, img, 5, 20) < FP32_NCC
    assert _view_cost(img, img, 60, 20) == 2.0
    # a plane behind the source camera
    views = [V(), V(np.eye(3), [0.0, 0.0, -3.0])]
    assert _view_cost(img, img, 20, 20, views=views) == 2.0


def _bilinear64(img, u, v):
    h, w = img.shape
    x0, y0 = math.floor(u), math.floor(v)
    fx, fy = u - x0, v - y0
    x1, y1 = min(x0 + 1, w - 1), min(y0 + 1, h - 1)
    f = img.astype(np.float64) / 255
    top = f[y0, x0] * (1 - fx) + f[y0, x1] * fx
    bot = f[y1, x0] * (1 - fx) + f[y1, x1] * fx
    return top * (1 - fy) + bot * fy


def test_ncc_matches_fp64_brute_force_in_reference_units():
    rng = np.random.default_rng(11)
    ref, src = texture(6), texture(7)
    views = [V(), V(rot("y", 6), [0.1, 0.01, 0.02])]
    p = params(views, [1])
    imgs = O.Images([ref, src])
    checked = 0
    for _ in range(300):
        x, y = int(rng.integers(6, W - 6)), int(rng.integers(6, H - 6))
        n = np.array([rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3), -1.0])
        n /= np.linalg.norm(n)
        d = rng.uniform(0.8, 2.0)
        win = O.window(p, imgs, x, y)
        nn = n.astype(np.float32)
        rho = np.float32(float(nn.astype(np.float64) @ (ray(x, y) * d)))
        Hf = O.homography(p, 0, nn, rho)
        got = O.view_cost(p, imgs, 0, Hf, x, y, win)
        Hm = Hf.astype(np.float64).reshape(3, 3)
        values, weights, stats = win
        ss = sss = srs = 0.0
        sw = float(weights.astype(np.float64).sum())
        invalid = False
        for i in range(36):
            ox, oy = (i % 6) * 2 - 5, (i // 6) * 2 - 5
            q = Hm @ np.array([x + ox, y + oy, 1.0])
            if q[2] <= 1e-6:
                invalid = True
                break
            u, v = q[0] / q[2], q[1] / q[2]
            if not (0 <= u <= W - 1 and 0 <= v <= H - 1):
                invalid = True
                break
            s = _bilinear64(src, u, v)
            w = float(weights[i])
            r = float(values[i]) / 255
            ss += w * s
            sss += w * s * s
            srs += w * s * r
        if invalid:
            assert got == 2.0
            continue
        mean_r = float(stats[1]) / 255
        var_r = max(sum(float(weights[i]) * (float(values[i]) / 255) ** 2 for i in range(36)) / sw - mean_r ** 2, 0)
        mean = ss / sw
        var_s = max(sss / sw - mean * mean, 0)
        den = math.sqrt(var_r * var_s)
        want = 1.0 if den < 1e-5 else min(max(1 - (srs / sw - mean_r * mean) / den, 0), 2)
        assert abs(got - want) < 3e-4, (x, y, got, want)
        checked += 1
    assert checked > 200


@pytest.mark.parametrize("topk", [0, 1, 2, 3])
def test_hypothesis_cost_is_mean_of_k_smallest_views(topk):
    img = texture(8)
    srcs = [texture(9 + i) for i in range(4)]
    views = [V()] + [V(rot("y", a), [0.05 * a, 0.0, 0.0]) for a in (-6, -3, 3, 6)]
    p = params(views, [1, 2, 3, 4], topk=topk)
    imgs = O.Images([img] + srcs)
    x, y = 30, 24
    win = O.window(p, imgs, x, y)
    n = np.array([0.1, -0.05, -1.0], dtype=np.float32)
    n /= np.float32(np.linalg.norm(n))
    d = np.float32(1.3)
    X = ray(x, y).astype(np.float32) * d
    rho = np.float32(n[0] * X[0] + n[1] * X[1] + n[2] * X[2])
    costs = sorted(O.view_cost(p, imgs, s, O.homography(p, s, n, rho), x, y, win) for s in range(4))
    k = topk if topk else max(1, min(3, 4 - 1))
    total = np.float32(0)
    for c in costs[:k]:
        total = np.float32(total + np.float32(c))
    assert O.hypothesis_cost(p, imgs, x, y, win, n, float(d)) == pytest.approx(float(total / np.float32(k)), abs=0)


def test_hypothesis_validity_rules():
    img = texture(10)
    views = [V(), V(rot("y", 4), [0.05, 0, 0])]
    p = params(views, [1], lo=0.5, hi=3.0)
    imgs = O.Images([img, texture(11)])
    win = O.window(p, imgs, 30, 24)
    facing = [0.0, 0.0, -1.0]
    assert O.hypothesis_cost(p, imgs, 30, 24, win, facing, 0.0) == 2.0
    assert O.hypothesis_cost(p, imgs, 30, 24, win, facing, -1.0) == 2.0
    assert O.hypothesis_cost(p, imgs, 30, 24, win, facing, 0.4) == 2.0   # nearer than min depth
    assert O.hypothesis_cost(p, imgs, 30, 24, win, facing, 3.5) == 2.0   # beyond max depth
    assert O.hypothesis_cost(p, imgs, 30, 24, win, [0.0, 0.0, 1.0], 1.0) == 2.0  # faces away
    assert O.hypothesis_cost(p, imgs, 30, 24, win, facing, 1.0) < 2.0


def test_degenerate_threshold_on_the_square_is_exact():
    """The kernels test den < 1e-5 * 255^2 (den = sqrtf(var_r * var_s)) as var_r * var_s < T before
    taking the square root; T must be the smallest float whose IEEE square root reaches the floor,
    so that the two tests agree on every float."""
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                        "luvas.io-3d_amd", "csrc", "pm_device.h")
    text = open(path).read()
    T = np.float32(float.fromhex(re.search(r"kDegenerateSq = (0x[0-9a-fp.+-]+)f;", text).group(1)))
    floor = np.float32(1e-5 * 255.0 * 255.0)
    below = np.nextafter(T, np.float32(0))
    assert np.sqrt(T) >= floor and np.sqrt(below) < floor
    # and on a band of floats around it, the two tests agree
    bits = T.view(np.uint32) + np.arange(-4096, 4096, dtype=np.int64)
    xs = bits.astype(np.uint32).view(np.float32)
    assert np.array_equal(np.sqrt(xs) < floor, xs < T)
