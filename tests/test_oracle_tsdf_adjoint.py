"""Pins of the FP64 TSDF integration adjoint (oracle/tsdf_adjoint_oracle.c).

Closed forms on a fronto-parallel view (one update gives tsdf = o, so dL/dd = 2 (o - t) / mu; two
updates give the weighted mean), the weight cap (96) against central differences, and central
differences of the depth used at single pixels and of 24 pose entries on the sphere workload.
The loss is piecewise smooth: the nearest-pixel choice and the skips (cam.z <= 0.01,
sdf < -mu) make it jump, the clamps (o <= 1, weight >= 0.2, the confidence ratio) make kinks.  A
difference quotient is used only where steps h and h/2 agree (no jump or kink inside the step)."""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes

H_STEP = 1e-8


def fd(f, h=H_STEP):
    g1 = (f(h) - f(-h)) / (2 * h)
    g2 = (f(h / 2) - f(-h / 2)) / h
    if not (np.isfinite(g1) and np.isfinite(g2)) or abs(g1 - g2) > 1e-5 * max(1.0, abs(g1)):
        return np.nan
    return g1


# ---- fronto-parallel closed forms ---------------------------------------------------------------

F, W, H = 100.0, 80, 60


def fronto_grid(voxel=0.01, trunc=4.0):
    coords = np.array([(bx, by, bz, 0) for bz in range(4) for by in range(2) for bx in range(2)], dtype=np.int32)
    return O.Grid((-0.08, -0.08, 0.5), voxel, (2, 2, 4), coords.reshape(-1), voxel * trunc)


def fronto_view(conf=0.0, t=(0.0, 0.0, 0.0)):
    return O.integration_view(W, H, F, F, (W - 1) / 2, (H - 1) / 2, np.eye(3), np.array(t), 5.0, conf)


def flat_input(depth, normal=(0.0, 0.0, -1.0), cost=0, prior=None):
    d = np.full(W * H, depth, dtype=np.float64)
    n = np.tile(np.asarray(normal, dtype=np.float64), (W * H, 1))
    c = np.full(W * H, cost, dtype=np.uint8)
    return (d, n, c), prior


def centre_voxel(g):
    # brick (1, 1, 1), local (0, 0, 2): world (0.005, 0.005, 0.605)
    b = 1 * 4 + 1 * 2 + 1
    return b * 512 + 0 + 8 * 0 + 64 * 2, np.array([0.005, 0.005, 0.605])


def test_one_update_closed_form():
    g = fronto_grid()
    vox, world = centre_voxel(g)
    (inp, _) = flat_input(0.62)
    r = O.tsdf_backward64(g, [fronto_view()], [inp], None, [vox], [0.1])
    mu = 0.04
    o = (0.62 - world[2]) / mu
    assert r["tsdf"][0] == pytest.approx(o, rel=1e-12)
    assert r["loss"] == pytest.approx((o - 0.1) ** 2, rel=1e-12)
    gd = r["grad_depth"][0]
    assert np.count_nonzero(gd) == 1 and gd.sum() == pytest.approx(2 * (o - 0.1) / mu, rel=1e-12)
    # dL/dt_z = dL/dcam_z = -2 (o - t) / mu (the weight, |n . c|, also depends on cam but a single
    # update's tsdf does not depend on its weight)
    assert r["grad_pose"][0][11] == pytest.approx(-2 * (o - 0.1) / mu, rel=1e-9)


def test_two_updates_weighted_mean_closed_form():
    g = fronto_grid()
    vox, world = centre_voxel(g)
    mu = 0.04
    inp1, _ = flat_input(0.62)
    inp2, _ = flat_input(0.63, cost=0)
    v1, v2 = fronto_view(), fronto_view(conf=0.63 * 1.5)  # w2 = 1 * 1.5^2 = 2.25 (|n . c| ~ 1)
    r = O.tsdf_backward64(g, [v1, v2], [inp1, inp2], None, [vox], [0.0])
    c = world / np.linalg.norm(world)
    w1 = abs(c[2])
    w2 = w1 * 1.5 ** 2
    o1, o2 = (0.62 - world[2]) / mu, (0.63 - world[2]) / mu
    ts = (o1 * w1 + o2 * w2) / (w1 + w2)
    assert r["tsdf"][0] == pytest.approx(ts, rel=1e-12)
    # d tsdf / d d1 = w1 / (w1 + w2) / mu; d tsdf / d d2 = (w2 / mu + (o2 - ts) dw2/dd2) / (w1 + w2)
    dw2 = w1 * 2 * 1.5 * (-0.63 * 1.5 / 0.63 ** 2)
    g1 = 2 * ts * w1 / (w1 + w2) / mu
    g2 = 2 * ts * (w2 / mu + (o2 - ts) * dw2) / (w1 + w2)
    assert r["grad_depth"][0].sum() == pytest.approx(g1, rel=1e-10)
    assert r["grad_depth"][1].sum() == pytest.approx(g2, rel=1e-10)


def test_weight_cap_against_central_differences():
    """40 updates of weight ~3.6 each: the weight reaches the cap (96) after ~27 of them."""
    g = fronto_grid()
    vox, world = centre_voxel(g)
    rng = np.random.default_rng(7)
    depths = 0.6 + 0.03 * rng.random(40)
    views = [fronto_view(conf=d * 1.9) for d in depths]
    inputs = [flat_input(d)[0] for d in depths]
    r = O.tsdf_backward64(g, views, inputs, None, [vox], [0.2])
    checked = 0
    for k in (0, 5, 20, 26, 27, 30, 39):
        def loss(h, k=k):
            inp = list(inputs)
            d, n, c = inp[k]
            inp[k] = (d + h, n, c)
            return O.tsdf_backward64(g, views, inp, None, [vox], [0.2])["loss"]
        f = fd(loss, 1e-7)
        if np.isfinite(f):
            assert f == pytest.approx(r["grad_depth"][k].sum(), rel=1e-5, abs=1e-9)
            checked += 1
    assert checked >= 6
    # every update before the cap moves the result; the contribution decays but stays nonzero
    assert all(abs(r["grad_depth"][k].sum()) > 0 for k in range(40))


# ---- sphere workload ----------------------------------------------------------------------------

def camera_normals(depth, k):
    Hh, Ww = depth.shape
    u, v = np.meshgrid(np.arange(Ww), np.arange(Hh))
    P = np.stack([(u - k.cx) / k.fx * depth, (v - k.cy) / k.fy * depth, depth], axis=-1)
    du = np.zeros_like(P)
    dv = np.zeros_like(P)
    du[:, 1:-1] = P[:, 2:] - P[:, :-2]
    dv[1:-1] = P[2:] - P[:-2]
    n = np.cross(du, dv)
    norm = np.linalg.norm(n, axis=-1, keepdims=True)
    n = np.where(norm > 0, n / np.maximum(norm, 1e-30), np.array([0.0, 0.0, -1.0]))
    flip = (n * P).sum(-1) > 0
    n[flip] = -n[flip]
    return n.reshape(-1, 3)


@pytest.fixture(scope="module")
def sphere():
    w = scenes.sphere_small().render([0, 1, 2, 3])
    spec = w.tsdf
    nb = spec.voxels // 8
    coords = np.array([(i % nb, (i // nb) % nb, i // (nb * nb), 0) for i in range(nb ** 3)], dtype=np.int32)
    mu = spec.voxel * spec.truncation_voxels
    g = O.Grid(spec.origin, spec.voxel, (nb, nb, nb), coords.reshape(-1), mu)
    rng = np.random.default_rng(11)
    views, inputs, priors = [], [], []
    for i in range(4):
        v = w.views[i]
        k = v.intrinsics
        views.append(O.integration_view(v.width, v.height, k.fx, k.fy, k.cx, k.cy, v.pose.R, v.pose.t,
                                        w.ranges[i][1], 1.0 if i >= 2 else 0.0))
        depth = w.truth[i].astype(np.float64)
        n = camera_normals(depth, k)
        d = depth.reshape(-1).copy()
        holes = rng.random(d.size) < 0.2
        d[holes] = 0.0                     # these pixels use the prior
        cost = rng.integers(0, 8, d.size).astype(np.uint8)
        inputs.append((d, n, cost))
        priors.append(depth.reshape(-1) * (1 + 0.01 * rng.standard_normal(d.size)))
    # seeded voxels near the surface
    cand = rng.integers(0, g.n_bricks * 512, 400000).astype(np.uint32)
    b, l = cand // 512, cand % 512
    c = coords.reshape(-1, 4)[b, :3] * 8 + np.stack([l % 8, (l // 8) % 8, l // 64], axis=1)
    world = np.asarray(spec.origin) + (c + 0.5) * spec.voxel
    sd = w.scene.signed_distance(world)
    near = np.flatnonzero(np.abs(sd) < 0.8 * mu)[:300]
    return g, views, inputs, priors, cand[near], np.clip(sd[near] / mu, -1, 1)


def test_sphere_depth_gradients_against_central_differences(sphere):
    g, views, inputs, priors, vox, tgt = sphere
    r = O.tsdf_backward64(g, views, inputs, priors, vox, tgt)
    assert r["loss"] > 0
    flat = np.concatenate(r["grad_depth"])
    top = np.argsort(-np.abs(flat))[:16]
    sizes = np.cumsum([v.width * v.height for v in views])
    checked = 0
    for j in top:
        vi = int(np.searchsorted(sizes, j, side="right"))
        p = int(j - (sizes[vi - 1] if vi else 0))

        def loss(h, vi=vi, p=p):
            inp = list(inputs)
            pri = list(priors)
            d, n, c = inp[vi]
            if d[p] > 0:
                d = d.copy()
                d[p] += h
                inp[vi] = (d, n, c)
            else:
                q = pri[vi].copy()
                q[p] += h
                pri[vi] = q
            return O.tsdf_backward64(g, views, inp, pri, vox, tgt)["loss"]
        f = fd(loss)
        if np.isfinite(f):
            assert f == pytest.approx(flat[j], rel=1e-5, abs=1e-7), (vi, p)
            checked += 1
    assert checked >= 12


def test_sphere_pose_gradients_against_central_differences(sphere):
    g, views, inputs, priors, vox, tgt = sphere
    r = O.tsdf_backward64(g, views, inputs, priors, vox, tgt)
    checked = 0
    for vi in (1, 2):
        for k in range(12):
            def loss(h, vi=vi, k=k):
                vs = list(views)
                v = O.integration_view(vs[vi].width, vs[vi].height, vs[vi].fx, vs[vi].fy, vs[vi].cx, vs[vi].cy,
                                       list(vs[vi].R), list(vs[vi].t), vs[vi].max_depth, vs[vi].confidence)
                if k < 9:
                    v.R[k] += h
                else:
                    v.t[k - 9] += h
                vs[vi] = v
                return O.tsdf_backward64(g, vs, inputs, priors, vox, tgt)["loss"]
            f = fd(loss)
            if np.isfinite(f):
                gk = r["grad_pose"][vi][k]
                assert f == pytest.approx(gk, rel=1e-5, abs=1e-6 * max(1.0, np.abs(r["grad_pose"][vi]).max())), (vi, k)
                checked += 1
    assert checked >= 20


def test_parallel_loss_equals_the_replay_loss(sphere):
    """orc64_tsdf_loss (OpenMP over voxels, for difference quotients over large voxel lists) is the
    forward of orc64_tsdf_backward: the same per-voxel terms, a total that adds them in voxel order
    whatever the thread count, with the culling flags computed inside or passed in."""
    g, views, inputs, priors, vox, tgt = sphere
    r = O.tsdf_backward64(g, views, inputs, priors, vox, tgt)
    direct = (r["tsdf"] - tgt) ** 2
    a = O.tsdf_loss64(g, views, inputs, priors, vox, tgt)
    b = O.tsdf_loss64(g, views, inputs, priors, vox, tgt, vis=O.visibility_flags(g, views))
    terms = O.tsdf_loss64(g, views, inputs, priors, vox, tgt, per_voxel=True)
    assert np.allclose(terms, direct, rtol=1e-12, atol=1e-15)
    assert a == pytest.approx(r["loss"], rel=1e-12) and a == b
    total = 0.0
    for t in terms:
        total += t
    assert total == a
    old = O.get_threads()
    try:
        O.set_threads(1)
        one = O.tsdf_loss64(g, views, inputs, priors, vox, tgt)
        O.set_threads(3)
        three = O.tsdf_loss64(g, views, inputs, priors, vox, tgt)
    finally:
        O.set_threads(old)
    assert one == three == a
    # a voxel list in another order gives the same loss up to summation order
    perm = np.random.default_rng(3).permutation(len(vox))
    assert O.tsdf_loss64(g, views, inputs, priors, vox[perm], tgt[perm]) == pytest.approx(a, rel=1e-12)
