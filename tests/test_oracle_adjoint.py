"""The FP64 reference gradient of the aggregated NCC cost (oracle/adjoint_oracle.c) against FP64
central finite differences in depth, normal and relative pose, on seeded pixels of the sphere
workload.  The cost is piecewise smooth (bilinear kinks at texel edges, top-k switches) and
strongly curved: a step of 1e-4 m moves the taps by ~0.03 px, so its difference quotient is off
by up to ~0.5 % from curvature and edge crossings alone.  The check uses h = 1e-6 (taps move
~3e-4 px; FP64 keeps ~10 digits of the quotient) and keeps a sample only where the quotients with
h and h/2 agree (no kink inside the step) and the selected views do not change."""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes

H_STEP = 1e-6


@pytest.fixture(scope="module")
def setup():
    w = scenes.sphere_small().render([0, 1, 7, 2, 6])
    imgs = O.Images([im if im is not None else np.zeros((240, 320), np.uint8) for im in w.images])
    truth = w.truth[0]
    return w, imgs, truth


def plane_at(w, truth, x, y):
    """True depth; the sphere / ground normal from a finite-difference of the truth map."""
    k = w.views[0].intrinsics
    def pt(u, v):
        d = truth[v, u]
        return np.array([(u - k.cx) / k.fx * d, (v - k.cy) / k.fy * d, d])
    n = np.cross(pt(x + 1, y) - pt(x - 1, y), pt(x, y + 1) - pt(x, y - 1))
    n /= np.linalg.norm(n)
    if n[2] > 0:
        n = -n
    return n, float(truth[y, x])


def fd(f, x0, h):
    """Central difference with step h, or nan where the step crosses a kink (h vs h/2 disagree)."""
    g1 = (f(x0 + h) - f(x0 - h)) / (2 * h)
    g2 = (f(x0 + h / 2) - f(x0 - h / 2)) / h
    if not (np.isfinite(g1) and np.isfinite(g2)) or abs(g1 - g2) > 1e-6 * max(1.0, abs(g1)):
        return np.nan
    return g1


def test_fp64_gradient_matches_central_differences(setup):
    w, imgs, truth = setup
    rng = np.random.default_rng(5)
    p = O.params64(w.views, 0, w.sources[0], *w.ranges[0], topk=w.topk)
    checked = {"depth": 0, "normal": 0, "R": 0, "t": 0}
    for _ in range(400):
        x, y = int(rng.integers(20, 300)), int(rng.integers(20, 220))
        if truth[y, x] <= 0 or (truth[y - 1:y + 2, x - 1:x + 2] <= 0).any():
            continue
        n, d = plane_at(w, truth, x, y)
        d *= 1 + 0.002 * rng.standard_normal()
        r = O.cost_grad64(p, imgs, 0, x, y, n, d)
        if not (0 < r.cost < 1.5):
            continue
        sel = list(r.selected)

        def cost(n_, d_, p_=p):
            out = O.cost_grad64(p_, imgs, 0, x, y, n_, d_)
            return out.cost if list(out.selected) == sel else np.nan

        g = fd(lambda v: cost(n, v), d, H_STEP)
        if np.isfinite(g):
            assert abs(g - r.d_depth) <= 1e-5 * max(1.0, abs(r.d_depth)), (x, y, g, r.d_depth)
            checked["depth"] += 1
        for k in range(3):
            e = np.eye(3)[k]
            g = fd(lambda v: cost(n + e * (v - 0.0), d), 0.0, H_STEP)
            if np.isfinite(g):
                assert abs(g - r.d_normal[k]) <= 1e-5 * max(1.0, abs(r.d_normal[k])), (x, y, k, g, r.d_normal[k])
                checked["normal"] += 1
        s = int(np.argmax(sel))
        base = [(np.array(list(p.src[j].R)).reshape(3, 3), np.array(list(p.src[j].t))) for j in range(p.n_src)]
        for k in range(12):
            def cost_pose(v):
                rel = [(R.copy(), t.copy()) for R, t in base]
                if k < 9:
                    rel[s][0].reshape(-1)[k] += v
                else:
                    rel[s][1][k - 9] += v
                return cost(n, d, O.params64(w.views, 0, w.sources[0], *w.ranges[0], topk=w.topk, rel=rel))
            g = fd(cost_pose, 0.0, H_STEP)
            want = r.d_R[s][k] if k < 9 else r.d_t[s][k - 9]
            if np.isfinite(g):
                assert abs(g - want) <= 1e-5 * max(1.0, abs(want)), (x, y, k, g, want)
                checked["R" if k < 9 else "t"] += 1
    assert checked["depth"] > 20 and checked["normal"] > 60 and checked["R"] > 150 and checked["t"] > 50, checked


def test_world_pose_chain_rule_matches_finite_differences(setup):
    """lr3d.refine.world_pose_gradients (relative-pose gradient -> 6-DoF world-pose gradients of
    the reference and source cameras) against FP64 differences of perturbed world poses."""
    from lr3d.refine import perturb, world_pose_gradients
    w, imgs, truth = setup
    rng = np.random.default_rng(9)
    pixels = []
    while len(pixels) < 24:
        x, y = int(rng.integers(30, 290)), int(rng.integers(30, 210))
        if truth[y, x] > 0 and (truth[y - 1:y + 2, x - 1:x + 2] > 0).all():
            pixels.append((x, y) + plane_at(w, truth, x, y))
    views = w.views
    srcs = w.sources[0]

    def loss(vs):
        p = O.params64(vs, 0, srcs, *w.ranges[0], topk=w.topk)
        total, sel = 0.0, []
        for x, y, n, d in pixels:
            r = O.cost_grad64(p, imgs, 0, x, y, n, d)
            total += r.cost
            sel.append(tuple(r.selected))
        return total, sel

    p = O.params64(views, 0, srcs, *w.ranges[0], topk=w.topk)
    grad_rel = np.zeros((len(srcs), 12))
    for x, y, n, d in pixels:
        r = O.cost_grad64(p, imgs, 0, x, y, n, d)
        for s in range(len(srcs)):
            grad_rel[s, :9] += list(r.d_R[s])
            grad_rel[s, 9:] += list(r.d_t[s])
    grads = world_pose_gradients(views, 0, srcs, grad_rel)
    base, sel0 = loss(views)
    h = 1e-7
    checked = 0
    for layer in (0, srcs[0], srcs[1]):
        for k in range(6):
            e = np.zeros(6)
            e[k] = h
            vals = []
            for sign in (1, -1):
                vs = [type(v)(v.intrinsics, v.pose, v.width, v.height) for v in views]
                vs[layer] = type(views[layer])(views[layer].intrinsics,
                                               perturb(views[layer].pose, sign * e[:3], sign * e[3:]),
                                               views[layer].width, views[layer].height)
                val, sel = loss(vs)
                vals.append(val if sel == sel0 else np.nan)
            fd_g = (vals[0] - vals[1]) / (2 * h)
            want = grads[layer][0][k] if k < 3 else grads[layer][1][k - 3]
            if np.isfinite(fd_g):
                assert abs(fd_g - want) <= 2e-3 * max(1.0, abs(want)), (layer, k, fd_g, want)
                checked += 1
    assert checked >= 12
