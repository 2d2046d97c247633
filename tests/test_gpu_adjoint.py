"""GPU adjoint of the aggregated NCC cost (lr3d_ncc_backward / lr3d_ncc_refine_step) against the
FP64 CPU reference gradient (oracle/adjoint_oracle.c, itself pinned by central differences in
tests/test_oracle_adjoint.py) and against FP64 central differences of 16 world-pose parameters.

Bars: per pixel, |g_gpu - g_fp64| <= 1e-3 * max(1, |g_fp64|) for dL/dd and dL/dn on >= 99 % of the
1024 seeded pixels whose FP32 and FP64 costs agree (the top-k choice is the same); the summed
relative-pose gradient within 1e-3 relative; world-pose gradients within 2e-2 of FP64 central
differences (the summed loss crosses texel edges inside any step, see test_oracle_adjoint.py)."""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def solved(ctx):
    import lr3d
    w = scenes.sphere_small().render()
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    o = lr3d.SolveOptions(sources=w.sources[0], min_depth=w.ranges[0][0], max_depth=w.ranges[0][1],
                          iterations=w.iterations, topk=w.topk)
    st.solve(0, o)
    hyp, cost = st.hypotheses(0)
    yield w, st, o, hyp, cost
    st.destroy()


def seeded_pixels(cost, n=1024, seed=12):
    rng = np.random.default_rng(seed)
    ok = np.flatnonzero((cost > 0) & (cost < 1.0))
    return np.sort(rng.choice(ok, size=min(n, ok.size), replace=False))


def test_pixel_gradients_match_fp64_reference(solved):
    w, st, o, hyp, cost = solved
    say("case: NCC adjoint per pixel, view 0 of sphere 8x320x240, 1024 seeded pixels")
    out = st.ncc_backward(0, o)
    assert np.array_equal(out["cost"].view(np.uint32), cost.view(np.uint32))  # forward is exact
    imgs = O.Images(w.images)
    p = O.params64(w.views, 0, w.sources[0], *w.ranges[0], topk=w.topk)
    idx = seeded_pixels(cost)
    good = total = 0
    for i in idx:
        y, x = divmod(int(i), 320)
        h = hyp[i].astype(np.float64)
        r = O.cost_grad64(p, imgs, 0, x, y, h[:3], h[3])
        if abs(r.cost - cost[i]) > 1e-4:
            continue  # FP32 and FP64 chose different views
        total += 1
        ok = abs(out["grad_depth"][i] - r.d_depth) <= 1e-3 * max(1.0, abs(r.d_depth))
        for k in range(3):
            ok &= abs(out["grad_normal"][i, k] - r.d_normal[k]) <= 1e-3 * max(1.0, abs(r.d_normal[k]))
        good += ok
    say(f"adjoint: {good}/{total} pixels within 1e-3 of the FP64 reference")
    assert total >= 900 and good >= 0.99 * total


def test_pose_gradient_matches_fp64_reference(solved):
    w, st, o, hyp, cost = solved
    say("case: NCC adjoint, relative-pose gradient summed over 1024 pixels")
    idx = seeded_pixels(cost)
    weights = np.zeros(320 * 240, np.float32)
    weights[idx] = 1.0
    out = st.ncc_backward(0, o, weights)
    imgs = O.Images(w.images)
    p = O.params64(w.views, 0, w.sources[0], *w.ranges[0], topk=w.topk)
    ref = np.zeros((len(w.sources[0]), 12))
    loss = 0.0
    for i in idx:
        y, x = divmod(int(i), 320)
        h = hyp[i].astype(np.float64)
        r = O.cost_grad64(p, imgs, 0, x, y, h[:3], h[3])
        loss += r.cost
        for s in range(len(w.sources[0])):
            ref[s, :9] += list(r.d_R[s])
            ref[s, 9:] += list(r.d_t[s])
    err = np.abs(out["grad_pose"] - ref).max() / np.abs(ref).max()
    say(f"pose gradient: max relative error {err:.2e}, loss gpu {out['loss']:.6f} fp64 {loss:.6f}")
    assert err <= 1e-3
    assert abs(out["loss"] - loss) <= 1e-4 * loss


def test_world_pose_gradients_match_central_differences(solved):
    from lr3d.refine import perturb, world_pose_gradients
    w, st, o, hyp, cost = solved
    say("case: 16 world-pose parameters, FP64 central differences of the summed cost")
    idx = seeded_pixels(cost, n=256, seed=3)
    weights = np.zeros(320 * 240, np.float32)
    weights[idx] = 1.0
    out = st.ncc_backward(0, o, weights)
    srcs = w.sources[0]
    grads = world_pose_gradients(w.views, 0, srcs, out["grad_pose"])
    imgs = O.Images(w.images)
    planes = [(divmod(int(i), 320), hyp[i].astype(np.float64)) for i in idx]

    def loss(views):
        p = O.params64(views, 0, srcs, *w.ranges[0], topk=w.topk)
        return sum(O.cost_grad64(p, imgs, 0, x, y, h[:3], h[3]).cost for (y, x), h in planes)

    h = 1e-6
    params = [(srcs[0], k) for k in range(6)] + [(srcs[1], k) for k in range(6)] + [(0, k) for k in (0, 1, 3, 5)]
    errs = []
    for layer, k in params:
        e = np.zeros(6)
        e[k] = h
        vals = []
        for sign in (1, -1):
            views = list(w.views)
            v = views[layer]
            views[layer] = type(v)(v.intrinsics, perturb(v.pose, sign * e[:3], sign * e[3:]), v.width, v.height)
            vals.append(loss(views))
        fd = (vals[0] - vals[1]) / (2 * h)
        g = grads[layer][0][k] if k < 3 else grads[layer][1][k - 3]
        errs.append(abs(fd - g) / max(1.0, abs(g)))
    say(f"world-pose gradients vs FD: median relative error {np.median(errs):.2e}, max {max(errs):.2e}")
    assert len(params) == 16
    assert np.median(errs) <= 5e-3 and max(errs) <= 2e-2


def test_refinement_lowers_the_cost_with_noisy_poses(ctx):
    import lr3d
    from lr3d.refine import perturb, refine
    say("case: joint plane + pose refinement, 8 views, poses noised (0.5 deg, 5 mm), 20 steps")
    w = scenes.sphere_small().render()
    rng = np.random.default_rng(4)
    noisy = []
    for i, v in enumerate(w.views):
        pose = v.pose if i == 0 else perturb(v.pose, rng.normal(0, np.radians(0.5) / np.sqrt(3), 3),
                                             rng.normal(0, 0.005 / np.sqrt(3), 3))
        noisy.append(lr3d.StereoView(v.intrinsics, pose, v.width, v.height))
    st = lr3d.HipDenseStereo(ctx, noisy)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    solves = [(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                                    iterations=w.iterations, topk=w.topk)) for r in range(8)]
    st.solve_batched(solves)

    def pose_err():
        return np.mean([np.linalg.norm(st.views[i].pose.center - w.views[i].pose.center) for i in range(1, 8)])

    e0 = pose_err()
    hist = refine(st, solves, steps=20)
    e1 = pose_err()
    say(f"refinement: loss {hist[0]:.1f} -> {hist[-1]:.1f}, mean centre error {e0 * 1000:.2f} -> {e1 * 1000:.2f} mm")
    st.destroy()
    assert hist[-1] < hist[0]
