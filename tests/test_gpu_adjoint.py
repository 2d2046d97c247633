"""GPU adjoint of the aggregated NCC cost (lr3d_ncc_backward / lr3d_ncc_refine_step) against the
FP64 CPU reference gradient (oracle/adjoint_oracle.c, itself pinned by central differences in
tests/test_oracle_adjoint.py) and against FP64 central differences of 16 world-pose parameters.

Compared pixels are the well-conditioned ones among 1024 seeded pixels: FP32 and FP64 costs
agree within 1e-4 (in low-texture windows the forward's FP32 variance cancels, and FP32 can see a
cost where FP64 sees a degenerate window, or the reverse) and no tap of a selected view lies
within 1e-4 px of a texel edge (FP32 and FP64 tap positions can fall in different bilinear cells
there, where the gradient jumps).

The backward passes run in FP64 from the FP32 planes, so the bars are tight: every output finite
over the whole view; per compared pixel, |g_gpu - g_fp64| <= 1e-4 * max(1, |g_fp64|) for dL/dd and
dL/dn on >= 99 % of them and <= 1e-3 on all; the summed relative-pose gradient within 1e-4
relative; world-pose gradients within 2e-2 of FP64 central differences, relative to the norm of
the view's gradient (the summed loss crosses texel edges inside any step, see
test_oracle_adjoint.py)."""
import os
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


def selection_gap(r, n_src):
    """Gap between the largest selected view cost and the smallest unselected one of an FP64
    result (infinite when every view is selected).  Where two views nearly tie, the FP32 forward
    costs may select the other one, and the gradient belongs to another branch of the top-k."""
    costs = np.array(r.view_cost[:n_src])
    sel = np.array(r.selected[:n_src], dtype=bool)
    if sel.all() or not sel.any():
        return np.inf
    return float(costs[~sel].min() - costs[sel].max())


def clamp_gap(r, n_src):
    """Smallest distance of a selected view's FP64 cost to the clamps at 0 and 2 (PM:158).  Near
    them the FP32 cost of a highly correlated window (cancellation in the variance) can be clamped,
    carrying no gradient, where the FP64 one is not."""
    costs = np.array(r.view_cost[:n_src])
    sel = np.array(r.selected[:n_src], dtype=bool)
    return float(np.min(np.minimum(costs[sel], 2.0 - costs[sel]))) if sel.any() else np.inf


def well_conditioned(r, cost, n_src):
    """FP32 and FP64 costs agree, no tap within 1e-4 px of a texel edge, no near-tie in the top-k,
    no selected view within 1e-3 of a clamp."""
    return (abs(r.cost - cost) <= 1e-4 and r.min_kink >= 1e-4 and selection_gap(r, n_src) >= 1e-4
            and clamp_gap(r, n_src) >= 1e-3)


@pytest.fixture(scope="module")
def reference(solved):
    """FP64 cost and gradient of the seeded pixels; the well-conditioned ones."""
    w, st, o, hyp, cost = solved
    imgs = O.Images(w.images)
    p = O.params64(w.views, 0, w.sources[0], *w.ranges[0], topk=w.topk)
    res = {}
    for i in seeded_pixels(cost):
        y, x = divmod(int(i), 320)
        h = hyp[i].astype(np.float64)
        res[int(i)] = O.cost_grad64(p, imgs, 0, x, y, h[:3], h[3])
    n_src = min(len(w.sources[0]), 8)
    clean = [i for i, r in res.items() if well_conditioned(r, cost[i], n_src)]
    return imgs, res, np.array(clean)


def finite_or_report(out, tag):
    bad = ~np.isfinite(out["grad_depth"]) | ~np.isfinite(out["grad_normal"]).all(axis=1)
    if bad.any() or not np.isfinite(out["grad_pose"]).all():
        dump = os.environ.get("LR3D_TEST_DUMP")  # a directory to keep the outputs in, if set
        if dump:
            os.makedirs(dump, exist_ok=True)
            np.savez(os.path.join(dump, f"adjoint_{tag}.npz"), **{k: v for k, v in out.items() if k != "loss"})
        px = [divmod(int(i), 320)[::-1] for i in np.flatnonzero(bad)[:8]]
        say(f"{tag}: {int(bad.sum())} pixels with non-finite gradients, first (x, y) {px}; "
            f"pose finite {bool(np.isfinite(out['grad_pose']).all())}")
    return not bad.any() and np.isfinite(out["grad_pose"]).all()


def test_pixel_gradients_match_fp64_reference(solved, reference):
    w, st, o, hyp, cost = solved
    imgs, res, clean = reference
    say("case: NCC adjoint per pixel, view 0 of sphere 8x320x240, 1024 seeded pixels")
    out = st.ncc_backward(0, o)
    assert np.array_equal(out["cost"].view(np.uint32), cost.view(np.uint32))  # forward is exact
    assert finite_or_report(out, "unweighted")
    good = 0
    worst = (0.0, None)
    for i in clean:
        r = res[int(i)]
        ref = np.array([r.d_depth, *r.d_normal])
        got = np.array([out["grad_depth"][i], *out["grad_normal"][i]], dtype=np.float64)
        e = np.max(np.abs(got - ref) / np.maximum(1.0, np.abs(ref)))
        good += e <= 1e-4
        worst = max(worst, (e, (int(i) % 320, int(i) // 320)), key=lambda a: a[0])
    say(f"adjoint: {good}/{len(clean)} compared pixels within 1e-4 of the FP64 reference (of {len(res)}), "
        f"worst {worst[0]:.2e} at {worst[1]}")
    assert len(clean) >= 512 and good >= 0.99 * len(clean) and worst[0] <= 1e-3


def test_pose_gradient_matches_fp64_reference(solved, reference):
    w, st, o, hyp, cost = solved
    imgs, res, clean = reference
    say(f"case: NCC adjoint, relative-pose gradient summed over {len(clean)} compared pixels")
    weights = np.zeros(320 * 240, np.float32)
    weights[clean] = 1.0
    out = st.ncc_backward(0, o, weights)
    assert finite_or_report(out, "weighted")
    ref = np.zeros((len(w.sources[0]), 12))
    loss = 0.0
    for i in clean:
        r = res[int(i)]
        loss += r.cost
        for s in range(len(w.sources[0])):
            ref[s, :9] += list(r.d_R[s])
            ref[s, 9:] += list(r.d_t[s])
    err = np.abs(out["grad_pose"] - ref).max() / np.abs(ref).max()
    say(f"pose gradient: max relative error {err:.2e}, loss gpu {out['loss']:.6f} fp64 {loss:.6f}")
    assert err <= 1e-4
    assert abs(out["loss"] - loss) <= 1e-4 * loss


def test_world_pose_gradients_match_central_differences(solved, reference):
    from lr3d.refine import perturb, world_pose_gradients
    w, st, o, hyp, cost = solved
    say("case: 16 world-pose parameters, FP64 central differences of the summed cost")
    imgs, res, clean = reference
    idx = clean[::max(1, len(clean) // 256)][:256]
    weights = np.zeros(320 * 240, np.float32)
    weights[idx] = 1.0
    out = st.ncc_backward(0, o, weights)
    assert finite_or_report(out, "fd")
    srcs = w.sources[0]
    grads = world_pose_gradients(w.views, 0, srcs, out["grad_pose"])
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
        norm = np.linalg.norm(np.concatenate(grads[layer]))
        errs.append(abs(fd - g) / max(1.0, norm))
    say(f"world-pose gradients vs FD: median relative error {np.median(errs):.2e}, max {max(errs):.2e}")
    assert len(params) == 16
    assert np.median(errs) <= 5e-3 and max(errs) <= 2e-2


def test_zero_step_refine_only_evaluates(solved):
    """lr3d_ncc_refine_step with both steps 0 evaluates: planes, costs and packed maps unchanged."""
    w, st, o, hyp, cost = solved
    say("case: refine step with zero steps leaves the set unchanged")
    packed = st.depth_map(0)
    gp, before, after = st.refine_step([(0, o)], 0.0, 0.0)
    h2, c2 = st.hypotheses(0)
    assert np.array_equal(h2.view(np.uint32), hyp.view(np.uint32))
    assert np.array_equal(c2.view(np.uint32), cost.view(np.uint32))
    assert np.array_equal(st.depth_map(0), packed)
    assert before[0] == after[0] and np.isfinite(gp).all()


def test_refinement_lowers_the_cost_with_noisy_poses(ctx):
    import lr3d
    from lr3d.refine import perturb, refine
    say("case: joint plane + pose refinement, sphere 8x320x240, poses noised (0.5 deg, 5 mm), 50 steps")
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
        c = np.mean([np.linalg.norm(st.views[i].pose.center - w.views[i].pose.center) for i in range(1, 8)])
        r = np.mean([np.degrees(np.arccos(np.clip((np.trace(st.views[i].pose.R @ w.views[i].pose.R.T) - 1) / 2, -1, 1)))
                     for i in range(1, 8)])
        return c, r

    (c0, r0) = pose_err()
    hist = refine(st, solves, steps=50)
    (c1, r1) = pose_err()
    say(f"refinement: loss {hist[0]:.1f} -> {hist[-1]:.1f}; mean centre error {c0 * 1000:.2f} -> {c1 * 1000:.2f} mm, "
        f"rotation error {r0:.3f} -> {r1:.3f} deg")
    st.destroy()
    assert np.isfinite(hist).all(), hist
    assert all(b <= a for a, b in zip(hist, hist[1:]))
    assert hist[-1] < 0.99 * hist[0]
