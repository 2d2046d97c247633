"""Differentiation of the aggregated NCC cost at the differentiation configuration: 16 views at
960x540 of the sphere scene with poses noised by seeded noise (0.5 deg RMS rotation, 5 mm RMS
translation; view 0 exact), 6 batched PatchMatch iterations at the noisy poses, then

  - the per-pixel adjoint (lr3d_ncc_backward, FP32 forward + FP64 backward, windows recomputed)
    on 1024 seeded pixels of view 0 against the FP64 CPU reference gradient
    (oracle/adjoint_oracle.c) and against FP64 central differences of the FP64 cost;
  - 16 world-pose parameters (6 of each of two source views, 4 of the reference view) against
    FP64 central differences of the summed cost;
  - joint refinement of every view's planes and poses (lr3d_ncc_refine_step batched over all 16
    views, planes re-solved after each pose step; 50 steps on each of an eighth-, quarter- and
    half-resolution copy of the set, then 25 at full resolution): non-increasing losses, and the
    mean rotation error below 35 % and the mean centre error below 70 % of the noise's.

Central differences use h = 1e-6 and keep a sample where the quotients with h and h/2 agree (no
texel-edge kink inside the step).  At h = 1e-4 the taps move far enough to cross texel edges now
and then, so the quotient is not always the derivative; the test reports the GPU's agreement with
it and checks that the GPU gradient and the FP64 analytic gradient are off from it by the same
amounts (any gap is the quotient's)."""
import time

import numpy as np
import pytest

import oracle as O
from lr3d import scenes
from test_gpu_adjoint import finite_or_report, seeded_pixels, well_conditioned

pytestmark = pytest.mark.gpu

W, H = 960, 540


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def diff(ctx):
    import lr3d
    w = scenes.sphere_diff().render()
    noisy = scenes.noisy_views(w.views)
    st = lr3d.HipDenseStereo(ctx, noisy)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    solves = [(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                                    iterations=w.iterations, topk=w.topk)) for r in range(len(w.views))]
    st.solve_batched(solves)
    hyp, cost = st.hypotheses(0)
    yield w, noisy, st, solves, hyp, cost
    st.destroy()


@pytest.fixture(scope="module")
def reference(diff):
    w, noisy, st, solves, hyp, cost = diff
    imgs = O.Images(w.images)
    p = O.params64(noisy, 0, w.sources[0], *w.ranges[0], topk=w.topk)
    res = {}
    for i in seeded_pixels(cost):
        y, x = divmod(int(i), W)
        h = hyp[i].astype(np.float64)
        res[int(i)] = O.cost_grad64(p, imgs, 0, x, y, h[:3], h[3])
    n_src = min(len(w.sources[0]), 8)
    clean = [i for i, r in res.items() if well_conditioned(r, cost[i], n_src)]
    return imgs, p, res, np.array(clean)


def central(f, x0, h):
    g1 = (f(x0 + h) - f(x0 - h)) / (2 * h)
    g2 = (f(x0 + h / 2) - f(x0 - h / 2)) / h
    return g1, g2


def test_pixel_adjoint_matches_fp64_and_central_differences(diff, reference):
    w, noisy, st, solves, hyp, cost = diff
    imgs, p, res, clean = reference
    say(f"case: NCC adjoint, view 0 of sphere 16x960x540 (noisy poses), 1024 seeded pixels, {len(clean)} compared")
    out = st.ncc_backward(0, solves[0][1])
    assert np.array_equal(out["cost"].view(np.uint32), cost.view(np.uint32))
    assert finite_or_report(out, "c4")
    e_ref, e_fd6, e_fd4_gpu, e_fd4_ref = [], [], [], []
    for i in clean:
        r = res[int(i)]
        y, x = divmod(int(i), W)
        h = hyp[i].astype(np.float64)
        n, d = h[:3], h[3]
        ref = np.array([r.d_depth, *r.d_normal])
        got = np.array([out["grad_depth"][i], *out["grad_normal"][i]], dtype=np.float64)
        scale = np.maximum(1.0, np.abs(ref))
        e_ref.append(np.max(np.abs(got - ref) / scale))
        sel = list(r.selected)

        def c(d_, n_):
            o = O.cost_grad64(p, imgs, 0, x, y, n_, d_)
            return o.cost if list(o.selected) == sel else np.nan

        fns = [lambda v: c(v, n)] + [lambda v, k=k: c(d, n + np.eye(3)[k] * v) for k in range(3)]
        x0s = [d, 0.0, 0.0, 0.0]
        for k, (f, x0) in enumerate(zip(fns, x0s)):
            g1, g2 = central(f, x0, 1e-6)
            if np.isfinite(g1) and np.isfinite(g2) and abs(g1 - g2) <= 1e-6 * max(1.0, abs(g1)):
                e_fd6.append(abs(got[k] - g1) / max(1.0, abs(g1)))
            q1, _ = central(f, x0, 1e-4)
            if np.isfinite(q1):
                e_fd4_gpu.append(abs(got[k] - q1) / max(1.0, abs(q1)))
                e_fd4_ref.append(abs(ref[k] - q1) / max(1.0, abs(q1)))
    e_ref, e_fd6 = np.array(e_ref), np.array(e_fd6)
    e_fd4_gpu, e_fd4_ref = np.array(e_fd4_gpu), np.array(e_fd4_ref)
    say(f"vs FP64 reference: {(e_ref <= 1e-4).mean() * 100:.2f} % within 1e-4, worst {e_ref.max():.2e}")
    say(f"vs central differences h=1e-6: {len(e_fd6)} kink-free quotients, {(e_fd6 <= 2e-4).mean() * 100:.2f} % "
        f"within 2e-4, median {np.median(e_fd6):.2e}")
    say(f"vs central differences h=1e-4: median error GPU {np.median(e_fd4_gpu):.2e}, FP64 analytic "
        f"{np.median(e_fd4_ref):.2e}; |GPU - FP64| of the two errors <= 1e-3 on "
        f"{(np.abs(e_fd4_gpu - e_fd4_ref) <= 1e-3).mean() * 100:.2f} %; GPU within 1e-3 of the quotient on "
        f"{(e_fd4_gpu <= 1e-3).mean() * 100:.2f} % (99th percentile {np.percentile(e_fd4_gpu, 99):.2e})")
    assert len(clean) >= 500
    assert (e_ref <= 1e-4).mean() >= 0.99 and e_ref.max() <= 1e-3
    assert len(e_fd6) >= 3 * len(clean) and (e_fd6 <= 2e-4).mean() >= 0.99
    assert (np.abs(e_fd4_gpu - e_fd4_ref) <= 1e-3).mean() >= 0.99


def test_world_pose_gradients_match_central_differences(diff, reference):
    from lr3d.refine import perturb, world_pose_gradients
    w, noisy, st, solves, hyp, cost = diff
    imgs, p, res, clean = reference
    say("case: 16 world-pose parameters of view 0's cost, FP64 central differences (sphere 16x960x540)")
    idx = clean[::max(1, len(clean) // 256)][:256]
    weights = np.zeros(W * H, np.float32)
    weights[idx] = 1.0
    out = st.ncc_backward(0, solves[0][1], weights)
    assert finite_or_report(out, "c4-pose")
    srcs = w.sources[0]
    grads = world_pose_gradients(noisy, 0, srcs, out["grad_pose"])
    planes = [(divmod(int(i), W), hyp[i].astype(np.float64)) for i in idx]

    def loss(views):
        q = O.params64(views, 0, srcs, *w.ranges[0], topk=w.topk)
        return sum(O.cost_grad64(q, imgs, 0, x, y, h[:3], h[3]).cost for (y, x), h in planes)

    params = [(srcs[0], k) for k in range(6)] + [(srcs[1], k) for k in range(6)] + [(0, k) for k in (0, 1, 3, 5)]
    errs, errs4 = [], []
    for layer, k in params:
        g = grads[layer][0][k] if k < 3 else grads[layer][1][k - 3]
        norm = np.linalg.norm(np.concatenate(grads[layer]))
        for hstep, sink in ((1e-6, errs), (1e-4, errs4)):
            vals = []
            for sign in (1, -1):
                e = np.zeros(6)
                e[k] = sign * hstep
                views = list(noisy)
                v = views[layer]
                views[layer] = type(v)(v.intrinsics, perturb(v.pose, e[:3], e[3:]), v.width, v.height)
                vals.append(loss(views))
            sink.append(abs((vals[0] - vals[1]) / (2 * hstep) - g) / max(1.0, norm))
    say(f"world-pose gradients vs FD: h=1e-6 median {np.median(errs):.2e}, max {max(errs):.2e}; "
        f"h=1e-4 median {np.median(errs4):.2e}, max {max(errs4):.2e}")
    assert len(params) == 16
    assert np.median(errs) <= 5e-3 and max(errs) <= 2e-2


def test_joint_refinement_pyramid(ctx, diff):
    from lr3d.refine import pyramid_levels, refine_pyramid
    w, noisy, st, solves, hyp, cost = diff
    say("case: joint plane + pose refinement over all 16 views (sphere 16x960x540): 50 steps at each of "
        "1/8, 1/4 and 1/2 resolution, 25 at full resolution")

    def pose_err():
        c = np.mean([np.linalg.norm(st.views[i].pose.center - w.views[i].pose.center) for i in range(1, 16)])
        r = np.mean([np.degrees(np.arccos(np.clip((np.trace(st.views[i].pose.R @ w.views[i].pose.R.T) - 1) / 2, -1, 1)))
                     for i in range(1, 16)])
        return c, r

    c0, r0 = pose_err()
    levels = pyramid_levels(st, w.images, 3)
    ctx.set_profiling(True)
    t0 = time.perf_counter()
    hc, hf = refine_pyramid(st, levels, solves, coarse_steps=(50, 50, 50), steps=25)
    elapsed = time.perf_counter() - t0
    launches, ms = ctx.kernel_stats("ncc_backward")
    ctx.set_profiling(False)
    for level in levels:
        level.destroy()
    c1, r1 = pose_err()
    say(f"refinement: losses per level (coarsest first) " +
        ", ".join(f"{h[0]:.1f} -> {h[-1]:.1f}" for h in hc + [hf]) +
        f"; mean centre error {c0 * 1000:.2f} -> {c1 * 1000:.2f} mm, rotation {r0:.3f} -> {r1:.3f} deg; "
        f"{elapsed:.1f} s for 175 steps; {launches} adjoint launches, {ms / max(launches, 1):.2f} ms each")
    for h in hc + [hf]:
        assert np.isfinite(h).all()
        assert all(b <= a for a, b in zip(h, h[1:]))
        assert h[-1] < h[0]
    assert hc[0][-1] < 0.98 * hc[0][0]
    assert r1 < 0.35 * r0 and c1 < 0.7 * c0
