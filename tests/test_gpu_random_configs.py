"""Seeded random solve configurations on the GPU against the CPU oracle, bit for bit.

Each case draws, from a fixed seed: a subset of the small sphere workload's cameras, an image
size per view (24-200 x 16-160, mixed within a case, so row pitches, partial tiles and source
views larger or smaller than the reference all occur), a reference view, its sources (1 to 7,
in a random order), the top-k (the reference's rule or any value from 1 to the source count), the
iteration count and the depth range.  Two cases run the geometric pass over the oracle's
photometric maps (mixed sizes: source maps of other sizes than the reference's); the last two take 8 and 7
cameras with 7 and 6 sources, one with top-5 (the kernels' 8-slot selection network).  Planes and
costs of every pixel must equal the oracle's bit for bit: the candidate early exit, the texel
layout and the launch shapes change no result.  On the photometric cases the NCC adjoint's forward
pass must reproduce the solved costs bit for bit, with finite gradients.  Four coarse-to-fine draws
start the fine solve from a coarse level of a random scale."""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes

pytestmark = pytest.mark.gpu

CASES = 12
GEOMETRIC_CASES = (7, 9)


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def base():
    return scenes.sphere_small()


def draw(base, case):
    from lr3d import StereoView, scale_intrinsics
    rng = np.random.default_rng(9100 + case)
    n = int(rng.integers(3, 7)) if case < 10 else 8 - (case - 10)
    cams = sorted(rng.choice(len(base.views), n, replace=False).tolist())
    geometric = case in GEOMETRIC_CASES
    sizes = [(int(rng.integers(24, 201)), int(rng.integers(16, 161))) for _ in range(n)]
    views, images = [], []
    for c, (W, H) in zip(cams, sizes):
        v = base.views[c]
        sv = StereoView(scale_intrinsics(v.intrinsics, W / v.width, H / v.height), v.pose, W, H)
        g, _ = base.scene.render(sv)
        views.append(sv)
        images.append(g.reshape(H, W))
    ref = int(rng.integers(0, n))
    others = [i for i in range(n) if i != ref]
    sources = rng.permutation(others)[: int(rng.integers(1, len(others) + 1))].tolist()
    topk = 0 if rng.random() < 0.4 else int(rng.integers(1, len(sources) + 1))
    if case >= 10:
        sources = rng.permutation(others).tolist()
        topk = 5 if case == 10 else 0
    iterations = int(rng.integers(1, 4))
    lo = float(rng.uniform(0.3, 1.0))
    hi = float(rng.uniform(2.0, 4.0))
    return views, images, ref, sources, topk, iterations, lo, hi, geometric


@pytest.mark.parametrize("case", range(CASES))
def test_random_configuration_matches_oracle(ctx, base, case):
    import lr3d
    views, images, ref, sources, topk, iterations, lo, hi, geometric = draw(base, case)
    say(f"case {case}: {len(views)} views {[(v.width, v.height) for v in views]}, ref {ref}, "
        f"sources {sources}, topk {topk}, iterations {iterations}, depth {lo:.2f}-{hi:.2f}, "
        f"geometric {geometric}")
    imgs = O.Images(images)
    st = lr3d.HipDenseStereo(ctx, views)
    for i, im in enumerate(images):
        st.set_image(i, im)
    depths = None
    if geometric:
        depths = O.map_pool(views)
        for r in range(len(views)):
            srcs = [i for i in range(len(views)) if i != r]
            O.solve(views, imgs, r, srcs, lo, hi, iterations=1, depths=depths)
        for r in range(len(views)):
            st.write_depth(r, O.layer_map(views, depths, r))
    o = lr3d.SolveOptions(sources=sources, min_depth=lo, max_depth=hi, iterations=iterations, topk=topk,
                          geometric=geometric)
    st.solve_batched([(ref, o)])
    hyp, cost = st.hypotheses(ref)
    if not geometric:
        # the NCC adjoint's forward pass recomputes the solved planes' costs bit for bit
        out = st.ncc_backward(ref, o)
        assert np.array_equal(out["cost"].view(np.uint32), cost.view(np.uint32))
        assert all(np.isfinite(out[k]).all() for k in ("grad_depth", "grad_normal", "grad_pose"))
    st.destroy()
    _, ohyp, ocost = O.solve(views, imgs, ref, sources, lo, hi, geometric=geometric, iterations=iterations,
                             topk=topk, depths=None if depths is None else depths.copy())
    planes = np.all(hyp.view(np.uint32) == ohyp.view(np.uint32), axis=1)
    costs = cost.view(np.uint32) == ocost.view(np.uint32)
    say(f"case {case}: planes bit-exact on {planes.mean() * 100:.3f} %, costs on {costs.mean() * 100:.3f} %")
    assert planes.all() and costs.all()


@pytest.mark.parametrize("case", range(4))
def test_random_coarse_to_fine_matches_oracle(ctx, base, case):
    """Coarse-to-fine (PM:296-333) with a coarse level of a random scale (0.3-0.7 of each view's
    size, so the fine-to-coarse pixel mapping is not a factor of 2): the coarse solve and the fine
    solve started from it equal the oracle's coarse solve, init_coarse and propagation."""
    import lr3d
    from lr3d import StereoView, scale_intrinsics
    views, images, ref, sources, topk, iterations, lo, hi, _ = draw(base, case)  # photometric draws
    rng = np.random.default_rng(9900 + case)
    scale = float(rng.uniform(0.3, 0.7))
    cviews = [StereoView(scale_intrinsics(v.intrinsics, max(8, round(v.width * scale)) / v.width,
                                          max(8, round(v.height * scale)) / v.height),
                         v.pose, max(8, round(v.width * scale)), max(8, round(v.height * scale)))
              for v in views]
    cimages = [base.scene.render(v, depth=False)[0].reshape(v.height, v.width) for v in cviews]
    say(f"coarse case {case}: {[(v.width, v.height) for v in views]} over {[(v.width, v.height) for v in cviews]}, "
        f"ref {ref}, sources {sources}, topk {topk}, fine iterations {iterations}")
    coarse = lr3d.HipDenseStereo(ctx, cviews)
    for i, im in enumerate(cimages):
        coarse.set_image(i, im)
    fine = lr3d.HipDenseStereo(ctx, views)
    for i, im in enumerate(images):
        fine.set_image(i, im)
    fine.set_coarse(coarse)
    coarse.solve(ref, lr3d.SolveOptions(sources=sources, min_depth=lo, max_depth=hi, iterations=2, topk=topk,
                                        pack=False))
    fine.solve_batched([(ref, lr3d.SolveOptions(sources=sources, min_depth=lo, max_depth=hi,
                                                iterations=iterations, topk=topk, from_coarse=True))])
    hyp, cost = fine.hypotheses(ref)
    chyp, _ = coarse.hypotheses(ref)
    fine.destroy()
    coarse.destroy()
    _, ochyp, _ = O.solve(cviews, O.Images(cimages), ref, sources, lo, hi, iterations=2, topk=topk, pack=False)
    assert np.array_equal(chyp.view(np.uint32), ochyp.view(np.uint32))
    imgs = O.Images(images)
    v, cv = views[ref], cviews[ref]
    ck = scale_intrinsics(v.intrinsics, cv.width / v.width, cv.height / v.height)
    p = O.make_params(views, ref, sources, lo, hi, topk=topk)
    ohyp, ocost = O.init_coarse(p, imgs, [ck.fx, ck.fy, ck.cx, ck.cy], cv.width, cv.height, ochyp)
    for it in range(iterations):
        for phase in (0, 1):
            p = O.make_params(views, ref, sources, lo, hi, iteration=it, phase=phase, step=0.5 ** it, topk=topk)
            O.propagate(p, imgs, ohyp, ocost)
    planes = np.all(hyp.view(np.uint32) == ohyp.view(np.uint32), axis=1)
    say(f"coarse case {case}: planes bit-exact on {planes.mean() * 100:.3f} %")
    assert planes.all()
    assert np.array_equal(cost.view(np.uint32), ocost.view(np.uint32))
