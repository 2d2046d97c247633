"""GPU parity: the HIP path through the C ABI against the CPU oracle, on the seeded 8-view
320x240 sphere workload.

Bars (stated per test): indices, RNG streams and schedule are bit-exact by construction, so the
planes and costs are compared bit for bit and must match on >= 99.9 % of pixels; packed depths
must agree within 1e-3 relative on >= 99.5 % of valid pixels; TSDF values within 1e-4 of the
truncation distance, weights exactly.
"""
import sys

import numpy as np
import pytest

import oracle as O
from lr3d import scenes
from lr3d.packing import unpack_depth

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def work():
    return scenes.sphere_small().render()


@pytest.fixture(scope="module")
def oracle_maps(work):
    """Oracle solves of every view (packed maps, layer-major) plus view 0's planes."""
    imgs = O.Images(work.images)
    W, H = 320, 240
    depths = np.zeros(len(work.views) * W * H, dtype=np.uint32)
    hyp0 = cost0 = None
    for ref in range(len(work.views)):
        _, hyp, cost = O.solve(work.views, imgs, ref, work.sources[ref], *work.ranges[ref],
                               iterations=work.iterations, topk=work.topk, depths=depths)
        if ref == 0:
            hyp0, cost0 = hyp.copy(), cost.copy()
    return depths, hyp0, cost0


def make_stereo(ctx, work):
    from lr3d import HipDenseStereo
    st = HipDenseStereo(ctx, work.views)
    for i, im in enumerate(work.images):
        st.set_image(i, im)
    return st


def opts(work, ref, **kw):
    from lr3d import SolveOptions
    lo, hi = work.ranges[ref]
    base = dict(sources=work.sources[ref], min_depth=lo, max_depth=hi, iterations=work.iterations,
                topk=work.topk)
    base.update(kw)
    return SolveOptions(**base)


def depth_agreement(a, b):
    """Fraction of pixels valid in either map whose depths agree within 1e-3 relative."""
    da, db = unpack_depth(a), unpack_depth(b)
    valid = (a != 0) | (b != 0)
    ok = (a != 0) & (b != 0) & (np.abs(da - db) <= 1e-3 * np.maximum(np.abs(db), 1e-12))
    return ok[valid].mean() if valid.any() else 1.0


def test_init_bit_exact(ctx, work):
    say("case: init (iterations=0) view 0, sphere 8x320x240")
    st = make_stereo(ctx, work)
    st.solve(0, opts(work, 0, iterations=0))
    hyp, cost = st.hypotheses(0)
    imgs = O.Images(work.images)
    p = O.make_params(work.views, 0, work.sources[0], *work.ranges[0], topk=work.topk)
    ohyp, ocost = O.init(p, imgs)
    exact = np.all(hyp.view(np.uint32) == ohyp.view(np.uint32), axis=1) & \
        (cost.view(np.uint32) == ocost.view(np.uint32))
    say(f"init bit-exact fraction {exact.mean():.6f}")
    st.destroy()
    assert exact.mean() >= 0.999


def test_solve_matches_oracle(ctx, work, oracle_maps):
    say("case: plain solve view 0, 4 iterations, top-2")
    depths, ohyp, ocost = oracle_maps
    st = make_stereo(ctx, work)
    st.solve(0, opts(work, 0))
    hyp, cost = st.hypotheses(0)
    packed = st.depth_map(0)
    st.destroy()
    exact = np.all(hyp.view(np.uint32) == ohyp.view(np.uint32), axis=1)
    agree = depth_agreement(packed, depths[:320 * 240])
    say(f"solve: planes bit-exact {exact.mean():.6f}, packed depth agreement {agree:.6f}")
    assert agree >= 0.995
    assert exact.mean() >= 0.99


def test_batched_equals_plain(ctx, work):
    say("case: batched (graph) vs plain solves, all 8 views")
    a = make_stereo(ctx, work)
    for ref in range(len(work.views)):
        a.solve(ref, opts(work, ref))
    b = make_stereo(ctx, work)
    b.solve_batched([(ref, opts(work, ref)) for ref in range(len(work.views))])
    for ref in range(len(work.views)):
        pa, pb = a.depth_map(ref), b.depth_map(ref)
        ha, hb = a.hypotheses(ref)[0], b.hypotheses(ref)[0]
        assert np.array_equal(pa, pb), f"view {ref} packed maps differ"
        assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32)), f"view {ref} planes differ"
    # replays of the captured graph, with eager work in between, stay identical
    first = [b.depth_map(r).copy() for r in range(len(work.views))]
    for replay in range(3):
        say(f"graph replay {replay}")
        a.solve(0, opts(work, 0))
        b.solve_batched([(ref, opts(work, ref)) for ref in range(len(work.views))])
        for ref in range(len(work.views)):
            assert np.array_equal(b.depth_map(ref), first[ref])
    a.destroy()
    b.destroy()


def test_all_views_match_oracle(ctx, work, oracle_maps):
    say("case: batched solve of all views vs oracle")
    depths = oracle_maps[0]
    st = make_stereo(ctx, work)
    st.solve_batched([(ref, opts(work, ref)) for ref in range(len(work.views))])
    n = 320 * 240
    for ref in range(len(work.views)):
        agree = depth_agreement(st.depth_map(ref), depths[ref * n:(ref + 1) * n])
        say(f"view {ref}: depth agreement {agree:.6f}")
        assert agree >= 0.995
    st.destroy()


def test_filter_matches_oracle(ctx, work, oracle_maps):
    say("case: consistency filter of view 0 on oracle maps")
    depths = oracle_maps[0]
    st = make_stereo(ctx, work)
    n = 320 * 240
    for ref in range(len(work.views)):
        st.write_depth(ref, depths[ref * n:(ref + 1) * n])
    checks = work.sources[0]
    gpu = st.filtered_map(0, checks, 2)
    ref_map = O.filter_map(work.views, 0, checks, 2, depths)
    st.destroy()
    say(f"filter: exact fraction {(gpu == ref_map).mean():.6f}, kept {(ref_map != 0).mean():.3f}")
    assert np.array_equal(gpu, ref_map)


def _tsdf_setup(work):
    from lr3d import IntegrationView, dense_grid
    spec = work.tsdf
    grid, layout = dense_grid(spec.origin, spec.voxel, spec.voxels)
    views = [IntegrationView(v.intrinsics, v.pose, work.ranges[i][1], v.width, v.height)
             for i, v in enumerate(work.views)]
    return grid, layout, views


def _oracle_tsdf(work, grid, layout, depths, n_views):
    spec = work.tsdf
    g = O.Grid(grid.origin, grid.voxel, grid.dims, layout[0], spec.voxel * spec.truncation_voxels)
    tsdf = np.ones(layout[2] * 512, dtype=np.float32)
    weight = np.zeros(layout[2] * 512, dtype=np.float32)
    n = 320 * 240
    updates = 0
    for i in range(n_views):
        v = work.views[i]
        iv = O.integration_view(v.width, v.height, v.intrinsics.fx, v.intrinsics.fy, v.intrinsics.cx,
                                v.intrinsics.cy, v.pose.R, v.pose.t, work.ranges[i][1])
        updates += O.tsdf_integrate(g, iv, depths[i * n:(i + 1) * n], None, tsdf, weight)
    return tsdf, weight, updates


def test_tsdf_matches_oracle(ctx, work, oracle_maps):
    say("case: TSDF 128^3 x 8 mm, 8 oracle maps, sequential host-map integration")
    from lr3d import HipTsdfVolume
    depths = oracle_maps[0]
    grid, layout, views = _tsdf_setup(work)
    vol = HipTsdfVolume(ctx, grid, layout, work.tsdf.truncation_voxels)
    vol.set_counting(True)
    n = 320 * 240
    for i, v in enumerate(views):
        vol.integrate_host(depths[i * n:(i + 1) * n], v)
    tsdf, weight = vol.read()
    count = vol.update_count()
    vol.destroy()
    otsdf, oweight, oupdates = _oracle_tsdf(work, grid, layout, depths, len(views))
    mu = work.tsdf.voxel * work.tsdf.truncation_voxels
    say(f"tsdf: max |diff| {np.abs(tsdf - otsdf).max():.3g} (bar {1e-4 * mu:.3g}), "
        f"updates gpu {count} oracle {oupdates}")
    assert np.array_equal(weight, oweight)
    assert np.abs(tsdf - otsdf).max() <= 1e-4 * mu
    assert count == oupdates


def test_tsdf_pass_counts(ctx, work, oracle_maps):
    """The counting pass's totals of one batched pass into a fresh volume: updates equal the
    oracle's, written voxels equal the voxels with weight > 0, and the evaluated pairs and bricks
    read bound them (512 pairs per kept (brick, view); a brick is read when a view is kept)."""
    say("case: TSDF pass counts, 8 oracle maps in one batched pass")
    from lr3d import HipTsdfVolume
    depths = oracle_maps[0]
    grid, layout, views = _tsdf_setup(work)
    st = make_stereo(ctx, work)
    n = 320 * 240
    for i in range(len(views)):
        st.write_depth(i, depths[i * n:(i + 1) * n])
    vol = HipTsdfVolume(ctx, grid, layout, work.tsdf.truncation_voxels)
    vol.set_counting(True)
    vol.integrate_batched(st, list(range(len(views))), views)
    counts = vol.pass_counts()
    tsdf, weight = vol.read()
    vol.destroy()
    st.destroy()
    _, _, oupdates = _oracle_tsdf(work, grid, layout, depths, len(views))
    touched = (weight.reshape(-1, 512) > 0).any(axis=1).sum()
    say(f"tsdf counts: {counts}, voxels with weight {(weight > 0).sum()}, bricks touched {touched}")
    assert counts["updates"] == oupdates
    assert counts["voxels_written"] == int((weight > 0).sum())
    assert counts["bricks_read"] >= touched and counts["bricks_read"] <= layout[2]
    assert counts["pairs"] >= counts["updates"] and counts["pairs"] % 512 == 0
    assert counts["pairs"] <= 512 * counts["bricks_read"] * len(views)


def test_tsdf_batched_equals_sequential(ctx, work, oracle_maps):
    say("case: TSDF batched (8 views, one pass) vs sequential")
    from lr3d import HipTsdfVolume
    depths = oracle_maps[0]
    grid, layout, views = _tsdf_setup(work)
    st = make_stereo(ctx, work)
    n = 320 * 240
    for i in range(len(views)):
        st.write_depth(i, depths[i * n:(i + 1) * n])
    a = HipTsdfVolume(ctx, grid, layout, work.tsdf.truncation_voxels)
    b = HipTsdfVolume(ctx, grid, layout, work.tsdf.truncation_voxels)
    for i, v in enumerate(views):
        a.integrate(st, i, v)
    b.integrate_batched(st, list(range(len(views))), views)
    ta, wa = a.read()
    tb, wb = b.read()
    for x in (a, b, st):
        x.destroy()
    assert np.array_equal(ta.view(np.uint32), tb.view(np.uint32))
    assert np.array_equal(wa.view(np.uint32), wb.view(np.uint32))


def test_depth_culling_changes_nothing(ctx, work, oracle_maps):
    """Depth culling drops only (brick, view) pairs that update no voxel: volumes and update
    counts are identical with it on and off, for raw maps, for views with a prior (dense, and
    holed with zeros and NaNs) and with confidence weighting, batched and one view per pass."""
    say("case: TSDF depth culling on vs off (8 views, priors on views 2 and 5, confidence on view 3)")
    from dataclasses import replace
    from lr3d import HipTsdfVolume
    depths = oracle_maps[0]
    grid, layout, views = _tsdf_setup(work)
    n = 320 * 240
    truth2 = work.truth[2].astype(np.float32).reshape(-1).copy()
    holed5 = work.truth[5].astype(np.float32).reshape(-1).copy()
    holed5[::7] = 0.0
    holed5[3::11] = np.nan
    views[2] = replace(views[2], prior=truth2)
    views[5] = replace(views[5], prior=holed5)
    views[3] = replace(views[3], confidence=1.2)
    st = make_stereo(ctx, work)
    for i in range(len(views)):
        st.write_depth(i, depths[i * n:(i + 1) * n])
    out = {}
    for cull in (True, False):
        for batched in (True, False):
            v = HipTsdfVolume(ctx, grid, layout, work.tsdf.truncation_voxels)
            v.set_depth_culling(cull)
            v.set_counting(True)
            if batched:
                v.integrate_batched(st, list(range(len(views))), views)
            else:
                for i, iv in enumerate(views):
                    v.integrate(st, i, iv)
            out[cull, batched] = (*v.read(), v.update_count())
            v.destroy()
    st.destroy()
    ref = out[False, False]
    say(f"culling: {ref[2]} voxel updates")
    for key, (t, w, c) in out.items():
        assert np.array_equal(t.view(np.uint32), ref[0].view(np.uint32)), key
        assert np.array_equal(w.view(np.uint32), ref[1].view(np.uint32)), key
        assert c == ref[2], key


def test_geometric_pass_matches_oracle(ctx, work, oracle_maps):
    say("case: geometric (ACMM) pass of view 0 over the oracle's photometric maps")
    depths = oracle_maps[0].copy()
    st = make_stereo(ctx, work)
    n = 320 * 240
    for ref in range(len(work.views)):
        st.write_depth(ref, depths[ref * n:(ref + 1) * n])
    st.solve(0, opts(work, 0, geometric=True, iterations=2))
    gpu = st.depth_map(0)
    hyp, _ = st.hypotheses(0)
    st.destroy()
    imgs = O.Images(work.images)
    _, ohyp, _ = O.solve(work.views, imgs, 0, work.sources[0], *work.ranges[0], geometric=True,
                         iterations=2, topk=work.topk, depths=depths)
    exact = np.all(hyp.view(np.uint32) == ohyp.view(np.uint32), axis=1).mean()
    agree = depth_agreement(gpu, depths[:n])
    say(f"geometric: planes bit-exact {exact:.6f}, depth agreement {agree:.6f}")
    assert agree >= 0.995 and exact >= 0.99


def test_batched_prior_after_a_batch_without_one(ctx, work):
    """A batched solve replays a HIP graph captured per batch layout; the prior buffer is
    allocated by the first solve that passes a prior, after which graphs captured without it are
    captured again.  Same layout without, with, then without a prior: each equals the oracle."""
    say("case: batched solves of view 0 without, with and again without a prior (one batch layout)")
    rng = np.random.default_rng(8)
    truth = work.truth[0].reshape(-1)
    prior = np.where(rng.random(truth.size) < 0.7, truth * (1 + 0.02 * rng.standard_normal(truth.size)),
                     0).astype(np.float32)
    st = make_stereo(ctx, work)
    imgs = O.Images(work.images)
    for with_prior in (False, True, False):
        p = prior if with_prior else None
        st.solve_batched([(0, opts(work, 0, prior=p, iterations=1))])
        hyp, cost = st.hypotheses(0)
        _, ohyp, ocost = O.solve(work.views, imgs, 0, work.sources[0], *work.ranges[0], prior=p,
                                 iterations=1, topk=work.topk)
        assert np.array_equal(hyp.view(np.uint32), ohyp.view(np.uint32)), with_prior
        assert np.array_equal(cost.view(np.uint32), ocost.view(np.uint32)), with_prior
    st.destroy()


def test_prior_matches_oracle(ctx, work):
    say("case: solve with a phone-depth prior (noisy truth), view 0")
    rng = np.random.default_rng(7)
    truth = work.truth[0].reshape(-1)
    prior = np.where(rng.random(truth.size) < 0.7, truth * (1 + 0.02 * rng.standard_normal(truth.size)),
                     0).astype(np.float32)
    st = make_stereo(ctx, work)
    st.solve(0, opts(work, 0, prior=prior, iterations=1))
    hyp, cost = st.hypotheses(0)
    st.destroy()
    imgs = O.Images(work.images)
    _, ohyp, ocost = O.solve(work.views, imgs, 0, work.sources[0], *work.ranges[0], prior=prior,
                             iterations=1, topk=work.topk)
    exact = np.all(hyp.view(np.uint32) == ohyp.view(np.uint32), axis=1).mean()
    say(f"prior: planes bit-exact {exact:.6f}")
    assert exact >= 0.99


def test_coarse_to_fine_matches_oracle(ctx, work):
    say("case: coarse-to-fine (160x120 coarse level, 320x240 fine), view 0")
    from lr3d import HipDenseStereo, StereoView, scale_intrinsics
    cw, ch = 160, 120
    cviews = [StereoView(scale_intrinsics(v.intrinsics, cw / v.width, ch / v.height), v.pose, cw, ch)
              for v in work.views]
    cimages = [work.scene.render(v, depth=False)[0].reshape(ch, cw) for v in cviews]
    coarse = HipDenseStereo(ctx, cviews)
    for i, im in enumerate(cimages):
        coarse.set_image(i, im)
    fine = make_stereo(ctx, work)
    fine.set_coarse(coarse)
    coarse.solve(0, opts(work, 0, iterations=2, pack=False))
    fine.solve(0, opts(work, 0, iterations=1, from_coarse=True))
    hyp, cost = fine.hypotheses(0)
    chyp, _ = coarse.hypotheses(0)
    fine.destroy()
    coarse.destroy()
    # oracle: coarse solve, then init_coarse + propagate + pack at the fine level
    _, ochyp, _ = O.solve(cviews, O.Images(cimages), 0, work.sources[0], *work.ranges[0],
                          iterations=2, topk=work.topk, pack=False)
    assert np.array_equal(chyp.view(np.uint32), ochyp.view(np.uint32))
    imgs = O.Images(work.images)
    ck = scale_intrinsics(work.views[0].intrinsics, cw / 320, ch / 240)
    p = O.make_params(work.views, 0, work.sources[0], *work.ranges[0], topk=work.topk)
    ohyp, ocost = O.init_coarse(p, imgs, [ck.fx, ck.fy, ck.cx, ck.cy], cw, ch, ochyp)
    for phase in (0, 1):
        p = O.make_params(work.views, 0, work.sources[0], *work.ranges[0], iteration=0, phase=phase,
                          step=1.0, topk=work.topk)
        O.propagate(p, imgs, ohyp, ocost)
    exact = np.all(hyp.view(np.uint32) == ohyp.view(np.uint32), axis=1).mean()
    say(f"coarse-to-fine: planes bit-exact {exact:.6f}")
    assert exact >= 0.99


def test_mixed_resolution_batch_matches_oracle(ctx, work):
    say("case: one batch of views at 320x240, 400x300 and 256x192")
    from lr3d import HipDenseStereo, SolveOptions, StereoView, scale_intrinsics
    sizes = [(320, 240), (400, 300), (256, 192)]
    views, images = [], []
    for i, v in enumerate(work.views[:6]):
        w, h = sizes[i % 3]
        k = scale_intrinsics(v.intrinsics, w / v.width, h / v.height)
        sv = StereoView(k, v.pose, w, h)
        views.append(sv)
        images.append(work.scene.render(sv, depth=False)[0].reshape(h, w))
    srcs = {0: [1, 5, 2], 1: [0, 2, 3, 5], 2: [1, 3]}
    st = HipDenseStereo(ctx, views)
    for i, im in enumerate(images):
        st.set_image(i, im)
    st.solve_batched([(r, SolveOptions(sources=s, min_depth=0.5, max_depth=3.0, iterations=2))
                      for r, s in srcs.items()])
    imgs = O.Images(images)
    for r, s in srcs.items():
        hyp, _ = st.hypotheses(r)
        _, ohyp, _ = O.solve(views, imgs, r, s, 0.5, 3.0, iterations=2)
        exact = np.all(hyp.view(np.uint32) == ohyp.view(np.uint32), axis=1).mean()
        say(f"mixed: view {r} ({views[r].width}x{views[r].height}, {len(s)} sources) bit-exact {exact:.6f}")
        assert exact >= 0.99
    st.destroy()


def test_room_mixed_batch_matches_oracle(ctx):
    """The batched room configuration at quarter scale: 64 views at 160x120, 240x135, 320x180 and
    480x270 with 4-8 neighbours each, all solved in one batch; one view of each size is checked
    against the oracle."""
    from lr3d import HipDenseStereo, SolveOptions, scenes
    w = scenes.room_batched(scale=0.25, iterations=2).render()
    say(f"case: room batch x0.25, {len(w.views)} views, {w.megapixels:.2f} Mpx, 2 iterations")
    st = HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    st.solve_batched([(r, SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                                       iterations=w.iterations, topk=w.topk)) for r in range(len(w.views))])
    imgs = O.Images(w.images)
    for r in (0, 1, 2, 3):
        hyp, _ = st.hypotheses(r)
        _, ohyp, _ = O.solve(w.views, imgs, r, w.sources[r], *w.ranges[r], iterations=w.iterations, topk=w.topk)
        exact = np.all(hyp.view(np.uint32) == ohyp.view(np.uint32), axis=1).mean()
        say(f"room: view {r} ({w.views[r].width}x{w.views[r].height}, {len(w.sources[r][:8])} sources) "
            f"planes bit-exact {exact:.6f}")
        assert exact >= 0.99
    st.destroy()
