"""Edge cases of the C ABI on the GPU: empty and degenerate calls (empty batches, views without
sources, a zero-view TSDF pass, volumes nothing can see, empty meshes and clouds, an empty voxel
list for the adjoint), rejected calls (duplicate or unknown layers, a coarse start without a
coarse level, a set as its own coarse level, a coarse level destroyed before its fine set), a
context destroyed before its objects, ragged and extreme image sizes (1x1, 17x13, 16384x16,
16x16384: partial tiles and the widest rows) solved bit-exact to the oracle, and a TSDF pass outside
the ranges of the pass's correctly rounded short forms (the IEEE fallback) equal to the oracle."""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def sized(w, sizes):
    """The sphere workload's first views at other sizes (intrinsics scaled, frames re-rendered)."""
    from lr3d import StereoView, scale_intrinsics
    views, images = [], []
    for i, (W, H) in enumerate(sizes):
        v = w.views[i]
        k = scale_intrinsics(v.intrinsics, W / v.width, H / v.height)
        sv = StereoView(k, v.pose, W, H)
        g, _ = w.scene.render(sv)
        views.append(sv)
        images.append(g.reshape(H, W))
    return views, images


def test_empty_and_degenerate_solves(ctx):
    import lr3d
    say("case: empty batch, a view without sources, rejected batches")
    w = scenes.sphere_small().render()
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    st.solve_batched([])
    st.solve(0, lr3d.SolveOptions(sources=[], min_depth=0.5, max_depth=3.0, iterations=2))  # DS:187 no-op
    assert not st.depth_map(0).any()
    o = lr3d.SolveOptions(sources=w.sources[0], min_depth=0.5, max_depth=3.0, iterations=1)
    with pytest.raises(lr3d.Lr3dError):
        st.solve_batched([(0, o), (0, o)])
    with pytest.raises(lr3d.Lr3dError):
        st.solve(99, o)
    with pytest.raises(lr3d.Lr3dError):
        st.solve(0, lr3d.SolveOptions(sources=w.sources[0], min_depth=0.5, max_depth=3.0, iterations=1,
                                      from_coarse=True))
    # still usable after the rejected calls
    st.solve(0, o)
    assert st.depth_map(0).any()
    st.destroy()


def test_coarse_level_lifetime(ctx):
    import lr3d
    from lr3d import StereoView, scale_intrinsics
    say("case: a set cannot be its own coarse level; destroying the coarse level detaches it")
    w = scenes.sphere_small().render()
    coarse_views = [StereoView(scale_intrinsics(v.intrinsics, 0.5, 0.5), v.pose, 160, 120) for v in w.views]
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    coarse = lr3d.HipDenseStereo(ctx, coarse_views)
    with pytest.raises(lr3d.Lr3dError):
        st.set_coarse(st)
    st.set_coarse(coarse)
    coarse.destroy()
    o = lr3d.SolveOptions(sources=w.sources[0], min_depth=0.5, max_depth=3.0, iterations=1, from_coarse=True)
    with pytest.raises(lr3d.Lr3dError):
        st.solve_batched([(0, o)])
    st.solve(0, lr3d.SolveOptions(sources=w.sources[0], min_depth=0.5, max_depth=3.0, iterations=1))
    assert st.depth_map(0).any()
    st.destroy()


def test_context_outlives_its_objects():
    import lr3d
    say("case: a context destroyed while a stereo set and a volume remain is released by the last destroy")
    w = scenes.sphere_small().render()
    c = lr3d.Context(0)
    st = lr3d.HipDenseStereo(c, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    grid, layout = lr3d.dense_grid(w.tsdf.origin, w.tsdf.voxel, w.tsdf.voxels)
    vol = lr3d.HipTsdfVolume(c, grid, layout, w.tsdf.truncation_voxels)
    c.destroy()  # two users remain: the context stays until the last one goes
    st.solve(0, lr3d.SolveOptions(sources=w.sources[0], min_depth=0.5, max_depth=3.0, iterations=1))
    v = w.views[0]
    vol.integrate(st, 0, lr3d.IntegrationView(v.intrinsics, v.pose, 3.0, v.width, v.height))
    m = vol.extract_surface(0.6)  # a mesh of a live volume is one more user
    st.destroy()
    vol.destroy()
    assert len(m.read()["positions"]) > 0
    m.destroy()  # the last user: the context is released here


@pytest.mark.parametrize("sizes", [[(17, 13), (17, 13), (17, 13)], [(1, 1), (16, 16), (16, 16)],
                                   [(16384, 16), (16384, 16), (16384, 16)], [(16, 16384), (16, 16384), (16, 16384)]])
def test_ragged_and_extreme_sizes_match_oracle(ctx, sizes):
    import lr3d
    w = scenes.sphere_small()
    views, images = sized(w, sizes)
    say(f"case: sizes {sizes[0]} (+ sources {sizes[1:]}), 2 iterations, batched vs oracle")
    st = lr3d.HipDenseStereo(ctx, views)
    for i, im in enumerate(images):
        st.set_image(i, im)
    o = lr3d.SolveOptions(sources=[1, 2], min_depth=0.5, max_depth=3.0, iterations=2)
    st.solve_batched([(0, o)])
    hyp, cost = st.hypotheses(0)
    packed = st.depth_map(0)
    st.destroy()
    _, ohyp, ocost = O.solve(views, O.Images(images), 0, [1, 2], 0.5, 3.0, iterations=2)
    assert np.array_equal(hyp.view(np.uint32), ohyp.view(np.uint32))
    assert np.array_equal(cost.view(np.uint32), ocost.view(np.uint32))
    assert packed.size == sizes[0][0] * sizes[0][1]


def test_tsdf_degenerate_passes(ctx):
    import lr3d
    say("case: zero-view pass, views that see nothing, an all-empty map, an empty voxel list")
    w = scenes.sphere_small().render()
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    st.solve(0, lr3d.SolveOptions(sources=w.sources[0], min_depth=0.5, max_depth=3.0, iterations=1))
    grid, layout = lr3d.dense_grid(w.tsdf.origin, w.tsdf.voxel, w.tsdf.voxels)
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, w.tsdf.truncation_voxels)
    vol.set_counting(True)
    vol.integrate_batched(st, [], [])
    v = w.views[0]
    # a camera 100 m away, looking away from the volume
    from lr3d import Pose
    away = Pose(v.pose.R, v.pose.t + np.array([0.0, 0.0, -100.0]))
    vol.integrate(st, 0, lr3d.IntegrationView(v.intrinsics, away, 3.0, v.width, v.height))
    vol.integrate_host(np.zeros(v.width * v.height, np.uint32),
                       lr3d.IntegrationView(v.intrinsics, v.pose, 3.0, v.width, v.height))
    tsdf, weight = vol.read()
    assert vol.update_count() == 0
    assert np.all(tsdf == 1.0) and np.all(weight == 0.0)
    # the empty volume meshes to nothing, and so does its cleaned copy
    m = vol.extract_surface(0.6)
    empty = m.read()
    assert len(empty["positions"]) == 0 and len(empty["indices"]) == 0
    c = m.remove_small_components(0.01, 200)
    assert len(c.read()["indices"]) == 0
    c.destroy()
    m.destroy()
    # adjoint over no voxels: zero loss and gradients
    out = vol.backward(st, [0], [lr3d.IntegrationView(v.intrinsics, v.pose, 3.0, v.width, v.height)],
                       np.zeros(0, np.uint32), np.zeros(0, np.float32))
    assert out["loss"] == 0.0 and not np.any(out["grad_pose"]) and not np.any(out["grad_depth"][0])
    vol.destroy()
    st.destroy()


def test_fusion_of_empty_maps_and_a_tiny_budget(ctx):
    import lr3d
    say("case: fusion of empty maps; an 8-cell budget coarsens until the cloud fits (one cell per octant)")
    w = scenes.sphere_small().render()
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    thumbs = [np.zeros((60, 80, 3), np.uint8) for _ in range(2)]
    cloud, cell = lr3d.fuse_depth_maps(st, [0, 1], thumbs, 0.005, 1000, map=lr3d.MAP_RAW)
    assert len(cloud["positions"]) == 0 and cell == 0.005
    st.solve_batched([(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=0.5, max_depth=3.0, iterations=2))
                      for r in (0, 1)])
    # the sphere straddles the origin, so cells of any size split the samples by octant
    cloud, cell = lr3d.fuse_depth_maps(st, [0, 1], thumbs, 0.005, 8, map=lr3d.MAP_RAW)
    samples = int((st.depth_map(0) != 0).sum() + (st.depth_map(1) != 0).sum())
    assert 1 <= len(cloud["positions"]) <= 8 and cell > 0.5
    assert int(cloud["support"].astype(np.int64).sum()) == samples or cloud["support"].max() == 65535
    st.destroy()


def test_tsdf_ieee_fallback_matches_oracle(ctx):
    """A pass outside the ranges of the TSDF pass's correctly rounded short forms runs the IEEE
    divisions and square roots instead (focal lengths of 2^41 px: only voxels on the optical axis
    project into the map), and one inside them runs the short forms: both equal the oracle bit for
    bit, applied one after the other to the same volume."""
    import lr3d
    from lr3d import Intrinsics, Pose
    say("case: TSDF pass with fx = fy = 2^41 (IEEE forms) then fx = fy = 64 (short forms), 64^3 grid")
    voxel = 2.0 ** -7                      # exact in FP32, so the axis voxels' centres are exactly 0
    origin = (-31.5 * voxel, -31.5 * voxel, 1.0)
    grid, layout = lr3d.dense_grid(origin, voxel, 64)
    W = H = 64
    packed = np.full(W * H, O.pack_sample(1.25, (0.0, 0.0, -1.0), 0.1), np.uint32)
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, 4.0)
    vol.set_counting(True)
    g = O.Grid(grid.origin, grid.voxel, grid.dims, layout[0], voxel * 4.0)
    otsdf = np.ones(layout[2] * 512, np.float32)
    oweight = np.zeros(layout[2] * 512, np.float32)
    counts = []
    for f in (2.0 ** 41, 64.0):
        k = Intrinsics(f, f, 31.5, 31.5)
        vol.integrate_host(packed, lr3d.IntegrationView(k, Pose(), 3.0, W, H))
        iv = O.integration_view(W, H, f, f, 31.5, 31.5, np.eye(3), np.zeros(3), 3.0)
        counts.append(O.tsdf_integrate(g, iv, packed, None, otsdf, oweight))
        tsdf, weight = vol.read()
        assert np.array_equal(tsdf.view(np.uint32), otsdf.view(np.uint32))
        assert np.array_equal(weight.view(np.uint32), oweight.view(np.uint32))
    say(f"tsdf fallback: {counts[0]} updates on the axis, then {counts[1]}; device counted {vol.update_count()}")
    assert 0 < counts[0] <= 64 and counts[1] > 10000
    assert vol.update_count() == sum(counts)
    vol.destroy()
