"""Launch-shape autotuner (lr3d_stereo_autotune, lr3d_tsdf_autotune): every candidate shape gives
bit-identical maps and volumes, tuning leaves the set and the volume as it found them, and the
chosen shapes are among the candidates."""
import numpy as np
import pytest

from lr3d import scenes

pytestmark = pytest.mark.gpu

PROPAGATE = [(8, 8), (16, 4), (4, 16), (32, 2), (64, 1), (16, 8), (8, 16), (32, 4), (16, 16), (64, 4)]
INIT = [(8, 8), (16, 4), (32, 2), (64, 1), (16, 8), (64, 4)]
PACK = [(64, 4), (64, 1), (64, 2), (32, 8), (256, 1)]


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def work():
    return scenes.sphere_small().render()


def _stereo(ctx, w):
    import lr3d
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    solves = [(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                                    iterations=w.iterations, topk=w.topk)) for r in range(len(w.views))]
    return st, solves


def test_every_shape_gives_identical_maps(ctx, work):
    say("case: sphere 8x320x240 solved with every candidate shape of init, propagate and pack")
    ref, solves = _stereo(ctx, work)
    ref.solve_batched(solves)
    expect = [ref.depth_map(r) for r in range(8)]
    ref.destroy()
    st, _ = _stereo(ctx, work)
    for kernel, shapes in (("propagate", PROPAGATE), ("init", INIT), ("pack", PACK)):
        for tx, ty in shapes:
            st.set_tile(kernel, tx, ty)
            st.solve_batched(solves)
            for r in range(8):
                assert np.array_equal(st.depth_map(r), expect[r]), (kernel, tx, ty, r)
        st.set_tile(kernel, *shapes[0])
    st.destroy()


def test_autotune_keeps_the_set_and_picks_a_candidate(ctx, work):
    say("case: stereo autotune on the sphere batch")
    st, solves = _stereo(ctx, work)
    st.solve_batched(solves)
    before = [st.depth_map(r) for r in range(8)]
    hyp_before = st.hypotheses(3)
    tiles = st.autotune(solves, repeats=2)
    say(f"autotune: {tiles}")
    assert tiles["propagate"] in PROPAGATE and tiles["init"] in INIT and tiles["pack"] in PACK
    for r in range(8):
        assert np.array_equal(st.depth_map(r), before[r])
    hyp_after = st.hypotheses(3)
    assert np.array_equal(hyp_after[0], hyp_before[0]) and np.array_equal(hyp_after[1], hyp_before[1])
    st.solve_batched(solves)
    for r in range(8):
        assert np.array_equal(st.depth_map(r), before[r])
    st.destroy()


def test_tsdf_autotune_and_every_block_size(ctx, work):
    import lr3d
    say("case: TSDF 128^3 x 8 mm, bricks per block 1 / 2 / 4, XCD chunks 16 / 64 / 4096 and autotune")
    st, solves = _stereo(ctx, work)
    st.solve_batched(solves)
    spec = work.tsdf
    grid, layout = lr3d.dense_grid(spec.origin, spec.voxel, spec.voxels)
    views = [lr3d.IntegrationView(v.intrinsics, v.pose, work.ranges[i][1], v.width, v.height)
             for i, v in enumerate(work.views)]
    results = []
    for bricks in (1, 2, 4):
        vol = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
        vol.bricks_per_block = bricks
        vol.integrate_batched(st, list(range(8)), views)
        results.append(vol.read())
        vol.destroy()
    for chunk in (16, 64, 4096):  # the last is beyond the launch's blocks: every block keeps its index
        vol = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
        vol.bricks_per_block = 2
        vol.xcd_chunk = chunk
        assert vol.xcd_chunk == chunk
        vol.integrate_batched(st, list(range(8)), views)
        results.append(vol.read())
        vol.destroy()
    for t, w in results[1:]:
        assert np.array_equal(t.view(np.uint32), results[0][0].view(np.uint32))
        assert np.array_equal(w.view(np.uint32), results[0][1].view(np.uint32))
    # the 8-view passes above launch a wavefront per pool brick; passes of one or two views walk the
    # list of bricks with a kept view (the other schedule).  The 8 maps three times over in one pass
    # with more shapes, and the same 24 views applied one or two per pass with several chunk sizes
    layers24, views24 = list(range(8)) * 3, views * 3
    pool = []
    for bricks, chunk in ((1, 1), (4, 1), (1, 64), (1, 4096)):
        vol = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
        vol.bricks_per_block = bricks
        vol.xcd_chunk = chunk
        vol.integrate_batched(st, layers24, views24, map=lr3d.MAP_RAW)
        pool.append(vol.read())
        vol.destroy()
    listed = []
    for per, bricks, chunk in ((1, 1, 1), (1, 2, 64), (2, 4, 16), (1, 1, 4096)):
        vol = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
        vol.bricks_per_block = bricks
        vol.xcd_chunk = chunk
        for a in range(0, 24, per):
            vol.integrate_batched(st, layers24[a:a + per], views24[a:a + per], map=lr3d.MAP_RAW)
        listed.append(vol.read())
        vol.destroy()
    say(f"tsdf: 24 views in one pass with {len(pool)} launch shapes, and in one- and two-view list "
        f"passes with {len(listed)}")
    for t, w in pool[1:] + listed:
        assert np.array_equal(t.view(np.uint32), pool[0][0].view(np.uint32))
        assert np.array_equal(w.view(np.uint32), pool[0][1].view(np.uint32))
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
    for bad in (0, 4097):
        with pytest.raises(lr3d.Lr3dError):
            vol.xcd_chunk = bad
    vol.integrate_batched(st, list(range(4)), views[:4])
    before = vol.read()
    chosen = vol.autotune(st, list(range(8)), views, repeats=2)
    say(f"tsdf autotune: {chosen} bricks per block, chunks of {vol.xcd_chunk} blocks per XCD")
    assert chosen in (1, 2, 4)
    assert vol.xcd_chunk in (1, 16, 32, 64, 128, grid.dims[0])
    after = vol.read()
    assert np.array_equal(before[0].view(np.uint32), after[0].view(np.uint32))
    assert np.array_equal(before[1].view(np.uint32), after[1].view(np.uint32))
    vol.destroy()
    st.destroy()
