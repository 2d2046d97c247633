"""Seeded random TSDF passes on the GPU against the CPU oracle.

Each case draws: a subset of the small sphere workload's cameras at mixed image sizes (their maps
solved on the GPU and read back), a sparse brick pool (random cubes of bricks marked on the grid,
so the pool order and the brick lookup differ from a dense cube), the voxel size and truncation,
and per view a prior (the true depth with holes, zeros and NaNs) or none and a confidence weight
or none.  The views are integrated in one batched pass on the GPU and one view at a time by the
oracle; weights and update counts must be equal and tsdf values within 1e-4 of the truncation
distance (the TSDF bar; they have been bit-identical in every run), and the GPU's surface nets over
the sparse pool must equal the oracle's on the same volume."""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes

pytestmark = pytest.mark.gpu

CASES = 6


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def base():
    return scenes.sphere_small()


@pytest.mark.parametrize("case", range(CASES))
def test_random_sparse_pass_matches_oracle(ctx, base, case):
    import lr3d
    from lr3d import BrickGrid, IntegrationView, StereoView, scale_intrinsics
    rng = np.random.default_rng(7300 + case)
    n = int(rng.integers(2, 6))
    cams = sorted(rng.choice(len(base.views), n, replace=False).tolist())
    views, images, truths = [], [], []
    for c in cams:
        W, H = int(rng.integers(40, 241)), int(rng.integers(30, 181))
        v = base.views[c]
        sv = StereoView(scale_intrinsics(v.intrinsics, W / v.width, H / v.height), v.pose, W, H)
        g, d = base.scene.render(sv)
        views.append(sv)
        images.append(g.reshape(H, W))
        truths.append(d.astype(np.float32))
    voxel = float(rng.choice([0.006, 0.008, 0.012]))
    trunc = float(rng.choice([3.0, 4.0, 5.0]))
    grid = BrickGrid((-0.512, -0.512, -0.512), voxel, (1.024, 1.024, 1.024))
    for p in ((0.5, 0.0, 0.0), (-0.5, 0.0, 0.0), (0.0, 0.0, 0.5), (0.0, 0.0, -0.5), (0.0, 0.5, 0.0)):
        grid.mark(*p, 0.04)  # patches of the sphere's surface that every camera of the ring sees some of
    for _ in range(int(rng.integers(3, 9))):
        p = rng.uniform(-0.5, 0.5, 3)
        grid.mark(*p, float(rng.uniform(0.05, 0.3)))
    layout = grid.build()
    priors = [None if rng.random() < 0.5 else t.copy() for t in truths]
    for p in priors:
        if p is not None:
            p[rng.random(p.size) < 0.2] = 0.0
            p[rng.random(p.size) < 0.05] = np.nan
    conf = [0.0 if rng.random() < 0.5 else float(rng.uniform(0.5, 2.0)) for _ in range(n)]
    say(f"case {case}: {n} views {[(v.width, v.height) for v in views]}, voxel {voxel * 1000:g} mm, "
        f"truncation {trunc:g}, {layout[2]} of {grid.cells} bricks, priors "
        f"{[p is not None for p in priors]}, confidence {conf}")

    st = lr3d.HipDenseStereo(ctx, views)
    for i, im in enumerate(images):
        st.set_image(i, im)
    st.solve_batched([(r, lr3d.SolveOptions(sources=[j for j in range(n) if j != r], min_depth=0.5,
                                            max_depth=3.0, iterations=1)) for r in range(n)])
    maps = [st.depth_map(r) for r in range(n)]
    ivs = [IntegrationView(v.intrinsics, v.pose, 3.0, v.width, v.height, prior=priors[i], confidence=conf[i])
           for i, v in enumerate(views)]
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, trunc)
    vol.set_counting(True)
    vol.integrate_batched(st, list(range(n)), ivs)
    tsdf, weight = vol.read()
    count = vol.update_count()
    m = vol.extract_surface(0.6)
    mesh = m.read()
    m.destroy()
    vol.destroy()
    st.destroy()

    mu = voxel * trunc
    g = O.Grid(grid.origin, grid.voxel, grid.dims, layout[0], mu)
    otsdf = np.ones(layout[2] * 512, dtype=np.float32)
    oweight = np.zeros(layout[2] * 512, dtype=np.float32)
    updates = 0
    for i, v in enumerate(views):
        k = v.intrinsics
        iv = O.integration_view(v.width, v.height, k.fx, k.fy, k.cx, k.cy, v.pose.R, v.pose.t, 3.0, conf[i])
        updates += O.tsdf_integrate(g, iv, maps[i], priors[i], otsdf, oweight)
    diff = float(np.abs(tsdf - otsdf).max())
    say(f"case {case}: {count} updates (oracle {updates}), max |tsdf diff| {diff:.3g}, "
        f"bit-identical {np.array_equal(tsdf.view(np.uint32), otsdf.view(np.uint32))}")
    assert updates > 0
    assert count == updates
    assert np.array_equal(weight.view(np.uint32), oweight.view(np.uint32))
    assert diff <= 1e-4 * mu
    # surface nets over the sparse pool (the brick lookup) equal the oracle's on the same volume
    ref = O.extract_surface(grid.origin, grid.voxel, grid.dims, layout[1], tsdf, weight, 0.6)
    say(f"case {case}: mesh {len(mesh['positions'])} vertices, {len(mesh['indices']) // 3} triangles")
    for k in ("positions", "normals", "indices"):
        a, r = np.asarray(mesh[k]), np.asarray(ref[k])
        assert a.shape == r.shape and np.array_equal(a.view(np.uint32), r.view(np.uint32)), k
