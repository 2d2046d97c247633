"""GPU dense-cloud fusion (lr3d_fuse_depth_maps) against the oracle's restatement of the reference's
fuseDepthMaps + PointAccumulator (oracle/fusion_oracle.c) on the sphere workload's filtered maps.

Without coarsening every cell's sums equal the reference's (float sums in sample order), so the
clouds are the same set of points bit for bit (compared after sorting, since the reference's
points follow its hash table).  With coarsening, the reference merges cells in hash-table order
and this build in cell-key order: the merged float sums may differ in their last bits, so the bar
there is the same total support, the same number of points within 0.1 %, and >= 99.9 % of the
points within 1 um of an oracle point with the same support and colour."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle as O
from lr3d import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def fused(ctx):
    import lr3d
    w = scenes.sphere_small().render()
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    st.solve_batched([(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                                            iterations=w.iterations, topk=w.topk)) for r in range(8)])
    for r in range(8):
        st.filter(r, w.sources[r], w.mesh_needed(r))
    maps = [st.read_filtered(r) for r in range(8)]
    rng = np.random.default_rng(8)
    thumbs = []
    for im in w.images:  # an RGB thumbnail of a quarter of the size, channels tinted differently
        g = im[::4, ::4].astype(np.float64)
        thumbs.append(np.stack([g, 0.8 * g + 20, 255 - g], axis=2).clip(0, 255).astype(np.uint8))
    yield w, st, maps, thumbs
    st.destroy()


def _sorted(c):
    rows = np.concatenate([c["positions"].view(np.uint32), c["normals"].view(np.uint32),
                           c["colors"].astype(np.uint32), c["support"][:, None].astype(np.uint32)], axis=1)
    return rows[np.lexsort(rows.T[::-1])]


def test_fused_cloud_equals_oracle_without_coarsening(fused):
    import lr3d
    w, st, maps, thumbs = fused
    say("case: dense-cloud fusion of 8 filtered maps at 4 mm cells, no coarsening")
    got, cell = lr3d.fuse_depth_maps(st, list(range(8)), thumbs, 0.004, 10_000_000)
    ref, rcell = O.fuse_depth_maps(w.views, maps, thumbs, 0.004, 10_000_000)
    say(f"fusion: {len(got['support'])} points (oracle {len(ref['support'])}), {int(got['support'].sum())} samples")
    assert cell == rcell == 0.004
    assert np.array_equal(_sorted(got), _sorted(ref))


def test_fused_cloud_with_coarsening(fused):
    import lr3d
    w, st, maps, thumbs = fused
    say("case: dense-cloud fusion with a 20000-cell budget (coarsening)")
    got, cell = lr3d.fuse_depth_maps(st, list(range(8)), thumbs, 0.002, 20000)
    ref, rcell = O.fuse_depth_maps(w.views, maps, thumbs, 0.002, 20000)
    n, m = len(got["support"]), len(ref["support"])
    say(f"coarsened: cell {cell:.5f} (oracle {rcell:.5f}), {n} points (oracle {m})")
    assert cell == rcell and cell > 0.002
    assert int(got["support"].sum()) == int(ref["support"].sum())
    assert abs(n - m) <= 0.001 * m
    d, j = cKDTree(ref["positions"].astype(np.float64)).query(got["positions"].astype(np.float64))
    same = (d < 1e-6) & (ref["support"][j] == got["support"]) & np.all(ref["colors"][j] == got["colors"], axis=1)
    say(f"coarsened: {same.mean() * 100:.3f} % of points matched within 1 um")
    assert same.mean() >= 0.999
