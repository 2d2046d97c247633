"""Dense-cloud fusion over chunks of views (lr3d_fuse_depth_maps) on the phone set's filtered maps
(32 views 1280x720): the cloud does not depend on how the views are chunked, and equals the
oracle's view-by-view fusion (oracle/fusion_oracle.c).

A chunk's samples are keyed, sorted and added in one pass.  A chunk ends where the reference's
accumulator changes its cells: after the first view that takes it over the budget (it coarsens,
DF:36), or at the sample whose new cell fills its hash table past 60 % (before the next sample
the table doubles and every cell is re-binned by its mean position at the same cell size,
PA:67,87-110, which moves a cell whose float mean rounds outside it; on this set that moves
~100-850 of 650K-750K cells once 16 or more views are in).  LR3D_FUSION_CHUNK_VIEWS=k fixes the
chunks at k views (1: one view per pass, the reference's own loop); unset, they grow from 4 views.
Every chunking must give the same cloud and cell bit for bit: without coarsening, with a budget
that coarsens a few times across the views, and with one so small that the first view alone
coarsens repeatedly."""
import os

import numpy as np
import pytest

import oracle as O
from lr3d import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def phone(ctx):
    import lr3d
    w = scenes.sphere_phone().render()
    n = len(w.views)
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    st.solve_batched([(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                                            iterations=w.iterations, topk=w.topk)) for r in range(n)])
    for r in range(n):
        st.filter(r, w.sources[r], w.mesh_needed(r))
    thumbs = []
    for im in w.images:  # RGB thumbnails of a quarter of the size, channels tinted differently
        g = im[::4, ::4].astype(np.float64)
        thumbs.append(np.stack([g, 0.8 * g + 20, 255 - g], axis=2).clip(0, 255).astype(np.uint8))
    yield w, st, thumbs
    st.destroy()


def _fuse(st, n, thumbs, cell, budget, chunk):
    import lr3d
    old = os.environ.get("LR3D_FUSION_CHUNK_VIEWS")
    if chunk is None:
        os.environ.pop("LR3D_FUSION_CHUNK_VIEWS", None)
    else:
        os.environ["LR3D_FUSION_CHUNK_VIEWS"] = str(chunk)
    try:
        return lr3d.fuse_depth_maps(st, list(range(n)), thumbs, cell, budget)
    finally:
        if old is None:
            os.environ.pop("LR3D_FUSION_CHUNK_VIEWS", None)
        else:
            os.environ["LR3D_FUSION_CHUNK_VIEWS"] = old


def _sorted(c):
    rows = np.concatenate([c["positions"].view(np.uint32), c["normals"].view(np.uint32),
                           c["colors"].astype(np.uint32), c["support"][:, None].astype(np.uint32)], axis=1)
    return rows[np.lexsort(rows.T[::-1])]


@pytest.mark.parametrize("cell,budget", [(0.004, 100_000_000), (0.002, 400_000), (0.001, 2_000)])
def test_every_chunking_gives_the_same_cloud(phone, cell, budget):
    w, st, thumbs = phone
    n = len(w.views)
    say(f"case: fusion of {n} filtered maps 1280x720 at {cell * 1000:g} mm cells, budget {budget} cells, "
        f"chunks of 1, 3, 8 and 32 views and growing chunks")
    ref, rcell = _fuse(st, n, thumbs, cell, budget, 1)
    say(f"one view per pass: {len(ref['support'])} points, final cell {rcell:.6f}")
    for chunk in (3, 8, 32, None):
        got, gcell = _fuse(st, n, thumbs, cell, budget, chunk)
        assert gcell == rcell, chunk
        for k in ("positions", "normals", "colors", "support"):
            assert np.array_equal(got[k], ref[k]), (chunk, k)
    if budget < 100_000_000:
        assert rcell > cell  # the budget did coarsen


def test_growing_chunks_equal_the_oracle_without_coarsening(phone):
    import lr3d
    w, st, thumbs = phone
    n = len(w.views)
    say(f"case: fusion of {n} filtered maps at 4 mm cells in growing chunks against the oracle's view-by-view fusion")
    maps = [st.read_filtered(r) for r in range(n)]
    got, cell = _fuse(st, n, thumbs, 0.004, 100_000_000, None)
    ref, rcell = O.fuse_depth_maps(w.views, maps, thumbs, 0.004, 100_000_000)
    say(f"fusion: {len(got['support'])} points (oracle {len(ref['support'])}), {int(got['support'].sum())} samples")
    assert cell == rcell == 0.004
    assert np.array_equal(_sorted(got), _sorted(ref))


def test_growing_chunks_with_coarsening_match_the_oracle(phone):
    """With coarsening the reference merges cells in hash-table order and this build in key order, so
    merged sums may differ in their last bits: the same bars as the sphere set's coarsening test."""
    from scipy.spatial import cKDTree
    w, st, thumbs = phone
    n = len(w.views)
    say(f"case: fusion of {n} filtered maps at 2 mm cells with a 400000-cell budget against the oracle")
    maps = [st.read_filtered(r) for r in range(n)]
    got, cell = _fuse(st, n, thumbs, 0.002, 400_000, None)
    ref, rcell = O.fuse_depth_maps(w.views, maps, thumbs, 0.002, 400_000)
    k, m = len(got["support"]), len(ref["support"])
    say(f"coarsened: cell {cell:.6f} (oracle {rcell:.6f}), {k} points (oracle {m})")
    assert cell == rcell and cell > 0.002
    assert int(got["support"].sum()) == int(ref["support"].sum())
    assert abs(k - m) <= 0.001 * m
    d, j = cKDTree(ref["positions"].astype(np.float64)).query(got["positions"].astype(np.float64))
    same = (d < 1e-6) & (ref["support"][j] == got["support"]) & np.all(ref["colors"][j] == got["colors"], axis=1)
    say(f"coarsened: {same.mean() * 100:.3f} % of points matched within 1 um")
    assert same.mean() >= 0.999
