"""The geometric (ACMM) pass over every view of a set, batched, against single solves and the
oracle, bit for bit.

The reference runs the geometric pass one layer after another (…/pipeline/reconstructor.ts:96-108):
layer L's solve reads the packed maps that layers < L have just rewritten (Gauss-Seidel over the
views).  lr3d_stereo_solve_batched promises the result of n single solves in array order, so a
batched geometric pass must equal (a) the same solves called one by one and (b) the oracle's
geometric solves run in the same order over one map pool.  The oracle is pinned against the
Jacobi order as well: running every geometric solve against the maps from before the pass gives
other planes, so a batch that read stale maps would fail here.

Workloads: the small sphere configuration (8 views 320x240, 4 sources, top-2) with all views in
order and in reverse order, and the batched room configuration at quarter scale (64 views at
160x120, 240x135, 320x180 and 480x270 with 4-8 sources each), whose geometric pass and
consistency filter mix layer sizes, compared on every view.
"""
import sys

import numpy as np
import pytest

import oracle as O
from lr3d import HipDenseStereo, SolveOptions, scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def options(work, ref, iterations, geometric):
    lo, hi = work.ranges[ref]
    return SolveOptions(sources=work.sources[ref], min_depth=lo, max_depth=hi, iterations=iterations,
                        topk=work.topk, geometric=geometric)


def stereo_of(ctx, work):
    st = HipDenseStereo(ctx, work.views)
    for i, im in enumerate(work.images):
        st.set_image(i, im)
    return st


def oracle_passes(work, photometric_its, geometric_its, order, jacobi=False):
    """Photometric solves of every view, then the geometric pass in `order` over the map pool.
    jacobi: every geometric solve reads the maps from before the pass (the order the batch must
    not use)."""
    imgs = O.Images(work.images)
    depths = O.map_pool(work.views)
    for r in range(len(work.views)):
        O.solve(work.views, imgs, r, work.sources[r], *work.ranges[r], iterations=photometric_its,
                topk=work.topk, depths=depths)
    before = depths.copy()
    planes = {}
    for r in order:
        src = before.copy() if jacobi else depths
        _, hyp, cost = O.solve(work.views, imgs, r, work.sources[r], *work.ranges[r], geometric=True,
                               iterations=geometric_its, topk=work.topk, depths=src)
        if jacobi:
            O.layer_map(work.views, depths, r)[:] = O.layer_map(work.views, src, r)
        planes[r] = (hyp.copy(), cost.copy())
    return depths, planes


def compare(st, work, depths, planes, label):
    for r in range(len(work.views)):
        hyp, cost = st.hypotheses(r)
        ohyp, ocost = planes[r]
        assert np.array_equal(hyp.view(np.uint32), ohyp.view(np.uint32)), f"{label}: planes of view {r}"
        assert np.array_equal(cost.view(np.uint32), ocost.view(np.uint32)), f"{label}: costs of view {r}"
        assert np.array_equal(st.depth_map(r), O.layer_map(work.views, depths, r)), f"{label}: map of view {r}"


@pytest.mark.parametrize("order", ["forward", "reverse"])
def test_batched_geometric_pass_is_gauss_seidel(ctx, order):
    work = scenes.sphere_small().render()
    n = len(work.views)
    seq = list(range(n)) if order == "forward" else list(range(n - 1, -1, -1))
    say(f"case: geometric pass of all {n} views of {work.name}, {order} order, batched vs single vs oracle")
    photo = [(r, options(work, r, 2, False)) for r in range(n)]
    geo = [(r, options(work, r, 2, True)) for r in seq]
    batched = stereo_of(ctx, work)
    batched.solve_batched(photo)
    batched.solve_batched(geo)
    single = stereo_of(ctx, work)
    for r, o in photo:
        single.solve(r, o)
    for r, o in geo:
        single.solve(r, o)
    depths, planes = oracle_passes(work, 2, 2, seq)
    compare(batched, work, depths, planes, "batched")
    compare(single, work, depths, planes, "single")
    say(f"{order}: batched == single solves == oracle on every pixel of {n} views")
    # the Jacobi order changes planes: a batch reading the maps from before the pass would fail
    jdepths, jplanes = oracle_passes(work, 2, 2, seq, jacobi=True)
    changed = sum(int(np.any(jplanes[r][0].view(np.uint32) != planes[r][0].view(np.uint32))) for r in seq[1:])
    say(f"{order}: the Jacobi order gives other planes on {changed} of {n - 1} views after the first")
    assert changed >= n // 2
    assert np.array_equal(jplanes[seq[0]][0].view(np.uint32), planes[seq[0]][0].view(np.uint32))
    batched.destroy()
    single.destroy()


def test_mixed_resolution_geometric_pass_and_filter_match_oracle(ctx):
    work = scenes.room_batched(scale=0.25, iterations=2).render()
    n = len(work.views)
    say(f"case: {work.name}, {n} views of 4 sizes, {work.megapixels:.2f} Mpx: photometric (2 its), "
        f"geometric (1 it) and filter, batched vs oracle")
    st = stereo_of(ctx, work)
    st.solve_batched([(r, options(work, r, 2, False)) for r in range(n)])
    st.solve_batched([(r, options(work, r, 1, True)) for r in range(n)])
    depths, planes = oracle_passes(work, 2, 1, list(range(n)))
    compare(st, work, depths, planes, "room")
    kept = 0
    for r in range(n):
        needed = work.mesh_needed(r)
        st.filter(r, work.sources[r], needed)
        ofilt = O.filter_map(work.views, r, work.sources[r][:8], needed, depths)
        assert np.array_equal(st.read_filtered(r), ofilt), f"filtered map of view {r}"
        kept += int(np.count_nonzero(ofilt))
    say(f"room: planes, costs, packed and filtered maps of all {n} views equal the oracle's "
        f"({kept} filtered depths kept)")
    assert kept > 0
    st.destroy()
