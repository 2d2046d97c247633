"""The order of the source views changes no result of a solve.

`hypothesis_cost` (PM:192-233) sums the k smallest view costs after sorting them ascending, so
the aggregate of a hypothesis depends on the set of view costs, not on the order the views are
visited in; the random draws (PM:235-245, PM:401-420) do not depend on the sources at all.  The
HIP propagate kernel relies on this when it visits the sources last to first, so that the
candidate lower bound rises sooner.  Here the oracle solves one view with its source list in
three orders, with top-k below and at the source count and with the geometric term (which adds
a per-view reprojection error before the selection), and every plane, cost and packed sample
must be identical bit for bit.  A selection that summed costs in visiting order, or kept the
first k instead of the k smallest, would fail: the sums would round differently or pick other
views."""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes


@pytest.fixture(scope="module")
def work():
    return scenes.sphere_small().render()


def solve(work, sources, topk, depths=None, geometric=False):
    imgs = O.Images(work.images)
    packed, hyp, cost = O.solve(work.views, imgs, 0, sources, *work.ranges[0], iterations=2,
                                topk=topk, depths=depths, geometric=geometric)
    return packed.copy(), hyp.copy(), cost.copy()


def same(a, b):
    return all(np.array_equal(np.asarray(x).view(np.uint32), np.asarray(y).view(np.uint32)) for x, y in zip(a, b))


@pytest.mark.parametrize("topk", [2, 4])
def test_source_order_changes_no_result(work, topk):
    sources = list(work.sources[0])
    assert len(sources) == 4
    base = solve(work, sources, topk)
    assert np.isfinite(base[2]).all() and (base[2] < 2.0).mean() > 0.5
    for order in (sources[::-1], [sources[2], sources[0], sources[3], sources[1]]):
        assert same(base, solve(work, order, topk)), order


def test_source_order_changes_no_result_with_the_geometric_term(work):
    W, H = work.views[0].width, work.views[0].height
    depths = np.zeros(len(work.views) * W * H, dtype=np.uint32)
    imgs = O.Images(work.images)
    for ref in range(len(work.views)):
        O.solve(work.views, imgs, ref, work.sources[ref], *work.ranges[ref], iterations=1,
                topk=work.topk, depths=depths)
    sources = list(work.sources[0])
    base = solve(work, sources, work.topk, depths=depths.copy(), geometric=True)
    assert same(base, solve(work, sources[::-1], work.topk, depths=depths.copy(), geometric=True))
