"""The allocation-free batched solve (lr3d_stereo_solve_batched_noalloc) against the existing
batched entry point (lr3d_stereo_solve_batched), bit for bit, through the C ABI.

Two calls in a row on the same pair of sets: first every view of the small sphere workload
(photometric, 2 iterations, priors on two views), then a smaller batch (three views, a geometric
pass reading the first call's maps, in a non-sequential order).  Planes, costs and packed maps of
every layer must be identical after each call; the second call reuses the scratch array the first
one filled.  Rejections: a solve with a prior on a set without a prior pool, and more solves than
the scratch array holds; the set stays usable after both."""
import numpy as np
import pytest

from lr3d import HipDenseStereo, Lr3dError, SolveOptions, scenes
from lr3d import _native as N

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def options(work, r, iterations, geometric=False, prior=None):
    lo, hi = work.ranges[r]
    return SolveOptions(sources=work.sources[r], min_depth=lo, max_depth=hi, iterations=iterations,
                        topk=work.topk, geometric=geometric, prior=prior)


def stereo_of(ctx, work, prior_pool=False):
    st = HipDenseStereo(ctx, work.views, prior_pool=prior_pool)
    for i, im in enumerate(work.images):
        st.set_image(i, im)
    return st


def assert_same(a, b, n, label):
    for r in range(n):
        ha, ca = a.hypotheses(r)
        hb, cb = b.hypotheses(r)
        assert np.array_equal(ha.view(np.uint32), hb.view(np.uint32)), f"{label}: planes of view {r}"
        assert np.array_equal(ca.view(np.uint32), cb.view(np.uint32)), f"{label}: costs of view {r}"
        assert np.array_equal(a.depth_map(r), b.depth_map(r)), f"{label}: packed map of view {r}"


def test_noalloc_solve_equals_the_batched_solve_on_two_calls(ctx):
    work = scenes.sphere_small().render()
    n = len(work.views)
    rng = np.random.default_rng(17)
    priors = {}
    for r in (0, 5):  # the analytic depth with holes, as a phone depth prior
        p = work.truth[r].astype(np.float32).copy()
        p[rng.random(p.shape) < 0.3] = 0.0
        priors[r] = p
    first = [(r, options(work, r, 2, prior=priors.get(r))) for r in range(n)]
    second = [(r, options(work, r, 1, geometric=True)) for r in (6, 2, 4)]
    a = stereo_of(ctx, work)
    b = stereo_of(ctx, work, prior_pool=True)
    say(f"case: no-alloc call 1, {len(first)} solves of {work.name} (priors on views 0 and 5)")
    a.solve_batched(first)
    b.solve_batched_noalloc(first)
    assert_same(a, b, n, "call 1")
    say(f"case: no-alloc call 2, {len(second)} geometric solves (views 6, 2, 4)")
    a.solve_batched(second)
    b.solve_batched_noalloc(second)
    assert_same(a, b, n, "call 2")
    say("no-alloc: planes, costs and packed maps equal the batched entry point's after both calls")
    a.destroy()
    b.destroy()


def test_noalloc_solve_rejects_what_it_cannot_hold(ctx):
    work = scenes.sphere_small().render()
    st = stereo_of(ctx, work)
    prior = work.truth[0].astype(np.float32)
    say("case: no-alloc rejections (prior without a prior pool, too many solves)")
    with pytest.raises(Lr3dError) as e:
        st.solve_batched_noalloc([(0, options(work, 0, 1, prior=prior))])
    assert e.value.code == -4 and "LR3D_STEREO_PRIOR_POOL" in str(e.value)
    many = [(r % len(work.views), options(work, r % len(work.views), 1)) for r in range(N.NOALLOC_MAX_SOLVES + 1)]
    with pytest.raises(Lr3dError) as e:
        st.solve_batched_noalloc(many)
    assert e.value.code == -1 and str(N.NOALLOC_MAX_SOLVES) in str(e.value)
    # still usable: the same solve through both entry points
    ref = stereo_of(ctx, work)
    solves = [(r, options(work, r, 1)) for r in (3, 1)]
    st.solve_batched_noalloc(solves)
    ref.solve_batched(solves)
    assert_same(st, ref, len(work.views), "after rejections")
    st.destroy()
    ref.destroy()
