"""Calls from several host threads: the C ABI serialises calls on one context with its mutex and
keeps contexts independent, so concurrent use must give the same bits as sequential use.

Two cases on the 8-view 320x240 sphere set, each thread running the whole path (batched solve of
every view, filter of every view, one batched TSDF pass of the filtered maps, read-back):
  - four threads, each with its own context, stereo set and volume;
  - four threads sharing one context, each with its own stereo set and volume.
Every thread's maps and volume must equal those of a single-threaded run.  ctypes releases the
interpreter lock during each foreign call, so the library calls overlap.
"""
import hashlib
import threading

import numpy as np
import pytest

from lr3d import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def work():
    return scenes.sphere_small().render()


def pipeline(ctx, w, repeats=2):
    from lr3d import HipDenseStereo, HipTsdfVolume, IntegrationView, SolveOptions, MAP_FILTERED, dense_grid
    st = HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    n = len(w.views)
    solves = [(r, SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                               iterations=w.iterations, topk=w.topk)) for r in range(n)]
    grid, layout = dense_grid(w.tsdf.origin, w.tsdf.voxel, w.tsdf.voxels)
    views = [IntegrationView(v.intrinsics, v.pose, w.ranges[i][1], v.width, v.height) for i, v in enumerate(w.views)]
    digests = []
    for _ in range(repeats):
        st.solve_batched(solves)
        for r in range(n):
            st.filter(r, w.sources[r], w.mesh_needed(r))
        vol = HipTsdfVolume(ctx, grid, layout, w.tsdf.truncation_voxels)
        vol.integrate_batched(st, list(range(n)), views, map=MAP_FILTERED)
        h = hashlib.sha256()
        for r in range(n):
            h.update(st.depth_map(r).tobytes())
            h.update(st.read_filtered(r).tobytes())
        for a in vol.read():
            h.update(np.ascontiguousarray(a).tobytes())
        vol.destroy()
        digests.append(h.hexdigest())
    st.destroy()
    return digests


def run_threads(target, count):
    results, errors = [None] * count, []

    def body(i):
        try:
            results[i] = target(i)
        except Exception as e:  # reported by the test, not swallowed
            errors.append(repr(e))

    threads = [threading.Thread(target=body, args=(i,)) for i in range(count)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors
    return results


def test_threads_with_their_own_contexts(ctx, work):
    import lr3d
    say("case: 4 threads, one context each, sphere 8x320x240 solve + filter + TSDF, twice per thread")
    reference = pipeline(ctx, work, repeats=1)[0]

    def one(i):
        c = lr3d.Context(0)
        try:
            return pipeline(c, work)
        finally:
            c.destroy()

    results = run_threads(one, 4)
    assert all(d == reference for r in results for d in r)


def test_threads_sharing_one_context(ctx, work):
    say("case: 4 threads on one context, sphere 8x320x240 solve + filter + TSDF, twice per thread")
    reference = pipeline(ctx, work, repeats=1)[0]
    results = run_threads(lambda i: pipeline(ctx, work), 4)
    assert all(d == reference for r in results for d in r)
