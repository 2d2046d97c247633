"""GPU parity of the consistency filter and the TSDF pass at full size, on the phone set (32 views
1280x720, TSDF 256^3 x 4 mm), the mixed-resolution room batch (64 views 640x480 ... 1920x1080,
TSDF 384^3 x 1 cm) and the 200-view room (1280x720, TSDF 512^3 x 8 mm), exactly as the bench step
runs them: batched PatchMatch of every view, the
filter of every view against its sources with the mesh agreement count, one batched TSDF pass of
the filtered maps.

The oracle takes the GPU's raw maps as its input (its own solve of every view would take minutes
on the CPU; the solve itself is compared at full size in test_gpu_parity_full_size.py): every
filtered map must equal the oracle's filter of the same raw maps bit for bit, and the volume must
equal the oracle's one-view-at-a-time integration of the same filtered maps bit for bit (tsdf,
weights and update counts).
"""
import os
import time

import numpy as np
import pytest

import oracle as O
from lr3d import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def run_step(ctx, w):
    import lr3d
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    n = len(w.views)
    st.solve_batched([(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0],
                                            max_depth=w.ranges[r][1], iterations=w.iterations,
                                            topk=w.topk)) for r in range(n)])
    needed = [w.mesh_needed(r) for r in range(n)]
    for r in range(n):
        st.filter(r, w.sources[r], needed[r])
    spec = w.tsdf
    grid, layout = lr3d.dense_grid(spec.origin, spec.voxel, spec.voxels)
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
    iviews = [lr3d.IntegrationView(v.intrinsics, v.pose, w.ranges[i][1], v.width, v.height)
              for i, v in enumerate(w.views)]
    vol.set_counting(True)
    vol.integrate_batched(st, list(range(n)), iviews, map=lr3d.MAP_FILTERED)
    updates = vol.update_count()
    tsdf, weight = vol.read()
    raw = np.concatenate([st.depth_map(r) for r in range(n)])
    filt = [st.read_filtered(r) for r in range(n)]
    vol.destroy()
    st.destroy()
    return grid, layout, raw, filt, needed, tsdf, weight, updates


def check(ctx, w):
    grid, layout, raw, filt, needed, tsdf, weight, updates = run_step(ctx, w)
    O.set_threads(min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    n = len(w.views)
    mismatched = []
    for r in range(n):
        ofilt = O.filter_map(w.views, r, list(w.sources[r]), needed[r], raw)
        if not np.array_equal(ofilt, filt[r]):
            mismatched.append((r, int((ofilt != filt[r]).sum())))
    say(f"filter: {n - len(mismatched)} of {n} filtered maps bit-identical to the oracle "
        f"({time.perf_counter() - t0:.1f} s); mismatches {mismatched[:5]}")
    assert not mismatched
    spec = w.tsdf
    g = O.Grid(grid.origin, grid.voxel, grid.dims, layout[0], spec.voxel * spec.truncation_voxels)
    otsdf = np.ones(layout[2] * 512, dtype=np.float32)
    oweight = np.zeros(layout[2] * 512, dtype=np.float32)
    oupdates = 0
    t0 = time.perf_counter()
    for r, v in enumerate(w.views):
        k = v.intrinsics
        iv = O.integration_view(v.width, v.height, k.fx, k.fy, k.cx, k.cy, v.pose.R, v.pose.t, w.ranges[r][1])
        oupdates += O.tsdf_integrate(g, iv, filt[r], None, otsdf, oweight)
    same_tsdf = np.array_equal(tsdf.view(np.uint32), otsdf.view(np.uint32))
    say(f"tsdf {spec.voxels}^3 x {spec.voxel * 1000:g} mm: updates gpu {updates} oracle {oupdates}, "
        f"weights bit-identical {np.array_equal(weight.view(np.uint32), oweight.view(np.uint32))}, tsdf "
        f"bit-identical {same_tsdf} (oracle {time.perf_counter() - t0:.1f} s)")
    assert updates == oupdates and updates > 0
    assert np.array_equal(weight.view(np.uint32), oweight.view(np.uint32))
    assert same_tsdf


def test_phone_set_filter_and_tsdf_match_oracle(ctx):
    w = scenes.sphere_phone().render()
    say("case: phone set 32 x 1280x720, filter of every view + TSDF 256^3 x 4 mm of the filtered maps")
    check(ctx, w)


def test_room_batch_filter_and_tsdf_match_oracle(ctx):
    w = scenes.room_batched().render()
    say("case: room batch 64 views mixed resolution, filter of every view + TSDF 384^3 x 1 cm of the "
        "filtered maps")
    check(ctx, w)


def test_room_large_filter_and_tsdf_match_oracle(ctx):
    w = scenes.room_large().render()
    say("case: room 200 x 1280x720, filter of every view + TSDF 512^3 x 8 mm of the filtered maps")
    check(ctx, w)
