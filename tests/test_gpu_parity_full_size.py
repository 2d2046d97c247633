"""GPU parity at full size: one 1280x720 view of the phone configuration (32 views, 8 sources,
top-3, 8 iterations), one 1920x1080 view of the mixed-resolution room batch (64 views, 4-10
sources) and one view of the 200-view room (1280x720, 4-8 sources, 8 iterations), each solved on
the GPU inside a batch and by the CPU oracle on the same seeded inputs.

Bars as in test_gpu_parity.py: planes and costs bit for bit on >= 99.9 % of pixels, packed depths
within 1e-3 relative on >= 99.5 % of valid pixels."""
import os
import time

import numpy as np
import pytest

import oracle as O
from lr3d import scenes
from test_gpu_parity import depth_agreement

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def check_view(ctx, w, batch, ref):
    import lr3d
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    solves = [(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                                    iterations=w.iterations, topk=w.topk)) for r in batch]
    st.solve_batched(solves)
    hyp, cost = st.hypotheses(ref)
    packed = st.depth_map(ref)
    st.destroy()
    O.set_threads(min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    imgs = O.Images(w.images)
    opacked, ohyp, ocost = O.solve(w.views, imgs, ref, w.sources[ref], *w.ranges[ref], iterations=w.iterations,
                                   topk=w.topk)
    v = w.views[ref]
    exact = (np.all(hyp.view(np.uint32) == ohyp.view(np.uint32), axis=1) &
             (cost.view(np.uint32) == ocost.view(np.uint32))).mean()
    agree = depth_agreement(packed, O.layer_map(w.views, opacked, ref))
    say(f"view {ref} ({v.width}x{v.height}, {len(w.sources[ref][:8])} sources): planes and costs bit-exact on "
        f"{exact * 100:.4f} % of pixels, depth agreement {agree * 100:.4f} % (oracle {time.perf_counter() - t0:.1f} s)")
    assert exact >= 0.999 and agree >= 0.995


def test_phone_view_matches_oracle(ctx):
    w = scenes.sphere_phone().render()
    say("case: phone set 32 x 1280x720, batched solve of views 0, 5, 17; view 0 against the oracle")
    check_view(ctx, w, [0, 5, 17], 0)


def test_room_full_hd_view_matches_oracle(ctx):
    w = scenes.room_batched().render()
    ref = 3  # 1920x1080
    assert (w.views[ref].width, w.views[ref].height) == (1920, 1080)
    say("case: room batch 64 views mixed resolution, batched solve of all views; view 3 (1920x1080) "
        "against the oracle")
    check_view(ctx, w, list(range(len(w.views))), ref)


def test_room_large_view_matches_oracle(ctx):
    w = scenes.room_large().render()
    ref = 117
    say("case: room 200 x 1280x720 (the bench workload), batched solve of all 200 views; view 117 "
        "against the oracle")
    check_view(ctx, w, list(range(len(w.views))), ref)
