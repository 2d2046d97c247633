"""The reference's PatchMatch known-answer check (frontend/tools/verify/stereo-check.ts:11-72) in
its own configuration, on the GPU: 5 views at 640x360 of a textured box and sphere on a table,
view 2 solved from views 0, 1, 3, 4 over 0.1-2 m with 5 iterations.  The reference passes when
more than 60 % of the surface pixels get a depth and the median relative error is below 0.4 %
(stereo-check.ts:69-72; published: 100 % and 0.04 %, frontend/README.md:262-263).  The GPU planes
and costs must also equal the oracle's bit for bit."""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes
from lr3d.packing import unpack_depth

pytestmark = pytest.mark.gpu


def test_known_answer_configuration(ctx):
    import lr3d
    w = scenes.reference_check().render()
    print("case: reference known-answer configuration, view 2 of 5 x 640x360", flush=True)
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    o = lr3d.SolveOptions(sources=w.sources[2], min_depth=0.1, max_depth=2.0, iterations=5)
    st.solve_batched([(2, o)])
    packed = st.depth_map(2)
    hyp, cost = st.hypotheses(2)
    st.destroy()
    truth = w.truth[2].reshape(-1)
    covered = truth > 0
    valid = covered & (packed != 0)
    err = np.abs(unpack_depth(packed[valid]) - truth[valid]) / truth[valid]
    frac, median = valid.sum() / covered.sum(), float(np.median(err))
    print(f"{frac * 100:.1f} % of surface pixels valid, median relative error {median * 100:.3f} % "
          f"(reference: > 60 % and < 0.4 % to pass; published 100 % and 0.04 %)", flush=True)
    assert frac > 0.6 and median < 0.004
    _, ohyp, ocost = O.solve(w.views, O.Images(w.images), 2, w.sources[2], 0.1, 2.0, iterations=5)
    assert np.array_equal(hyp.view(np.uint32), ohyp.view(np.uint32))
    assert np.array_equal(cost.view(np.uint32), ocost.view(np.uint32))
