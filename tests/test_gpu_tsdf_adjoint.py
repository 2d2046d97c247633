"""GPU adjoint of TSDF integration (lr3d_tsdf_backward) against the FP64 CPU reference
(oracle/tsdf_adjoint_oracle.c, pinned by closed forms and central differences in
test_oracle_tsdf_adjoint.py) on the sphere workload: 8 views at 320x240 solved and filtered on the
GPU, TSDF 128^3 x 8 mm, truncation 3 voxels, two views with a prior, four with confidence weighting.

Bars: the replayed tsdf equals lr3d_tsdf_integrate_batched's bit for bit; against FP64 (on voxels
that stay 1e-3 away from every decision boundary in every view -- the nearest-pixel rounding in
pixels, the cam.z > 0.01 and sdf >= -mu tests in units of mu -- so that FP32 and FP64 take the
same branches): tsdf within 1e-4 (in units of the truncation), loss within 1e-5 relative,
per-view pose gradients within 1e-3 of the view's largest entry (a sum over ~6e4 voxels of terms
that carry the FP32 state's rounding; measured 1.5e-4), and per-pixel depth gradients
within 1e-3 + 1e-4 |g| (L per metre) on >= 99.9 % of the pixels that receive one: the FP32 state
carries ~1e-6 of rounding, and dL/dd = 2 (tsdf - target) (...) / mu turns it into ~1e-4 absolute
where tsdf is close to its target."""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def setup(ctx):
    import lr3d
    w = scenes.sphere_small().render()
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    solves = [(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                                    iterations=w.iterations, topk=w.topk)) for r in range(8)]
    st.solve_batched(solves)
    for r in range(8):
        st.filter(r, w.sources[r], w.mesh_needed(r))
    maps = [st.read_filtered(r) for r in range(8)]
    rng = np.random.default_rng(21)
    views = []
    for i, v in enumerate(w.views):
        prior = (w.truth[i] * (1 + 0.01 * rng.standard_normal(w.truth[i].shape))).astype(np.float32) if i < 2 else None
        views.append(lr3d.IntegrationView(v.intrinsics, v.pose, w.ranges[i][1], v.width, v.height,
                                          prior=prior, confidence=1.0 if i >= 4 else 0.0))
    spec = w.tsdf
    grid, layout = lr3d.dense_grid(spec.origin, spec.voxel, spec.voxels)
    mu = spec.voxel * spec.truncation_voxels
    g = O.Grid(grid.origin, grid.voxel, grid.dims, layout[0], mu)
    # seeded voxels: near the surface, plus a few anywhere
    coords = layout[0].reshape(-1, 4)
    cand = rng.integers(0, layout[2] * 512, 2_000_000).astype(np.uint32)
    b, l = cand // 512, cand % 512
    c = coords[b, :3] * 8 + np.stack([l % 8, (l // 8) % 8, l // 64], axis=1)
    sd = w.scene.signed_distance(np.asarray(spec.origin) + (c + 0.5) * spec.voxel)
    pick = np.concatenate([np.flatnonzero(np.abs(sd) < mu)[:60000], np.arange(4000)])
    vox, tgt = cand[pick], np.clip(sd[pick] / mu, -1, 1).astype(np.float32)
    # FP64 reference on the same inputs
    ivs = [O.integration_view(v.width, v.height, v.intrinsics.fx, v.intrinsics.fy, v.intrinsics.cx, v.intrinsics.cy,
                              v.pose.R, v.pose.t, v.max_depth, v.confidence) for v in views]
    decoded = O.decode_maps(maps)
    priors = [None if v.prior is None else v.prior.astype(np.float64).reshape(-1) for v in views]
    ref = O.tsdf_backward64(g, ivs, decoded, priors, vox, tgt.astype(np.float64))
    yield w, st, grid, layout, views, vox, tgt, ref, g, ivs, decoded, priors
    st.destroy()


def test_replay_equals_batched_integration(ctx, setup):
    import lr3d
    w, st, grid, layout, views, vox, tgt, ref, *_ = setup
    say(f"case: TSDF adjoint replay vs integrate_batched, {len(vox)} voxels, 8 views")
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, w.tsdf.truncation_voxels)
    vol.integrate_batched(st, list(range(8)), views, map=lr3d.MAP_FILTERED)
    tsdf, _ = vol.read()
    out = vol.backward(st, list(range(8)), views, vox, tgt, map=lr3d.MAP_FILTERED)
    vol.destroy()
    assert np.array_equal(out["tsdf"].view(np.uint32), tsdf[vox].view(np.uint32))
    loss = float(((out["tsdf"].astype(np.float64) - tgt) ** 2).sum())
    assert out["loss"] == pytest.approx(loss, rel=1e-12)
    say(f"replay: bit-exact, loss {out['loss']:.6f}, updated voxels {(out['tsdf'] != 1).sum()}")


def test_gradients_match_fp64_reference(ctx, setup):
    import lr3d
    w, st, grid, layout, views, vox, tgt, ref, g, ivs, decoded, priors = setup
    keep = ref["kink"] >= 1e-3
    say(f"case: TSDF adjoint vs FP64 reference, {int(keep.sum())} of {len(vox)} voxels clear of decision boundaries")
    vk, tk = vox[keep], tgt[keep]
    ref = O.tsdf_backward64(g, ivs, decoded, priors, vk, tk.astype(np.float64))
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, w.tsdf.truncation_voxels)
    out = vol.backward(st, list(range(8)), views, vk, tk, map=lr3d.MAP_FILTERED)
    vol.destroy()
    assert np.isfinite(out["grad_pose"]).all() and all(np.isfinite(x).all() for x in out["grad_depth"])
    terr = np.abs(out["tsdf"] - ref["tsdf"]).max()
    lerr = abs(out["loss"] - ref["loss"]) / ref["loss"]
    gp, rp = out["grad_pose"], ref["grad_pose"]
    perr = max(np.abs(gp[v] - rp[v]).max() / max(np.abs(rp[v]).max(), 1e-12) for v in range(8))
    got = np.concatenate([x.reshape(-1) for x in out["grad_depth"]]).astype(np.float64)
    exp = np.concatenate(ref["grad_depth"])
    touched = np.flatnonzero(exp != 0)
    rel = np.abs(got[touched] - exp[touched]) / (1e-3 + 1e-4 * np.abs(exp[touched]))
    good = (rel <= 1.0).mean()
    say(f"tsdf max |diff| {terr:.2e}; loss rel {lerr:.2e}; pose max rel {perr:.2e}; "
        f"depth grads {len(touched)} px, {good * 100:.3f} % within the bar, worst {rel.max():.2f} x bar; "
        f"pixels with a gradient only on the GPU {int(((got != 0) & (exp == 0)).sum())}")
    assert terr <= 1e-4
    assert lerr <= 1e-5
    assert perr <= 1e-3
    assert good >= 0.999
