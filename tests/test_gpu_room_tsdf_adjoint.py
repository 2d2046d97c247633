"""Room-scale forward and TSDF adjoint: 200 views at 1280x720 of the procedural room with poses
noised by seeded noise (0.5 deg RMS rotation, 5 mm RMS translation; view 0 exact), batched
PatchMatch of all 200 views (8 iterations), one batched integration of the packed maps into a
512^3 x 8 mm volume (1.07 GB of tsdf + weight), then the adjoint (lr3d_tsdf_backward) of
L = sum (tsdf - target)^2 over 1M seeded voxels near the surfaces, target = the analytic signed
distance over the truncation, clamped to [-1, 1].

Checks: the adjoint's forward replay equals the integrated volume bit for bit at the 1M voxels;
on the voxels that stay 1e-3 away from every decision boundary in every view (the nearest-pixel
rounding in pixels, the cam.z > 0.01 and sdf >= -mu tests in units of mu: FP32 and FP64 then take
the same branches), the GPU gradients against the FP64 CPU reference
(oracle/tsdf_adjoint_oracle.c) with the bars of the small TSDF adjoint test (the pose bar 1e-3 of
the view's largest entry on 99 % of the views, 1e-2 on all), and against FP64
central differences on 256 samples: 224 depth pixels (8 seeded views x 28) and 32 pose entries
(4 seeded views x 8 entries of R and t).  A difference quotient is used where steps h and h/2
agree (no jump or kink inside the step); it sums the loss over the voxels the sample can move
(voxels within 1.5 px of the pixel, which is the whole loss's difference quotient since the other
voxels do not depend on the pixel; for a pose entry, the voxels of bricks the view's culling keeps
less those within 1e-5 m of a skip threshold, with the GPU gradient taken over the same voxels)."""
import time

import numpy as np
import pytest

import oracle as O
from lr3d import scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def fd(f, h, agree=1e-5):
    """Central difference of a sum of per-voxel terms, differenced voxel by voxel (voxels the step
    leaves alone cancel exactly instead of adding the rounding of the whole sum), checked against
    the quotient with h/2."""
    def diff(a, b):
        return float(np.sum(np.asarray(a) - np.asarray(b)))
    g1 = diff(f(h), f(-h)) / (2 * h)
    g2 = diff(f(h / 2), f(-h / 2)) / h
    if not (np.isfinite(g1) and np.isfinite(g2)) or abs(g1 - g2) > agree * max(1.0, abs(g1)):
        return np.nan
    return g1


@pytest.fixture(scope="module")
def room(ctx):
    import lr3d
    t0 = time.perf_counter()
    w = scenes.room_large().render()
    noisy = scenes.noisy_views(w.views, seed=5)
    n = len(w.views)
    say(f"room-200x1280x720 rendered in {time.perf_counter() - t0:.1f} s")
    st = lr3d.HipDenseStereo(ctx, noisy)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    solves = [(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                                    iterations=w.iterations, topk=w.topk)) for r in range(n)]
    ctx.synchronize()
    t0 = time.perf_counter()
    st.solve_batched(solves)
    ctx.synchronize()
    t_pm = time.perf_counter() - t0
    spec = w.tsdf
    grid, layout = lr3d.dense_grid(spec.origin, spec.voxel, spec.voxels)
    views = [lr3d.IntegrationView(v.intrinsics, v.pose, w.ranges[i][1], v.width, v.height) for i, v in enumerate(noisy)]
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
    ctx.set_profiling(True)
    vol.integrate_batched(st, list(range(n)), views, map=lr3d.MAP_RAW)
    t_int = ctx.kernel_stats("tsdf_integrate")[1]
    ctx.set_profiling(False)
    say(f"forward: batched PatchMatch of 200 views {t_pm:.2f} s ({w.megapixels * 113 / t_pm:.0f} Mpx*hyp/s); "
        f"integration into 512^3 x 8 mm {t_int:.2f} ms")
    mu = spec.voxel * spec.truncation_voxels
    rng = np.random.default_rng(55)
    cand = np.unique(rng.integers(0, layout[2] * 512, 8_000_000).astype(np.uint32))
    rng.shuffle(cand)
    coords = layout[0].reshape(-1, 4)
    b, l = cand // 512, cand % 512
    cell = coords[b, :3] * 8 + np.stack([l % 8, (l // 8) % 8, l // 64], axis=1)
    sd = w.scene.signed_distance(np.asarray(spec.origin) + (cell + 0.5) * spec.voxel)
    pick = np.flatnonzero(np.abs(sd) < 2 * mu)[:1_000_000]
    assert len(pick) == 1_000_000
    vox, tgt = cand[pick], np.clip(sd[pick] / mu, -1, 1).astype(np.float32)
    g = O.Grid(grid.origin, grid.voxel, grid.dims, layout[0], mu)
    ivs = [O.integration_view(v.width, v.height, v.intrinsics.fx, v.intrinsics.fy, v.intrinsics.cx, v.intrinsics.cy,
                              v.pose.R, v.pose.t, v.max_depth, v.confidence) for v in views]
    maps = [st.depth_map(r) for r in range(n)]
    yield w, noisy, st, vol, grid, layout, views, vox, tgt, g, ivs, maps
    vol.destroy()
    st.destroy()


def test_replay_equals_the_integrated_volume(ctx, room):
    import lr3d
    w, noisy, st, vol, grid, layout, views, vox, tgt, *_ = room
    say("case: TSDF adjoint over 1M voxels x 200 views, replay vs the integrated 512^3 volume")
    ctx.set_profiling(True)
    out = vol.backward(st, list(range(len(views))), views, vox, tgt, map=lr3d.MAP_RAW)
    launches, ms = ctx.kernel_stats("tsdf_backward")
    ctx.set_profiling(False)
    tsdf, _ = vol.read()
    assert np.array_equal(out["tsdf"].view(np.uint32), tsdf[vox].view(np.uint32))
    assert np.isfinite(out["grad_pose"]).all() and all(np.isfinite(x).all() for x in out["grad_depth"])
    loss = float(((out["tsdf"].astype(np.float64) - tgt) ** 2).sum())
    assert out["loss"] == pytest.approx(loss, rel=1e-12)
    say(f"backward: {ms / max(launches, 1):.2f} ms for 1M voxels x 200 views; loss {out['loss']:.2f}; "
        f"{int((out['tsdf'] != 1).sum())} voxels updated")


def test_gradients_match_fp64_reference_and_central_differences(ctx, room):
    import lr3d
    w, noisy, st, vol, grid, layout, views, vox, tgt, g, ivs, maps = room
    n = len(views)
    t0 = time.perf_counter()
    decoded = O.decode_maps(maps)
    full = O.tsdf_backward64(g, ivs, decoded, None, vox, tgt.astype(np.float64))
    keep = full["kink"] >= 1e-3
    vk, tk = vox[keep], tgt[keep].astype(np.float64)
    ref = O.tsdf_backward64(g, ivs, decoded, None, vk, tk)
    say(f"case: TSDF adjoint vs FP64, {len(vk)} of {len(vox)} voxels clear of decision boundaries "
        f"(FP64 reference {time.perf_counter() - t0:.1f} s)")
    out = vol.backward(st, list(range(n)), views, vk, tk.astype(np.float32), map=lr3d.MAP_RAW)
    assert np.isfinite(out["grad_pose"]).all() and all(np.isfinite(x).all() for x in out["grad_depth"])
    terr = np.abs(out["tsdf"] - ref["tsdf"]).max()
    lerr = abs(out["loss"] - ref["loss"]) / ref["loss"]
    gp, rp = out["grad_pose"], ref["grad_pose"]
    live = [v for v in range(n) if np.abs(rp[v]).max() > 0]
    # per view, relative to the view's largest entry: a sum over ~1e5 voxels of terms that carry
    # the FP32 state's rounding (~1e-6 absolute in tsdf, so ~1e-3 relative in a small residual),
    # which cancel less well in views with a small total
    verr = np.array([np.abs(gp[v] - rp[v]).max() / np.abs(rp[v]).max() for v in live])
    perr = float(np.percentile(verr, 99))
    got = np.concatenate([x.reshape(-1) for x in out["grad_depth"]]).astype(np.float64)
    exp = np.concatenate(ref["grad_depth"])
    touched = np.flatnonzero(exp != 0)
    rel = np.abs(got[touched] - exp[touched]) / (1e-3 + 1e-4 * np.abs(exp[touched]))
    good = (rel <= 1.0).mean()
    say(f"vs FP64: tsdf max |diff| {terr:.2e}; loss rel {lerr:.2e}; pose rel per view over {len(live)} views: "
        f"median {np.median(verr):.2e}, 99th percentile {perr:.2e}, max {verr.max():.2e}; "
        f"depth grads {len(touched)} px, {good * 100:.3f} % within 1e-3 + 1e-4|g|")
    assert terr <= 1e-4 and lerr <= 1e-5 and good >= 0.999
    assert np.median(verr) <= 1e-4 and perr <= 1e-3 and verr.max() <= 1e-2

    # central differences on 256 samples
    rng = np.random.default_rng(56)
    spec = w.tsdf
    coords = layout[0].reshape(-1, 4)
    b, l = vk // 512, vk % 512
    world = np.asarray(spec.origin) + (coords[b, :3] * 8 + np.stack([l % 8, (l // 8) % 8, l // 64], axis=1) + 0.5) * spec.voxel
    vis = O.visibility_flags(g, ivs)
    depth_err, pose_err = [], []
    sizes = [v.width * v.height for v in views]
    live_depth = [v for v in range(n) if np.count_nonzero(ref["grad_depth"][v]) >= 200]
    for v in rng.choice(live_depth, 8, replace=False):
        iv = ivs[v]
        R = np.array(list(iv.R)).reshape(3, 3)
        t = np.array(list(iv.t))
        cam = world @ R.T + t
        front = cam[:, 2] > 0.01
        u = np.where(front, iv.fx * cam[:, 0] / np.where(front, cam[:, 2], 1) + iv.cx, -1e9)
        q = np.where(front, iv.fy * cam[:, 1] / np.where(front, cam[:, 2], 1) + iv.cy, -1e9)
        gpx = np.flatnonzero(ref["grad_depth"][v])
        taken = 0
        for p in rng.permutation(gpx):
            py, px = divmod(int(p), iv.width)
            near = np.flatnonzero((np.abs(u - px) <= 1.5) & (np.abs(q - py) <= 1.5) & (vis[v][b] != 0))
            if len(near) == 0 or len(near) > 64:
                continue

            def loss(h, v=v, p=p, near=near):
                inp = list(decoded)
                d, nn, c = inp[v]
                d = d.copy()
                d[p] += h
                inp[v] = (d, nn, c)
                return O.tsdf_loss64(g, ivs, inp, None, vk[near], tk[near], vis=vis, per_voxel=True)
            f = fd(loss, 1e-7)
            if not np.isfinite(f):
                continue
            gv = float(out["grad_depth"][v].reshape(-1)[p])
            depth_err.append(abs(gv - f) / (1e-3 + 1e-4 * abs(f)))
            taken += 1
            if taken == 28:
                break
    # Pose entries: over the voxels of bricks the view's culling keeps, leaving out those within
    # 1e-5 m of a skip threshold of that view (cam.z = 0.01, sdf = -mu), where a step of the pose
    # switches an update on or off (a jump, not a kink); the GPU gradient is taken over the same set.
    live_pose = [v for v in live if v != 0]
    for v in rng.choice(live_pose, 4, replace=False):
        iv = ivs[v]
        R = np.array(list(iv.R)).reshape(3, 3)
        t = np.array(list(iv.t))
        cam = world @ R.T + t
        front = cam[:, 2] > 0.01
        z = np.where(front, cam[:, 2], 1.0)
        ru = np.rint(iv.fx * cam[:, 0] / z + iv.cx)
        rq = np.rint(iv.fy * cam[:, 1] / z + iv.cy)
        inside = front & (ru >= 0) & (rq >= 0) & (ru < iv.width) & (rq < iv.height)
        pix = np.where(inside, rq * iv.width + ru, 0).astype(np.int64)
        d = np.where(inside, decoded[v][0][pix], 0.0)
        ok = (vis[v][b] != 0) & (np.abs(cam[:, 2] - 0.01) > 1e-5)
        ok &= ~inside | (d <= 0) | (np.abs(d - cam[:, 2] + spec.voxel * spec.truncation_voxels) > 1e-5)
        sub = np.flatnonzero(ok)
        gv = vol.backward(st, list(range(n)), views, vk[sub], tk[sub].astype(np.float32),
                          map=lr3d.MAP_RAW)["grad_pose"][v]
        vmax = np.abs(gv).max()
        for k in (0, 2, 4, 6, 8, 9, 10, 11):
            def loss(h, v=v, k=k, sub=sub):
                iv = O.integration_view(ivs[v].width, ivs[v].height, ivs[v].fx, ivs[v].fy, ivs[v].cx, ivs[v].cy,
                                        np.array(list(ivs[v].R)).reshape(3, 3), np.array(list(ivs[v].t)),
                                        ivs[v].max_depth, ivs[v].confidence)
                if k < 9:
                    iv.R[k] += h
                else:
                    iv.t[k - 9] += h
                vs = list(ivs)
                vs[v] = iv
                return O.tsdf_loss64(g, vs, decoded, None, vk[sub], tk[sub], vis=vis, per_voxel=True)
            # a sum over ~1e5 voxels always holds a few kinks (o = 1, w = 0.2, the weight cap)
            # within the step: the quotients with h and h/2 are asked to agree to 1e-4
            f = fd(loss, 1e-7, agree=1e-4)
            if np.isfinite(f):
                pose_err.append(abs(gv[k] - f) / vmax)
    depth_err, pose_err = np.array(depth_err), np.array(pose_err)
    say(f"central differences: {len(depth_err)} depth samples, worst {depth_err.max():.2f} x (1e-3 + 1e-4|g|); "
        f"{len(pose_err)} pose entries, worst {pose_err.max():.2e} of the view's largest entry")
    assert len(depth_err) + len(pose_err) >= 240 and len(pose_err) >= 24
    assert (depth_err <= 1.0).all()
    assert (pose_err <= 1e-3).all()
