#!/usr/bin/env python3
"""Benchmark of the dense-reconstruction hot path on one MI355X.

One step = batched PatchMatch of every view (init + `iterations` red-black rounds + pack, all
views sharing each launch), the consistency filter of every view, one batched TSDF integration
pass of every filtered map, and the TSDF adjoint of that pass over 1M seeded voxels near the
surfaces (loss: squared residual against the analytic signed distance; gradients with respect to
every depth used and every pose, computed on the device; the pose gradients and the loss are read
back, the per-pixel depth gradients stay on the device).  The timed region launches eagerly with a
HIP event pair around every kernel (HIP cannot time events inside graphs); the same steps are then
replayed with the PatchMatch schedule from its captured HIP graph and reported under "graph".
Inputs (u8 frames, cameras, the voxel list and its targets) are resident in device memory before
the timed region.

Default workload: 200 views at 1280x720 of the seeded procedural room, 4-8 neighbours per view
(top-k by the reference rule), 8 iterations, TSDF 512^3 x 8 mm (1.07 GB, truncation 4 voxels).
Prints one JSON line.

  metric value   PatchMatch Mpx*hyp/s over the whole step (Mpx * (1 + 14*iterations) per step)
  tsdf           voxel updates per step and per second of the integration pass, and the adjoint
  roofline       the dominant kernel (pm_propagate) against the FP32 vector peak, timed with HIP
                 events on the library's stream over the timed region
  cpu_baseline   the C oracle (same arithmetic, OpenMP) on a bounded sample of the same workload;
                 under "downstream" the oracle's fusion, surface nets and small-component removal
                 (one thread) on the same inputs as the GPU's, with whether the results are equal
  plain          the per-view path (one solve / filter / integrate call per view, the same adjoint)
                 for the batched / plain ratio
  accuracy       bad pixels (> 1 % relative depth error vs the analytic depth) on the filtered maps
                 that are fused (the gated map) and on the raw PatchMatch maps
  adjoint        the NCC adjoint of every view in one launch (outside the step)
  downstream     dense-cloud fusion of the filtered maps, surface nets of the volume and small-
                 component removal (outside the step)
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "luvas.io-3d_amd"))

# MI355X published peaks
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
# the same ALUs one FP32 operation per lane per cycle (no packed FP32, which this build does not use)
FP32_UNPACKED_TFLOPS = 78.6
# Algorithmic FLOPs: 33 per NCC tap (homography 18, divide 3, bilinear 6, sums 6), 36 taps per
# source view, 20 per view for the NCC finish; 12 per window tap for the bilateral window.
FLOP_PER_TAP = 33
TAPS = 36
FLOP_VIEW_FINISH = 20
FLOP_WINDOW = 36 * 12


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="room-large")
    ap.add_argument("--adjoint-voxels", type=int, default=1_000_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint legs")
    ap.add_argument("--no-downstream", action="store_true", help="skip the fusion and meshing leg")
    ap.add_argument("--no-profile", action="store_true", help="time without kernel events")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--tiles", default=None,
                    help="pin launch shapes instead of tuning, e.g. init=16x8,propagate=8x8,pack=64x4,bricks=4,chunk=64")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import lr3d
    from lr3d import scenes

    t_setup = time.time()
    work = scenes.WORKLOADS[args.workload]().render()
    log(f"workload {work.name}: {len(work.views)} views, {work.megapixels:.2f} Mpx, rendered in "
        f"{time.time() - t_setup:.1f}s")
    ctx = lr3d.Context(0)
    st = lr3d.HipDenseStereo(ctx, work.views)
    for i, im in enumerate(work.images):
        st.set_image(i, im)
    spec = work.tsdf
    grid, layout = lr3d.dense_grid(spec.origin, spec.voxel, spec.voxels)
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
    n = len(work.views)
    solves = [(r, lr3d.SolveOptions(sources=work.sources[r], min_depth=work.ranges[r][0],
                                    max_depth=work.ranges[r][1], iterations=work.iterations,
                                    topk=work.topk)) for r in range(n)]
    iviews = [lr3d.IntegrationView(v.intrinsics, v.pose, work.ranges[i][1], v.width, v.height)
              for i, v in enumerate(work.views)]
    needed = [work.mesh_needed(r) for r in range(n)]
    layers = list(range(n))

    # The TSDF adjoint's voxels: seeded pool indices within two truncation distances of the
    # analytic surface, targets the clamped signed distance in truncation units.
    rng = np.random.default_rng(7)
    coords = layout[0].reshape(-1, 4)
    cand = np.unique(rng.integers(0, layout[2] * 512, 8 * args.adjoint_voxels).astype(np.uint32))
    bb, ll = cand // 512, cand % 512
    cell = coords[bb, :3] * 8 + np.stack([ll % 8, (ll // 8) % 8, ll // 64], axis=1)
    mu = spec.voxel * spec.truncation_voxels
    sd = work.scene.signed_distance(np.asarray(spec.origin) + (cell + 0.5) * spec.voxel)
    near = np.flatnonzero(np.abs(sd) < 2 * mu)
    pick = np.sort(rng.choice(near, min(len(near), args.adjoint_voxels), replace=False))
    vox, tgt = cand[pick], np.clip(sd[pick] / mu, -1, 1).astype(np.float32)
    log(f"TSDF adjoint: {len(vox)} voxels within {2 * mu * 1000:g} mm of the surfaces")
    last = {}

    def step(stereo=st, volume=vol, batched=True):
        if batched:
            stereo.solve_batched(solves)
        else:
            for r, o in solves:
                stereo.solve(r, o)
        for r in range(n):
            stereo.filter(r, work.sources[r], needed[r])
        if batched:
            volume.integrate_batched(stereo, layers, iviews, map=lr3d.MAP_FILTERED)
        else:
            for r in range(n):
                volume.integrate(stereo, r, iviews[r], map=lr3d.MAP_FILTERED)
        last["adjoint"] = volume.backward(stereo, layers, iviews, vox, tgt, map=lr3d.MAP_FILTERED,
                                          grad_depth=False)

    # launch shapes: timed on this batch before the warm-up (leaves the set and volume unchanged)
    tiles = None
    if args.tiles:
        for item in args.tiles.split(","):
            k, v = item.split("=")
            if k == "bricks":
                vol.bricks_per_block = int(v)
            elif k == "chunk":
                vol.xcd_chunk = int(v)
            else:
                tx, ty = (int(x) for x in v.split("x"))
                st.set_tile(k, tx, ty)
        tiles = {k: list(v) for k, v in st.tiles().items()}
        tiles["tsdf_bricks_per_block"] = vol.bricks_per_block
        tiles["tsdf_xcd_chunk"] = vol.xcd_chunk
        log(f"pinned launch shapes: {tiles}")
    elif not args.no_autotune:
        t_tune = time.perf_counter()
        tiles = st.autotune(solves, repeats=3)
        st.solve_batched(solves)
        for r in range(n):
            st.filter(r, work.sources[r], needed[r])
        bricks = vol.autotune(st, list(range(n)), iviews, map=lr3d.MAP_FILTERED, repeats=3)
        tiles = {k: list(v) for k, v in tiles.items()}
        tiles["tsdf_bricks_per_block"] = bricks
        tiles["tsdf_xcd_chunk"] = vol.xcd_chunk
        log(f"autotune ({time.perf_counter() - t_tune:.1f}s): {tiles}")

    for w in range(args.warmup):
        log(f"warmup step {w}")
        step()
    ctx.set_profiling(not args.no_profile)
    ctx.synchronize()
    log(f"timed region: {args.steps} steps")
    t0 = time.perf_counter()
    for s in range(args.steps):
        step()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    stats = {k: ctx.kernel_stats(k) for k in ("pm_init", "pm_propagate", "pm_pack", "pm_consistency",
                                            "tsdf_tiles", "tsdf_visibility", "tsdf_integrate",
                                            "tsdf_backward")}
    ctx.set_profiling(False)
    ms_per_step = elapsed * 1e3 / args.steps
    # The same steps replayed from the captured graph (HIP cannot time events inside graphs, so
    # the event-timed region above launches eagerly).
    graph_before = ctx.kernel_stats("graph_launch")[0]
    graph_steps = min(args.steps, 3)
    ctx.synchronize()
    tg = time.perf_counter()
    for s in range(graph_steps):
        step()
    ctx.synchronize()
    graph_ms = (time.perf_counter() - tg) * 1e3 / graph_steps
    graph_replays = ctx.kernel_stats("graph_launch")[0] - graph_before
    hyp_per_px = work.hypotheses_per_pixel()
    mpx_hyp = work.megapixels * hyp_per_px
    value = mpx_hyp * args.steps / elapsed
    log(f"{ms_per_step:.1f} ms/step, {value:.1f} Mpx*hyp/s")

    # voxel updates of one pass (counted in an untimed pass of the same kernels)
    vol.set_counting(True)
    vol.integrate_batched(st, list(range(n)), iviews, map=lr3d.MAP_FILTERED)
    updates = vol.update_count()
    counts = vol.pass_counts()
    vol.set_counting(False)

    # accuracy of the last step against the analytic depth (filtered maps, as fused)
    from lr3d.packing import unpack_depth
    bad = valid = raw_bad = raw_valid = 0
    for r in range(n):
        gt = work.truth[r].reshape(-1)
        for packed, is_raw in ((st.read_filtered(r), False), (st.depth_map(r), True)):
            m = (packed != 0) & (gt > 0)
            err = np.abs(unpack_depth(packed[m]) - gt[m]) / gt[m]
            if is_raw:
                raw_valid += int(m.sum())
                raw_bad += int((err > 0.01).sum())
            else:
                valid += int(m.sum())
                bad += int((err > 0.01).sum())

    # roofline of the dominant kernel: pm_propagate, FP32 vector bound
    prop_launches, prop_ms = stats["pm_propagate"]
    roofline = None
    if prop_launches:
        colour_px = sum(v.width * v.height for v in work.views) / 2.0
        per_hyp = sum(len(work.sources[r][:8]) for r in range(n)) / n * (TAPS * FLOP_PER_TAP + FLOP_VIEW_FINISH)
        flops = colour_px * (14 * per_hyp + FLOP_WINDOW)
        avg_s = prop_ms / prop_launches / 1e3
        achieved = flops / avg_s / 1e12
        roofline = {"kernel": "pm_propagate", "bound": "valu", "achieved": round(achieved, 3),
                    "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                    "traffic": None, "avg_launch_ms": round(prop_ms / prop_launches, 3),
                    "flop_per_launch": flops, "peak_unpacked": FP32_UNPACKED_TFLOPS,
                    "frac_unpacked": round(achieved / FP32_UNPACKED_TFLOPS, 4)}
    # bytes per launch between L2 and memory (FETCH_SIZE + WRITE_SIZE, copy-calibrated) measured by
    # rocprofv3 counters for this workload (profiles/traffic.json); an upper bound of HBM traffic
    measured = {}
    try:
        measured = json.load(open(os.path.join(ROOT, "profiles", "traffic.json"))).get(work.name, {})
    except (OSError, ValueError):
        pass
    if roofline is not None:
        roofline["traffic"] = measured.get("pm_propagate")
    tsdf_launches, tsdf_ms = stats["tsdf_integrate"]
    tb_launches, tb_ms = stats["tsdf_backward"]
    tsdf = {"voxel_updates_per_step": updates, "voxels": layout[2] * 512,
            "adjoint": {"voxels": int(len(vox)), "views": n, "loss": last["adjoint"]["loss"],
                        "avg_launch_ms": round(tb_ms / max(tb_launches, 1), 3),
                        "voxel_views_per_s": round(len(vox) * n / (tb_ms / max(tb_launches, 1)) * 1e3, 1),
                        "finite": bool(np.isfinite(last["adjoint"]["grad_pose"]).all())}}
    if tsdf_launches:
        t_avg = tsdf_ms / tsdf_launches / 1e3
        map_bytes = sum(v.width * v.height * 4 for v in work.views)
        bytes_alg = layout[2] * 512 * 16 + map_bytes
        # bytes of the bricks the pass actually touches: every brick with a kept view is read
        # (8 B per voxel), changed voxels are written (8 B), plus the maps
        bytes_touched = counts["bricks_read"] * 512 * 8 + counts["voxels_written"] * 8 + map_bytes
        traffic = measured.get("tsdf_integrate")
        tsdf.update({"avg_launch_ms": round(t_avg * 1e3, 3),
                     "voxel_upd_per_s": updates / t_avg,
                     "pass_counts": counts,
                     "pairs_per_update": round(counts["pairs"] / max(updates, 1), 3),
                     # readings against the HBM peak; the pass itself is limited by VALU issue (one
                     # projection and update per (voxel, view) pair; the counter profiles)
                     "roofline": {"kernel": "tsdf_integrate", "bound": "hbm", "limiter": "valu",
                                  "achieved": round(bytes_alg / t_avg / 1e9, 1), "peak": HBM_PEAK_GBS,
                                  "unit": "GB/s", "frac": round(bytes_alg / t_avg / 1e9 / HBM_PEAK_GBS, 4),
                                  "bytes_per_launch": bytes_alg, "traffic": traffic,
                                  "frac_touched": round(bytes_touched / t_avg / 1e9 / HBM_PEAK_GBS, 4),
                                  "bytes_touched": bytes_touched,
                                  "frac_traffic": round(traffic / t_avg / 1e9 / HBM_PEAK_GBS, 4) if traffic else None}})

    # The NCC adjoint on the same batch (outside the step): every view's aggregated cost, dL/dplane
    # and relative-pose gradients in one launch (lr3d_ncc_refine_step with zero steps), timed with
    # HIP events.
    adjoint = None
    if not args.no_adjoint:
        ctx.set_profiling(True)
        for _ in range(3):
            st.refine_step(solves, 0.0, 0.0)
        ncc_launches, ncc_ms = ctx.kernel_stats("ncc_backward")
        ctx.set_profiling(False)
        adjoint = {"ncc": {"views": n, "mpx": round(work.megapixels, 3),
                           "avg_launch_ms": round(ncc_ms / max(ncc_launches, 1), 3),
                           "mpx_per_s": round(work.megapixels / (ncc_ms / max(ncc_launches, 1)) * 1e3, 1)}}
        log(f"adjoint: NCC {adjoint['ncc']['avg_launch_ms']} ms for {n} views")

    plain = None
    if not args.no_plain:
        log("plain path: one call per view")
        st2 = lr3d.HipDenseStereo(ctx, work.views)
        for i, im in enumerate(work.images):
            st2.set_image(i, im)
        vol2 = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
        for k, (tx, ty) in st.tiles().items():  # same launch shapes as the batched path
            st2.set_tile(k, tx, ty)
        vol2.bricks_per_block = vol.bricks_per_block
        vol2.xcd_chunk = vol.xcd_chunk
        batched_adjoint = last["adjoint"]
        ctx.synchronize()
        t1 = time.perf_counter()
        step(st2, vol2, batched=False)
        ctx.synchronize()
        plain_s = time.perf_counter() - t1
        same = all(np.array_equal(st2.depth_map(r), st.depth_map(r)) for r in range(n))
        same = same and np.array_equal(batched_adjoint["tsdf"], last["adjoint"]["tsdf"])
        plain = {"ms_per_step": round(plain_s * 1e3, 2), "mpx_hyp_per_s": round(mpx_hyp / plain_s, 2),
                 "batched_speedup": round(plain_s * 1e3 / ms_per_step, 3), "same_results": bool(same),
                 "what": "one solve, filter and integrate call per view, then the same adjoint call"}
        st2.destroy()
        vol2.destroy()

    # TSDF integration in the reference's schedule, one map per pass (TV:103-154 per keyframe): the
    # same filtered maps into a fresh volume, view after view.  A counting run gives what each pass
    # touches (every brick with a kept view is read, 8 B per voxel; changed voxels are written, 8 B;
    # the view's map, 4 B per pixel); a second run times each integrate kernel with HIP events.
    per_view = None
    if not args.no_plain:
        counts_pv = None
        pv_ms = pv_launches = pv_cull_ms = 0
        for timed in (False, True):
            vol3 = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
            vol3.bricks_per_block = vol.bricks_per_block
            vol3.xcd_chunk = vol.xcd_chunk
            vol3.set_counting(not timed)
            ctx.synchronize()
            ctx.set_profiling(timed)
            for r in range(n):
                vol3.integrate(st, r, iviews[r], map=lr3d.MAP_FILTERED)
            ctx.synchronize()
            if timed:
                pv_launches, pv_ms = ctx.kernel_stats("tsdf_integrate")
                pv_cull_ms = ctx.kernel_stats("tsdf_tiles")[1] + ctx.kernel_stats("tsdf_visibility")[1]
            else:
                counts_pv = vol3.pass_counts()
            ctx.set_profiling(False)
            vol3.destroy()
        if pv_launches:
            map_bytes = sum(v.width * v.height * 4 for v in work.views)
            touched = counts_pv["bricks_read"] * 512 * 8 + counts_pv["voxels_written"] * 8 + map_bytes
            pv_s = pv_ms / 1e3
            per_view = {"passes": pv_launches, "avg_launch_ms": round(pv_ms / pv_launches, 4),
                        "total_ms": round(pv_ms, 3),
                        # depth tiles, culling and the brick list before each pass's integration
                        "culling_avg_ms": round(pv_cull_ms / pv_launches, 4), "pass_counts": counts_pv,
                        "bytes_touched": touched, "achieved_gbs": round(touched / pv_s / 1e9, 1),
                        "frac_touched": round(touched / pv_s / 1e9 / HBM_PEAK_GBS, 4),
                        "voxel_upd_per_s": counts_pv["updates"] / pv_s,
                        "batched_pass_speedup": round(pv_ms / (tsdf["avg_launch_ms"]), 3)
                        if tsdf.get("avg_launch_ms") else None}
            log(f"per-view TSDF passes: {pv_launches} x {per_view['avg_launch_ms']} ms, "
                f"{per_view['achieved_gbs']} GB/s on the bytes they touch")
        tsdf["per_view"] = per_view

    # The stages after the step (outside it), on its outputs, with the reference's settings:
    # dense-cloud fusion of every filtered map (fuseDepthMaps with 4 mm cells, the high-quality
    # profile's 600000-cell budget, quarter-size grey thumbnails), surface nets of the step's volume
    # (min weight 0.6) and the removal of components under max(200 faces, 1 %) (tsdf-meshing.ts:87-91).
    # The calls are synchronous: wall time per call (fusion: thumbnails up, cloud back), the median
    # of 3; the surface-net kernels also with HIP events, against the bytes they must move at least
    # (every pool voxel's weight read, 4 B, to find the observed bricks; 24 B per vertex and 4 B per
    # index written).
    downstream = None
    if not args.no_downstream:
        thumbs = []
        for im in work.images:
            g = np.asarray(im)[::4, ::4]
            thumbs.append(np.ascontiguousarray(np.stack([g, g, g], axis=2)))
        fus_s = []
        for _ in range(3):
            ctx.synchronize()
            t3 = time.perf_counter()
            cloud, final_cell = lr3d.fuse_depth_maps(st, list(range(n)), thumbs, 0.004, 600_000)
            fus_s.append(time.perf_counter() - t3)
        fus = sorted(fus_s)[1]
        samples = sum(int(np.count_nonzero(st.read_filtered(r))) for r in range(n))
        mesh_s, comp_s = [], []
        ctx.set_profiling(True)
        for _ in range(3):
            ctx.synchronize()
            t3 = time.perf_counter()
            mesh = vol.extract_surface(0.6)
            ctx.synchronize()
            t4 = time.perf_counter()
            kept = mesh.remove_small_components(0.01, 200)
            ctx.synchronize()
            mesh_s.append(t4 - t3)
            comp_s.append(time.perf_counter() - t4)
            (verts, indices), (kverts, kindices) = mesh.size, kept.size
            gpu_mesh, gpu_kept = mesh.read(), kept.read()
            mesh.destroy()
            kept.destroy()
        mesh_kernel_ms = sum(ctx.kernel_stats(k)[1]
                             for k in ("mesh_bricks", "mesh_active", "mesh_vertices", "mesh_quads")) / 3
        ctx.set_profiling(False)
        mesh_bytes = 4 * layout[2] * 512 + 24 * verts + 4 * indices
        downstream = {
            "fusion": {"ms": round(fus * 1e3, 2), "samples": samples, "samples_per_s": round(samples / fus, 1),
                       "points": int(len(cloud["support"])), "cell_m": 0.004, "cell_budget": 600_000,
                       "final_cell_m": round(final_cell, 6)},
            "surface_nets": {"ms": round(sorted(mesh_s)[1] * 1e3, 3), "kernel_ms": round(mesh_kernel_ms, 3),
                             "pool_voxels": layout[2] * 512, "vertices": verts, "triangles": indices // 3,
                             "bytes": mesh_bytes, "achieved_gbs": round(mesh_bytes / mesh_kernel_ms / 1e6, 1),
                             "frac": round(mesh_bytes / mesh_kernel_ms / 1e6 / HBM_PEAK_GBS, 4)},
            "small_components": {"ms": round(sorted(comp_s)[1] * 1e3, 3), "vertices_kept": kverts,
                                 "triangles_kept": kindices // 3}}
        log(f"downstream: fusion {downstream['fusion']['ms']} ms ({len(cloud['support'])} points), surface nets "
            f"{downstream['surface_nets']['ms']} ms, small components {downstream['small_components']['ms']} ms")

    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        threads = min(16, os.cpu_count() or 1)
        O.set_threads(threads)
        imgs = O.Images([work.images[i] for i in range(n)])
        # a workload small enough runs whole (every view, every iteration), repeated until ~10 s of
        # CPU work; others run a bounded sample: init + 1 iteration per view until 10 s are spent
        whole = mpx_hyp <= 300.0
        cpu_iters = work.iterations if whole else 1
        budget_s = 10.0
        log(f"cpu baseline: oracle, init + {cpu_iters} iteration(s) per view for {budget_s:g}s, {threads} threads")
        t2 = time.perf_counter()
        done = []
        passes = 0
        while True:
            for ref in range(n):
                lo, hi = work.ranges[ref]
                O.solve(work.views, imgs, ref, work.sources[ref], lo, hi, iterations=cpu_iters, topk=work.topk)
                if passes == 0:
                    done.append(ref)
                if not whole and time.perf_counter() - t2 >= budget_s:
                    break
            passes += 1
            if not whole or time.perf_counter() - t2 >= budget_s:
                break
        cpu_s = time.perf_counter() - t2
        sample = passes * sum(work.views[r].width * work.views[r].height for r in done) / 1e6 * (1 + 14 * cpu_iters)
        what = (f"the whole workload x {passes}" if whole
                else f"views {done[0]}-{done[-1]} of {work.name}")
        cpu = {"value": round(sample / cpu_s, 3), "unit": "Mpx*hyp/s", "cores": threads, "kind": "port",
               "sample": f"{what} (init + {cpu_iters} red-black iteration(s) each, same sources and top-k) = "
                         f"{sample:.2f} Mpx*hyp in {cpu_s:.1f}s on {cpu_model()}"}
        if downstream is not None:
            # the stages after the step on the CPU (the oracle, one thread) beside the GPU, same inputs:
            # surface nets and small-component removal of the step's volume, and fusion of the first
            # 10 filtered maps without coarsening (the GPU fuses the same maps again for the comparison)
            tsdf_h, weight_h = vol.read()
            t3 = time.perf_counter()
            ref_mesh = O.extract_surface(grid.origin, grid.voxel, grid.dims, layout[1], tsdf_h, weight_h, 0.6)
            mesh_cpu = time.perf_counter() - t3
            del tsdf_h, weight_h
            t3 = time.perf_counter()
            ref_kept = O.remove_small_components(ref_mesh, 0.01, 200)
            comp_cpu = time.perf_counter() - t3
            same_mesh = all(np.array_equal(np.asarray(gpu_mesh[k]).reshape(-1), np.asarray(ref_mesh[k]).reshape(-1))
                            for k in ("positions", "normals", "indices"))
            same_kept = all(np.array_equal(np.asarray(gpu_kept[k]).reshape(-1), np.asarray(ref_kept[k]).reshape(-1))
                            for k in ("positions", "normals", "indices"))
            k10 = min(10, n)
            maps10 = [st.read_filtered(r) for r in range(k10)]
            t3 = time.perf_counter()
            ref_cloud, _ = O.fuse_depth_maps(work.views[:k10], maps10, thumbs[:k10], 0.004, 1 << 40)
            fus_cpu = time.perf_counter() - t3
            gpu_cloud, _ = lr3d.fuse_depth_maps(st, list(range(k10)), thumbs[:k10], 0.004, 1 << 40)

            def rows(c):
                r = np.concatenate([c["positions"].view(np.uint32), c["normals"].view(np.uint32),
                                    c["colors"].astype(np.uint32), c["support"][:, None].astype(np.uint32)], axis=1)
                return r[np.lexsort(r.T[::-1])]
            same_cloud = bool(np.array_equal(rows(gpu_cloud), rows(ref_cloud)))
            samples10 = int(ref_cloud["support"].astype(np.int64).sum())
            cpu["downstream"] = {
                "cores": 1,
                "surface_nets": {"ms": round(mesh_cpu * 1e3, 1), "voxels_per_s": round(layout[2] * 512 / mesh_cpu, 1),
                                 "same_as_gpu": bool(same_mesh)},
                "small_components": {"ms": round(comp_cpu * 1e3, 1), "same_as_gpu": bool(same_kept)},
                "fusion": {"views": k10, "samples": samples10, "ms": round(fus_cpu * 1e3, 1),
                           "samples_per_s": round(samples10 / fus_cpu, 1), "same_as_gpu": same_cloud}}
            log(f"cpu downstream: surface nets {mesh_cpu:.1f}s (same as GPU: {same_mesh}), small components "
                f"{comp_cpu:.1f}s ({same_kept}), fusion of {k10} maps {fus_cpu:.1f}s ({same_cloud})")

    out = {
        "metric": "PatchMatch Mpx*hyp/s (batched PatchMatch + filter + TSDF integration + TSDF adjoint step)",
        "value": round(value, 3),
        "unit": "Mpx*hyp/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene, analytic ground truth)",
        "config": {"workload": work.name, "views": n,
                   "resolution": sorted({f"{v.width}x{v.height}" for v in work.views}),
                   "sources": sorted({len(x[:8]) for x in work.sources}),
                   "topk": work.topk or "max(1, min(3, sources - 1))", "iterations": work.iterations,
                   "launch_shapes": tiles,
                   "tsdf": f"{spec.voxels}^3 x {spec.voxel * 1000:g} mm, truncation {spec.truncation_voxels:g} voxels",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "tsdf": tsdf,
        "plain": plain,
        "adjoint": adjoint,
        "downstream": downstream,
        "accuracy": {"gate": "filtered maps (the maps that are fused) <= 1 % bad pixels",
                     "gate_pass": bad / max(valid, 1) <= 0.01,
                     "filtered_bad_frac": round(bad / max(valid, 1), 5), "filtered_valid_px": valid,
                     "raw_bad_frac": round(raw_bad / max(raw_valid, 1), 5), "raw_valid_px": raw_valid,
                     "bad_threshold": "relative depth error > 1 % vs analytic depth"},
        "kernel_ms": {k: {"launches": v[0], "total_ms": round(v[1], 3)} for k, v in stats.items()},
        "timed_region": "eager launches with a HIP event pair around every kernel",
        "graph": {"ms_per_step": round(graph_ms, 3), "steps": graph_steps, "replays": graph_replays,
                  "mpx_hyp_per_s": round(mpx_hyp / graph_ms * 1e3, 3)},
    }
    print(json.dumps(out), flush=True)
    vol.destroy()
    st.destroy()
    ctx.destroy()


if __name__ == "__main__":
    main()
