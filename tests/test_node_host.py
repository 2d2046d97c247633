"""The Node / TypeScript host (luvas.io-3d_amd/node/lr3d.js over the N-API addon lr3d.node):
it loads and exports the binding, fails loudly without a device, and on the GPU runs the
reference's stereo -> filter -> fusion sequence with results identical to the Python host."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "luvas.io-3d_amd", "lib", "lr3d.node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")

EXPORTS = ["ctxCreate", "ctxDestroy", "stereoCreate", "stereoDestroy", "stereoSetImage", "stereoSolve",
           "stereoFilter", "stereoReadDepth", "stereoReadFiltered", "stereoReadHyp", "stereoSetCoarse",
           "tsdfCreate", "tsdfDestroy", "tsdfIntegrate", "tsdfRead", "lastError", "nccBackward",
           "nccRefineStep", "stereoSetPose", "tsdfBackward", "tsdfExtractSurface", "stereoFuseDepthMaps",
           "stereoSolveNoalloc", "stereoAutotune", "stereoGetTile", "stereoSetTile", "stereoWriteDepth",
           "stereoWriteHyp", "tsdfAutotune", "tsdfIntegrateHost", "tsdfGetBricksPerBlock",
           "tsdfSetBricksPerBlock", "tsdfGetXcdChunk", "tsdfSetXcdChunk", "tsdfSetCounting", "tsdfSetDepthCulling", "tsdfUpdateCount",
           "tsdfPassCounts", "tsdfWrite", "ctxSynchronize", "ctxSetProfiling", "ctxKernelStats"]
# C-ABI entry points the addon reaches through another one
THROUGH = {"lr3d_tsdf_integrate": "HipTsdfVolume.integrate passes its one view to lr3d_tsdf_integrate_batched"}


def node(code, timeout=60):
    return subprocess.run([NODE, "-e", code], capture_output=True, text=True, timeout=timeout, cwd=ROOT)


def test_addon_exports_the_binding():
    r = node("const m=require('./luvas.io-3d_amd/node/lr3d.js');"
             "console.log(JSON.stringify(Object.keys(m.native)));"
             "console.log(Object.keys(m.STEREO_METHODS).join(','))")
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().split("\n")
    assert sorted(json.loads(lines[0])) == sorted(EXPORTS)
    assert lines[1] == "hip-single-scale,hip-coarse-to-fine"


def test_addon_binds_every_entry_point_of_the_header():
    import re
    header = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "lr3d.h")).read(), flags=re.S)
    declared = set(re.findall(r"\b(lr3d_[a-z0-9_]+)\s*\(", header))
    called = set(re.findall(r"\b(lr3d_[a-z0-9_]+)\(", open(os.path.join(ROOT, "luvas.io-3d_amd", "node", "lr3d_napi.c")).read()))
    assert sorted(declared - called - set(THROUGH)) == []


def test_brickgrid_matches_python_host():
    from lr3d import BrickGrid
    r = node("const m=require('./luvas.io-3d_amd/node/lr3d.js');"
             "const g=new m.BrickGrid([-0.1,0,0.2],0.01,[0.5,0.3,0.41]); g.mark(0.1,0.1,0.4,0.05); g.mark(0.3,0.2,0.5,0.02);"
             "const l=g.build(); console.log(JSON.stringify({dims:g.dims,coords:Array.from(l.coords),count:l.count}))")
    assert r.returncode == 0, r.stderr
    js = json.loads(r.stdout)
    g = BrickGrid((-0.1, 0, 0.2), 0.01, (0.5, 0.3, 0.41))
    g.mark(0.1, 0.1, 0.4, 0.05)
    g.mark(0.3, 0.2, 0.5, 0.02)
    coords, lookup, count = g.build()
    assert list(g.dims) == js["dims"] and count == js["count"] and list(coords) == js["coords"]


def test_context_without_device_rejects():
    from lr3d import Context, Lr3dError
    try:
        Context(0).destroy()
        pytest.skip("a HIP device is present")
    except Lr3dError:
        pass
    r = node("const m=require('./luvas.io-3d_amd/node/lr3d.js'); try { new m.Context(0) } catch (e) { console.log(e.message) }")
    assert "no HIP device" in r.stdout


@pytest.mark.gpu
def test_node_pipeline_matches_python_host(tmp_path, ctx):
    import lr3d
    from lr3d import scenes
    w = scenes.sphere_small().render()
    views = [{"intrinsics": {"fx": v.intrinsics.fx, "fy": v.intrinsics.fy, "cx": v.intrinsics.cx, "cy": v.intrinsics.cy},
              "pose": {"R": list(np.asarray(v.pose.R).reshape(9)), "t": list(np.asarray(v.pose.t))}}
             for v in w.views]
    job = {"method": "hip-single-scale", "width": 320, "height": 240, "views": views, "sources": w.sources,
           "ranges": w.ranges, "iterations": 3, "needed": 2, "origin": [-0.512] * 3, "voxel": 0.008,
           "extent": [1.024] * 3, "centre": [0.0, 0.0, 0.0], "radius": 0.6, "truncation": 4}
    g0 = lr3d.BrickGrid((-0.512,) * 3, 0.008, (1.024,) * 3)
    g0.mark(0.0, 0.0, 0.0, 0.6)
    rng = np.random.default_rng(3)
    voxels = rng.integers(0, g0.build()[2] * 512, 20000)
    job["voxels"] = [int(v) for v in voxels]
    job["targets"] = [float(v) for v in rng.uniform(-1, 1, voxels.size).astype(np.float32)]
    for i, im in enumerate(w.images):
        im.tofile(tmp_path / f"gray{i}.u8")
    # 80 x 60 colour thumbnails: the grey image subsampled, tinted per channel
    thumbs = [np.stack([im[::4, ::4], im[1::4, 2::4], 255 - im[2::4, 1::4]], axis=-1).astype(np.uint8)
              for im in w.images]
    for i, t in enumerate(thumbs):
        t.tofile(tmp_path / f"thumb{i}.rgb")
    job.update({"thumb": [80, 60], "cell": 0.005, "cellBudget": 20000})
    (tmp_path / "job.json").write_text(json.dumps(job))
    print("case: node host pipeline, 8 views 320x240", flush=True)
    r = subprocess.run([NODE, os.path.join(ROOT, "tests", "node", "pipeline.js"), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    # the same sequence through the Python host
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    for i in range(8):
        st.solve(i, lr3d.SolveOptions(sources=w.sources[i], min_depth=w.ranges[i][0], max_depth=w.ranges[i][1],
                                      iterations=3))
    grid = lr3d.BrickGrid((-0.512,) * 3, 0.008, (1.024,) * 3)
    grid.mark(0.0, 0.0, 0.0, 0.6)
    layout = grid.build()
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, 4)
    for i in range(8):
        assert np.array_equal(st.depth_map(i), np.fromfile(tmp_path / f"depth{i}.u32", np.uint32))
        st.filter(i, w.sources[i], 2)
        assert np.array_equal(st.read_filtered(i), np.fromfile(tmp_path / f"filtered{i}.u32", np.uint32))
        v = w.views[i]
        vol.integrate(st, i, lr3d.IntegrationView(v.intrinsics, v.pose, w.ranges[i][1], 320, 240), map=lr3d.MAP_FILTERED)
    # adjoints: the Node forward/backward pairs against the Python host
    losses = json.loads((tmp_path / "losses.json").read_text())
    o = lr3d.SolveOptions(sources=w.sources[0], min_depth=w.ranges[0][0], max_depth=w.ranges[0][1], iterations=3)
    nb = st.ncc_backward(0, o)
    assert np.array_equal(nb["cost"].view(np.uint32), np.fromfile(tmp_path / "ncc_cost.f32", np.uint32))
    assert np.array_equal(nb["grad_depth"].view(np.uint32), np.fromfile(tmp_path / "ncc_gd.f32", np.uint32))
    assert np.array_equal(nb["grad_normal"].reshape(-1).view(np.uint32), np.fromfile(tmp_path / "ncc_gn.f32", np.uint32))
    assert np.array_equal(nb["grad_pose"], np.fromfile(tmp_path / "ncc_gp.f64", np.float64).reshape(8, 12)[:len(w.sources[0])])
    assert nb["loss"] == losses["ncc"]
    vox = np.array(job["voxels"], dtype=np.uint32)
    ivs = [lr3d.IntegrationView(v.intrinsics, v.pose, w.ranges[i][1], 320, 240) for i, v in enumerate(w.views)]
    tb = vol.backward(st, list(range(8)), ivs, vox, np.array(job["targets"], np.float32), map=lr3d.MAP_FILTERED)
    assert np.array_equal(tb["tsdf"].view(np.uint32), np.fromfile(tmp_path / "tsdf_bw.f32", np.uint32))
    gd = np.concatenate([x.reshape(-1) for x in tb["grad_depth"]])
    assert np.allclose(gd, np.fromfile(tmp_path / "tsdf_gd.f32", np.float32), rtol=1e-6, atol=1e-9)
    assert np.allclose(tb["grad_pose"], np.fromfile(tmp_path / "tsdf_gp.f64", np.float64).reshape(8, 12), rtol=1e-10, atol=1e-12)
    assert abs(tb["loss"] - losses["tsdf"]) <= 1e-12 * max(1.0, abs(tb["loss"]))
    # dense cloud: the Node call against the Python host's
    cloud, cell = lr3d.fuse_depth_maps(st, list(range(8)), thumbs, 0.005, 20000)
    node_cell = json.loads((tmp_path / "cloud.json").read_text())["cell"]
    assert cell == node_cell
    for k, dt in (("positions", np.float32), ("normals", np.float32), ("colors", np.uint8), ("support", np.uint16)):
        assert np.array_equal(np.asarray(cloud[k]).reshape(-1), np.fromfile(tmp_path / f"cloud_{k}.bin", dt)), k
    # mesh: the Node call against the Python host's extraction
    m = vol.extract_surface(0.6)
    c = m.remove_small_components(0.01, 200)
    mesh = c.read()
    for k, dt in (("positions", np.float32), ("normals", np.float32), ("indices", np.uint32)):
        assert np.array_equal(np.asarray(mesh[k]).reshape(-1).view(np.uint32),
                              np.fromfile(tmp_path / f"mesh_{k}.bin", dt).view(np.uint32)), k
    c.destroy()
    m.destroy()
    tsdf, weight = vol.read()
    assert np.array_equal(np.fromfile(tmp_path / "coords.i32", np.int32), layout[0])
    assert np.array_equal(tsdf.view(np.uint32), np.fromfile(tmp_path / "tsdf.f32", np.uint32))
    assert np.array_equal(weight.view(np.uint32), np.fromfile(tmp_path / "weight.f32", np.uint32))
    # the Node host's refinement driver: a non-increasing loss that drops
    hist = json.loads((tmp_path / "refine.json").read_text())
    assert len(hist) == 6 and np.isfinite(hist).all()
    assert all(b <= a for a, b in zip(hist, hist[1:])) and hist[-1] < hist[0]
    # the rest of the C ABI through the Node host
    ex = json.loads((tmp_path / "extras.json").read_text())
    assert ex["noallocSame"] and ex["writeRoundTrip"] and ex["hostSame"]
    assert ex["propagateLaunches"] > 0
    assert ex["tilesSet"] == [16, 4] and set(ex["tiles"]) == {"init", "propagate", "pack"}
    pc = ex["passCounts"]
    assert pc["updates"] == ex["updateCount"] > 0 and pc["pairs"] >= pc["updates"] and pc["voxelsWritten"] > 0
    assert ex["bricksPerBlock"] in (1, 2, 4)
    assert 1 <= ex["xcdChunk"] <= 4096
    pyr = json.loads((tmp_path / "refine_pyramid.json").read_text())
    assert len(pyr["coarse"]) == 2
    for h in pyr["coarse"] + [pyr["fine"]]:
        assert len(h) == 3 and np.isfinite(h).all() and all(b <= a for a, b in zip(h, h[1:]))
    vol.destroy()
    st.destroy()


def test_node_pose_math_matches_python():
    """The Node host's refinement math (lr3d.js: worldPoseGradients, perturbPose) against the
    Python host's (lr3d/refine.py) on seeded poses and gradients."""
    from lr3d import Intrinsics, StereoView
    from lr3d.refine import perturb, world_pose_gradients
    from lr3d.camera import Pose, look_at
    rng = np.random.default_rng(8)
    views = [StereoView(Intrinsics(300.0, 300.0, 160.0, 120.0),
                        look_at(tuple(rng.uniform(-2, 2, 3) + np.array([0, 0, 3.0]))), 320, 240) for _ in range(4)]
    grad = rng.normal(0, 1, (8, 12))
    w, tau = rng.normal(0, 0.01, 3), rng.normal(0, 0.01, 3)
    js_views = [{"intrinsics": {"fx": 300.0, "fy": 300.0, "cx": 160.0, "cy": 120.0},
                 "pose": {"R": list(np.asarray(v.pose.R).reshape(9)), "t": list(np.asarray(v.pose.t))}} for v in views]
    code = ("const l = require('./luvas.io-3d_amd/node/lr3d.js');"
            f"const views = {json.dumps(js_views)};"
            f"const g = Float64Array.from({json.dumps(list(grad.reshape(-1)))});"
            "const m = l.worldPoseGradients(views, 0, [1, 2, 3], g);"
            "const out = {}; for (const [k, v] of m) out[k] = v;"
            f"const p = l.perturbPose(views[1].pose, {json.dumps(list(w))}, {json.dumps(list(tau))});"
            "console.log(JSON.stringify({grads: out, pose: p}));")
    r = node(code)
    assert r.returncode == 0, r.stderr
    res = json.loads(r.stdout.strip().splitlines()[-1])
    ref = world_pose_gradients(views, 0, [1, 2, 3], grad[:3])
    for layer, (gw, gt) in ref.items():
        got = res["grads"][str(layer)]
        assert np.allclose(got["w"], gw, rtol=1e-12, atol=1e-12) and np.allclose(got["tau"], gt, rtol=1e-12, atol=1e-12)
    p = perturb(views[1].pose, w, tau)
    assert np.allclose(np.reshape(res["pose"]["R"], (3, 3)), p.R, rtol=0, atol=1e-14)
    assert np.allclose(res["pose"]["t"], p.t, rtol=0, atol=1e-14)


def test_addon_rejects_wrong_typed_arrays():
    """Typed arrays are checked for their element type and length before anything is copied: a
    short prior, a prior of another element type, voxels of the wrong type and an image of the
    wrong type throw before any device work (no handle is needed to reach the checks)."""
    code = r"""
const m = require('./luvas.io-3d_amd/node/lr3d.js');
const n = m.native;
const out = {};
function tryCall(name, f) {
    try { f(); out[name] = 'no error'; } catch (e) { out[name] = e.constructor.name + ': ' + e.message; }
}
const desc = (prior) => ({ ref: 0, pixels: 12, sources: [1], minDepth: 0.5, maxDepth: 3, prior, iterations: 1 });
tryCall('short', () => n.stereoSolve(null, desc(new Float32Array(11))));
tryCall('u8', () => n.stereoSolve(null, desc(new Uint8Array(12))));
tryCall('f64', () => n.stereoSolve(null, desc(new Float64Array(12))));
tryCall('batched', () => n.stereoSolve(null, [desc(null), desc(new Float32Array(13))]));
tryCall('nopixels', () => n.stereoSolve(null, { ref: 0, sources: [1], minDepth: 0.5, maxDepth: 3,
    prior: new Float32Array(12), iterations: 1 }));
const iview = (prior) => ({ width: 4, height: 3, fx: 1, fy: 1, cx: 1, cy: 1, R: [1, 0, 0, 0, 1, 0, 0, 0, 1],
    t: [0, 0, 0], maxDepth: 3, prior, confidence: 0 });
tryCall('iview_short', () => n.tsdfIntegrate(null, null, [0], 0, [iview(new Float32Array(5))]));
tryCall('iview_u8', () => n.tsdfIntegrate(null, null, [0], 0, [iview(new Uint8Array(12))]));
tryCall('voxels_f32', () => n.tsdfBackward(null, null, [0], 0, [iview(null)], new Float32Array(3), new Float32Array(3)));
tryCall('image_f32', () => n.stereoSetImage(null, 0, new Float32Array(12), 4, 3));
tryCall('image_short', () => n.stereoSetImage(null, 0, new Uint8Array(11), 4, 3));
tryCall('coords_u32', () => n.tsdfCreate(null, [0, 0, 0], 0.01, [1, 1, 1], new Uint32Array(4), 4));
tryCall('volume_depth', () => new m.HipTsdfVolume(null, { voxel: 0.01 }, { coords: new Int32Array(4) }, {}, 4, 3));
// a well-formed prior reaches the C ABI, which rejects the missing handle asynchronously
const p = n.stereoSolve(null, desc(new Float32Array(12)));
p.then(() => { out.valid = 'resolved'; }, (e) => { out.valid = 'rejected: ' + e.message; })
 .then(() => console.log(JSON.stringify(out)));
"""
    r = node(code)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    for k in ("short", "u8", "f64", "batched", "nopixels", "iview_short", "iview_u8"):
        assert out[k].startswith("RangeError") and "Float32Array of width*height" in out[k], (k, out[k])
    assert out["voxels_f32"].startswith("TypeError") and "Uint32Array" in out["voxels_f32"]
    assert out["image_f32"].startswith("TypeError") and out["image_short"].startswith("TypeError")
    assert out["coords_u32"].startswith("TypeError") and "Int32Array" in out["coords_u32"]
    assert out["volume_depth"].startswith("TypeError") and "stereo.filtered" in out["volume_depth"]
    assert out["valid"].startswith("rejected") and "NULL" in out["valid"]
