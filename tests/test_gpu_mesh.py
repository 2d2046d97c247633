"""GPU surface nets and mesh post-processing (lr3d_tsdf_extract_surface,
lr3d_mesh_remove_small_components) against the oracle's restatement of the reference
(oracle/mesh_oracle.c): positions, normals and indices equal array for array, on the reference's
own sphere volume (fusion.spec.ts), on that volume plus a floating blob, and on the volume the
sphere workload fuses from its filtered depth maps."""
import numpy as np
import pytest

import oracle as O
from lr3d import scenes
from test_oracle_mesh import sphere_volume

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def same(gpu, ref, tag):
    for k in ("positions", "normals", "indices"):
        a, b = np.asarray(gpu[k]), np.asarray(ref[k])
        assert a.shape == b.shape, (tag, k, a.shape, b.shape)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (tag, k, int((a != b).sum()))


def _volume(ctx, vol):
    import lr3d
    grid = lr3d.BrickGrid(vol["origin"], vol["voxel"], (1, 1, 1))
    grid.dims = tuple(vol["dims"])
    coords = np.zeros(((vol["lookup"] >= 0).sum(), 4), np.int32)
    dx, dy = vol["dims"][0], vol["dims"][1]
    for i, b in enumerate(vol["lookup"]):
        if b >= 0:
            coords[b] = (i % dx, (i // dx) % dy, i // (dx * dy), 0)
    h = lr3d.HipTsdfVolume(ctx, grid, (coords.reshape(-1), vol["lookup"], len(coords)), 4.0)
    h.write(vol["tsdf"], vol["weight"])
    return h


def test_sphere_volume_mesh_equals_oracle(ctx):
    say("case: surface nets on the reference's sphere volume (r 5 cm, 2 mm voxels)")
    vol = sphere_volume()
    h = _volume(ctx, vol)
    m = h.extract_surface(0.5)
    got = m.read()
    ref = O.extract_surface(vol["origin"], vol["voxel"], vol["dims"], vol["lookup"], vol["tsdf"], vol["weight"], 0.5)
    say(f"mesh: {len(got['positions'])} vertices, {len(got['indices']) // 3} triangles")
    same(got, ref, "sphere")
    m.destroy()
    h.destroy()


def test_small_components_equal_oracle(ctx):
    say("case: sphere volume plus a floating blob, removeSmallComponents(0.01, 200)")
    vol = sphere_volume()
    # carve a small separate blob (2.5 voxel radius) into a corner of the allocated bricks
    coords_of = {int(b): i for i, b in enumerate(vol["lookup"]) if b >= 0}
    dx, dy = vol["dims"][0], vol["dims"][1]
    b0 = max(coords_of)  # last pool brick
    i = coords_of[b0]
    c = np.array([i % dx, (i // dx) % dy, i // (dx * dy)])
    local = np.arange(512)
    v = np.stack([local & 7, (local >> 3) & 7, local >> 6], axis=1)
    d = np.linalg.norm(v - 3.5, axis=1) - 2.5
    tsdf = vol["tsdf"].copy()
    weight = vol["weight"].copy()
    tsdf[b0 * 512:(b0 + 1) * 512] = np.clip(d / 4, -1, 1)
    # observed inside the brick only: no cell or quad crosses its boundary, so the blob is its own
    # component
    interior = np.all((v >= 1) & (v <= 6), axis=1)
    weight[b0 * 512:(b0 + 1) * 512] = np.where(interior, 2.0, 0.0)
    vol = dict(vol, tsdf=tsdf, weight=weight)
    h = _volume(ctx, vol)
    m = h.extract_surface(0.5)
    ref = O.extract_surface(vol["origin"], vol["voxel"], vol["dims"], vol["lookup"], tsdf, weight, 0.5)
    same(m.read(), ref, "with blob")
    cleaned = m.remove_small_components(0.01, 200)
    got = cleaned.read()
    ref_clean = O.remove_small_components(ref, 0.01, 200)
    say(f"components: {len(ref['indices']) // 3} -> {len(ref_clean['indices']) // 3} triangles")
    assert len(ref_clean["indices"]) < len(ref["indices"])
    same(got, ref_clean, "cleaned")
    for x in (cleaned, m, h):
        x.destroy()


def test_fused_sphere_workload_mesh_equals_oracle(ctx):
    import lr3d
    say("case: sphere 8x320x240 solved, filtered, fused (128^3 x 8 mm), meshed at min weight 0.6")
    w = scenes.sphere_small().render()
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    st.solve_batched([(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=w.ranges[r][0], max_depth=w.ranges[r][1],
                                            iterations=w.iterations, topk=w.topk)) for r in range(8)])
    for r in range(8):
        st.filter(r, w.sources[r], w.mesh_needed(r))
    spec = w.tsdf
    grid, layout = lr3d.dense_grid(spec.origin, spec.voxel, spec.voxels)
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
    views = [lr3d.IntegrationView(v.intrinsics, v.pose, w.ranges[i][1], v.width, v.height) for i, v in enumerate(w.views)]
    vol.integrate_batched(st, list(range(8)), views, map=lr3d.MAP_FILTERED)
    tsdf, weight = vol.read()
    m = vol.extract_surface(0.6)
    clean = m.remove_small_components(0.01, 200)
    got, got_clean = m.read(), clean.read()
    ref = O.extract_surface(grid.origin, grid.voxel, grid.dims, layout[1], tsdf, weight, 0.6)
    ref_clean = O.remove_small_components(ref, 0.01, 200)
    say(f"fused mesh: {len(ref['indices']) // 3} triangles, {len(ref_clean['indices']) // 3} after cleaning")
    same(got, ref, "fused")
    same(got_clean, ref_clean, "fused cleaned")
    # the surface lies near the sphere / ground
    r = np.linalg.norm(got_clean["positions"].astype(np.float64), axis=1)
    on_sphere = np.abs(r - 0.5) < 3 * spec.voxel
    on_ground = np.abs(got_clean["positions"][:, 1] + 0.5) < 3 * spec.voxel
    assert (on_sphere | on_ground).mean() > 0.95
    for x in (clean, m, vol, st):
        x.destroy()
