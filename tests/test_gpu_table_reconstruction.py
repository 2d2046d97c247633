"""End to end on the reference's verification scene with true poses: the two-ring capture of
frontend/tools/verify/trajectory.ts:5-12 (60 views, here at 640x360), batched PatchMatch, the
consistency filter, one batched TSDF pass of the filtered maps (256^3 x 2 mm) and surface nets,
and the fused dense cloud (2 mm cells), scored like the reference's own checks
(frontend/tools/verify/score.ts:199-296, the High profile's limits at :33-37): dense points'
distance to the true surface (median < 1.5 mm, 90th percentile < 5 mm), mesh vertex distance
(median < 1.5 mm, 90th percentile < 4 mm), completeness (> 85 % of true surface samples within
3 mm of a vertex) and the box's width, height and depth measured from the mesh like a caliper
(each within 3 mm of 140 x 120 x 100 mm).  The reference scores its whole pipeline (poses from its own tracking and
bundle adjustment); here the poses are exact, so this scores the hot path alone."""
import numpy as np
import pytest

from lr3d import scenes

pytestmark = pytest.mark.gpu

BOX_MIN = np.array([-0.07, 0.0, -0.05])
BOX_MAX = np.array([0.07, 0.12, 0.05])
SPHERE_C = np.array([0.13, 0.05, 0.07])
SPHERE_R = 0.05


def box_distance(p):
    """Signed distance to the box (negative inside)."""
    c, h = (BOX_MIN + BOX_MAX) / 2, (BOX_MAX - BOX_MIN) / 2
    q = np.abs(p - c) - h
    return np.linalg.norm(np.maximum(q, 0.0), axis=1) + np.minimum(q.max(axis=1), 0.0)


def surface_samples(rng, n):
    """Points on the surfaces the two rings see: the box's top and sides, the sphere above 3 cm,
    and the table top within 0.2 m of the centre away from the objects' footprints."""
    out = []
    # box: top and four sides, by area
    sx, sy, sz = BOX_MAX - BOX_MIN
    faces = [("top", sx * sz), ("x", sy * sz), ("x", sy * sz), ("z", sx * sy), ("z", sx * sy)]
    areas = np.array([a for _, a in faces])
    for f in rng.choice(len(faces), n // 3, p=areas / areas.sum()):
        u, v = rng.random(2)
        if faces[f][0] == "top":
            out.append([BOX_MIN[0] + u * sx, BOX_MAX[1], BOX_MIN[2] + v * sz])
        elif faces[f][0] == "x":
            out.append([BOX_MIN[0] if f == 1 else BOX_MAX[0], BOX_MIN[1] + u * sy, BOX_MIN[2] + v * sz])
        else:
            out.append([BOX_MIN[0] + u * sx, BOX_MIN[1] + v * sy, BOX_MIN[2] if f == 3 else BOX_MAX[2]])
    # sphere above y = 0.03
    while len(out) < 2 * n // 3:
        d = rng.standard_normal(3)
        p = SPHERE_C + SPHERE_R * d / np.linalg.norm(d)
        if p[1] > 0.03:
            out.append(list(p))
    # table top
    while len(out) < n:
        p = np.array([rng.uniform(-0.2, 0.2), 0.0, rng.uniform(-0.2, 0.2)])
        if box_distance(p[None])[0] > 0.01 and np.linalg.norm(p - SPHERE_C * [1, 0, 1]) > SPHERE_R + 0.01:
            out.append(list(p))
    return np.array(out)


def caliper(vertices, tol):
    """Box size from the mesh: width and depth between the medians of opposite side faces at
    mid-height, height from the median of the top face."""
    near = np.abs(box_distance(vertices)) < tol
    c, h = (BOX_MIN + BOX_MAX) / 2, (BOX_MAX - BOX_MIN) / 2
    sides = vertices[near & (vertices[:, 1] > 0.03) & (vertices[:, 1] < 0.09)]
    top = vertices[near & (np.abs(vertices[:, 1] - BOX_MAX[1]) < 2 * tol)]
    size = []
    for k in (0, 2):
        o = 2 - k
        face = sides[np.abs(sides[:, k] - c[k]) / h[k] > np.abs(sides[:, o] - c[o]) / h[o]]
        size.append(np.median(face[face[:, k] > c[k], k]) - np.median(face[face[:, k] < c[k], k]))
    return np.array([size[0], np.median(top[:, 1]) - BOX_MIN[1], size[1]])


def test_table_scene_cloud_and_mesh_are_accurate_and_complete(ctx):
    import lr3d
    from scipy.spatial import cKDTree
    w = scenes.table_rings().render()
    n = len(w.views)
    print(f"case: table scene, {n} views at 640x360, true poses", flush=True)
    st = lr3d.HipDenseStereo(ctx, w.views)
    for i, im in enumerate(w.images):
        st.set_image(i, im)
    st.solve_batched([(r, lr3d.SolveOptions(sources=w.sources[r], min_depth=0.1, max_depth=1.2,
                                            iterations=w.iterations)) for r in range(n)])
    for r in range(n):
        st.filter(r, w.sources[r], w.mesh_needed(r))
    spec = w.tsdf
    grid, layout = lr3d.dense_grid(spec.origin, spec.voxel, spec.voxels)
    vol = lr3d.HipTsdfVolume(ctx, grid, layout, spec.truncation_voxels)
    ivs = [lr3d.IntegrationView(v.intrinsics, v.pose, 1.2, v.width, v.height) for v in w.views]
    vol.integrate_batched(st, list(range(n)), ivs, map=lr3d.MAP_FILTERED)
    mesh = vol.extract_surface(0.6)
    verts = mesh.read()["positions"].astype(np.float64)
    mesh.destroy()
    vol.destroy()
    thumbs = [np.zeros((60, 80, 3), np.uint8) for _ in range(n)]
    cloud, cell = lr3d.fuse_depth_maps(st, list(range(n)), thumbs, spec.voxel, 4_000_000, map=lr3d.MAP_FILTERED)
    st.destroy()
    points = cloud["positions"].astype(np.float64)
    cerr = np.abs(w.scene.signed_distance(points[::7])) * 1000
    print(f"dense cloud: {len(points)} points in {cell * 1000:.1f} mm cells, error median "
          f"{np.median(cerr):.2f} mm, p90 {np.percentile(cerr, 90):.2f} mm (limits 1.5, 5)", flush=True)
    assert np.median(cerr) < 1.5 and np.percentile(cerr, 90) < 5.0

    err = np.abs(w.scene.signed_distance(verts)) * 1000
    med, p90 = float(np.median(err)), float(np.percentile(err, 90))
    samples = surface_samples(np.random.default_rng(17), 4000)
    reach = max(0.003, spec.voxel * 1.2)
    dist, _ = cKDTree(verts).query(samples)
    complete = float((dist < reach).mean())
    size = caliper(verts, max(0.0015, spec.voxel * 0.6))
    box_err = np.abs(size - (BOX_MAX - BOX_MIN)) * 1000
    print(f"{len(verts)} vertices: error median {med:.2f} mm, p90 {p90:.2f} mm (limits 1.5, 4); completeness "
          f"{complete * 100:.1f} % within {reach * 1000:.1f} mm (limit 85 %); box "
          f"{' x '.join(f'{v * 1000:.1f}' for v in size)} mm vs 140 x 120 x 100 (limit 3 mm)", flush=True)
    assert med < 1.5 and p90 < 4.0
    assert complete > 0.85
    assert box_err.max() < 3.0
