"""Pins of the oracle's surface nets and mesh post-processing (oracle/mesh_oracle.c) with the
reference's own known answers (frontend/src/app/vision/fusion/fusion.spec.ts:6-83): a sparse brick
TSDF of a sphere (radius 5 cm, 2 mm voxels, truncation 4 voxels) meshes to > 5000 triangles whose
vertices lie within 0.35 voxel of the sphere, the surface is closed (every edge in exactly two
triangles) with > 99 % outward normals, and a far floating triangle is dropped by
removeSmallComponents(0.01, 10) while the sphere is kept whole.  Plus closed forms: a single
sign-changing cell puts its vertex at the mean of its edge crossings."""
import numpy as np

import oracle as O
from lr3d import BrickGrid


def sphere_volume(radius=0.05, voxel=0.002):
    """fusion.spec.ts:6-32"""
    extent = radius * 2 + voxel * 8 * 2
    origin = (-extent / 2,) * 3
    grid = BrickGrid(origin, voxel, (extent,) * 3)
    mu = voxel * 4
    for a in range(2000):
        z = (a / 1999) * 2 - 1
        phi = a * 2.399963
        r = np.sqrt(1 - z * z)
        grid.mark(np.cos(phi) * r * radius, np.sin(phi) * r * radius, z * radius, mu)
    coords, lookup, count = grid.build()
    c = coords.reshape(-1, 4)[:, :3]
    local = np.arange(512)
    v = np.stack([local & 7, (local >> 3) & 7, local >> 6], axis=1)
    p = np.asarray(origin) + ((c[:, None, :] * 8 + v[None]) + 0.5) * voxel
    sdf = np.linalg.norm(p, axis=2) - radius
    tsdf = np.clip(sdf / mu, -1, 1).astype(np.float32).reshape(-1)
    weight = np.where(sdf > -mu, 2.0, 0.0).astype(np.float32).reshape(-1)
    return dict(origin=origin, voxel=voxel, dims=grid.dims, lookup=lookup, tsdf=tsdf, weight=weight)


def test_sphere_surface_nets_known_answers():
    vol = sphere_volume()
    m = O.extract_surface(vol["origin"], vol["voxel"], vol["dims"], vol["lookup"], vol["tsdf"], vol["weight"], 0.5)
    tris = m["indices"].reshape(-1, 3)
    assert len(tris) > 5000
    r = np.linalg.norm(m["positions"].astype(np.float64), axis=1)
    assert np.abs(r - 0.05).max() < 0.002 * 0.35
    edges = {}
    for t in tris:
        for k in range(3):
            a, b = int(t[k]), int(t[(k + 1) % 3])
            key = (min(a, b), max(a, b))
            edges[key] = edges.get(key, 0) + 1
    assert all(n == 2 for n in edges.values())
    outward = (np.sum(m["normals"] * m["positions"], axis=1) > 0).mean()
    assert outward > 0.99
    # unit normals
    assert np.allclose(np.linalg.norm(m["normals"], axis=1), 1, atol=1e-6)


def test_remove_small_components_drops_floaters():
    vol = sphere_volume()
    m = O.extract_surface(vol["origin"], vol["voxel"], vol["dims"], vol["lookup"], vol["tsdf"], vol["weight"], 0.5)
    nv = len(m["positions"])
    floater = {"positions": np.concatenate([m["positions"], [[1, 1, 1], [1.001, 1, 1], [1, 1.001, 1]]]).astype(np.float32),
               "normals": np.concatenate([m["normals"], np.zeros((3, 3))]).astype(np.float32),
               "indices": np.concatenate([m["indices"], [nv, nv + 1, nv + 2]]).astype(np.uint32)}
    cleaned = O.remove_small_components(floater, 0.01, 10)
    assert len(cleaned["indices"]) == len(m["indices"])
    # the sphere survives unchanged: same triangles, vertices in first-use order
    assert np.array_equal(cleaned["positions"][cleaned["indices"]], m["positions"][m["indices"]])


def test_single_cell_vertex_is_the_mean_of_edge_crossings():
    # one brick; corner (0,0,0) of cell (3,3,3) negative, every other corner positive
    tsdf = np.full(512, 0.5, dtype=np.float32)
    weight = np.full(512, 1.0, dtype=np.float32)
    tsdf[3 + 8 * 3 + 64 * 3] = -0.25
    m = O.extract_surface((0.0, 0.0, 0.0), 0.01, (1, 1, 1), np.array([0], np.int32), tsdf, weight, 0.5)
    # eight cells share that corner; the one at (3,3,3) has crossings at t = 0.25/0.75 = 1/3 on its
    # three edges from corner 0: mean offset (1/9, 1/9, 1/9)
    expect = (3 + 0.5 + 1 / 9) * 0.01
    cell = [i for i, p in enumerate(m["positions"]) if np.allclose(p, expect, atol=1e-7)]
    assert len(cell) == 1
    assert len(m["positions"]) == 8 and len(m["indices"]) == 6 * 6  # 6 quads around the corner
