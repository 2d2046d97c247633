"""ctypes front-end of the CPU oracle (liblr3d_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the parity checker / reported CPU baseline.  The product library never uses it.

Every function restates the reference file:line named in oracle/lr3d_oracle.h.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# LR3D_ORACLE_LIB selects another build of the same sources (the sanitizer build, for one).
LIB_PATH = os.environ.get("LR3D_ORACLE_LIB") or os.path.join(HERE, "liblr3d_oracle.so")
MAX_SOURCES = 8


class OrcView(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32),
                ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("R", C.c_double * 9), ("t", C.c_double * 3)]


class OrcSrc(C.Structure):
    _fields_ = [("k", C.c_float * 4), ("R", C.c_float * 9), ("t", C.c_float * 3),
                ("layer", C.c_int32), ("width", C.c_int32), ("height", C.c_int32),
                ("map_off", C.c_int64), ("G", C.c_float * 9), ("e", C.c_float * 3)]


class OrcParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("ref", C.c_int32),
                ("k", C.c_float * 4),
                ("inv_min", C.c_float), ("inv_max", C.c_float), ("step", C.c_float),
                ("geo_weight", C.c_float),
                ("iteration", C.c_uint32), ("phase", C.c_uint32), ("seed", C.c_uint32),
                ("flags", C.c_uint32), ("needed", C.c_uint32), ("n_src", C.c_int32),
                ("topk", C.c_int32), ("ref_off", C.c_int64), ("src", OrcSrc * MAX_SOURCES)]


class OrcImages(C.Structure):
    _fields_ = [("n_layers", C.c_int32), ("pixels", C.POINTER(C.c_void_p)),
                ("widths", C.POINTER(C.c_int32)), ("heights", C.POINTER(C.c_int32))]


class OrcTsdfGrid(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("voxel", C.c_double), ("dims", C.c_int32 * 3),
                ("n_bricks", C.c_int32), ("coords", C.POINTER(C.c_int32)), ("truncation", C.c_double)]


class OrcIntegrationView(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32),
                ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("R", C.c_double * 9), ("t", C.c_double * 3), ("max_depth", C.c_double),
                ("confidence", C.c_double)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    V, P = C.c_void_p, C.POINTER
    u32, i32, f32, dbl = C.c_uint32, C.c_int32, C.c_float, C.c_double
    sig = {
        "orc_pcg": (u32, [u32]),
        "orc_host_seed": (u32, [i32, u32, u32, u32]),
        "orc_pixel_rng": (u32, [u32, u32, u32, u32]),
        "orc_random": (f32, [P(u32)]),
        "orc_pack_sample": (u32, [f32, P(f32), f32]),
        "orc_unpack_depth": (f32, [u32]),
        "orc_unpack_normal": (None, [u32, P(f32)]),
        "orc_unpack_cost": (u32, [u32]),
        "orc_f32_to_f16": (C.c_uint16, [f32]),
        "orc_f16_to_f32": (f32, [C.c_uint16]),
        "orc_make_params": (None, [P(OrcView), i32, i32, P(i32), i32, dbl, dbl, u32, u32, dbl, u32,
                                   u32, i32, P(OrcParams)]),
        "orc_window": (None, [P(OrcParams), P(OrcImages), i32, i32, V, V, V]),
        "orc_homography": (None, [P(OrcParams), i32, P(f32), f32, V]),
        "orc_view_cost": (f32, [P(OrcParams), P(OrcImages), i32, V, i32, i32, V, V, V]),
        "orc_hypothesis_cost": (f32, [P(OrcParams), P(OrcImages), V, i32, i32, V, V, V, P(f32), f32]),
        "orc_init": (None, [P(OrcParams), P(OrcImages), V, V, V, V]),
        "orc_propagate": (None, [P(OrcParams), P(OrcImages), V, V, V]),
        "orc_pack": (None, [P(OrcParams), V, V, V]),
        "orc_consistency": (None, [P(OrcParams), V, V]),
        "orc_reproject": (None, [P(OrcParams), V, i32, f32, f32, P(f32), V]),
        "orc_init_coarse": (None, [P(OrcParams), P(OrcImages), P(f32), i32, i32, V, V, V]),
        "orc_solve": (i32, [P(OrcView), i32, P(OrcImages), i32, P(i32), i32, dbl, dbl, V, i32, i32,
                            i32, i32, V, V, V]),
        "orc_filter": (None, [P(OrcView), i32, i32, P(i32), i32, i32, V, V]),
        "orc_dense_coords": (i32, [P(i32), V]),
        "orc_visible_bricks": (i32, [P(OrcTsdfGrid), P(OrcIntegrationView), V]),
        "orc_tsdf_integrate": (C.c_int64, [P(OrcTsdfGrid), P(OrcIntegrationView), V, V, V, V]),
        "orc_set_threads": (None, [i32]),
        "orc_get_threads": (i32, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return C.c_void_p(a.ctypes.data) if a is not None else None


def set_threads(n: int) -> None:
    load().orc_set_threads(int(n))


def get_threads() -> int:
    return load().orc_get_threads()


# ---- plain helpers -----------------------------------------------------------------------------

def pcg(v: int) -> int:
    return load().orc_pcg(v & 0xFFFFFFFF)


def host_seed(ref: int, iteration: int, phase: int, flags: int) -> int:
    return load().orc_host_seed(ref, iteration, phase, flags)


def pixel_rng(seed: int, iteration: int, phase: int, index: int) -> int:
    return load().orc_pixel_rng(seed, iteration, phase, index)


def random_stream(state: int, n: int) -> list:
    s = C.c_uint32(state)
    return [load().orc_random(C.byref(s)) for _ in range(n)]


def f32_to_f16(v: float) -> int:
    return load().orc_f32_to_f16(v)


def f16_to_f32(h: int) -> float:
    return load().orc_f16_to_f32(h)


def pack_sample(depth: float, normal, cost: float) -> int:
    n = (C.c_float * 3)(*normal)
    return load().orc_pack_sample(depth, n, cost)


def unpack_depth(p: int) -> float:
    return load().orc_unpack_depth(p)


def unpack_normal(p: int):
    out = (C.c_float * 3)()
    load().orc_unpack_normal(p, out)
    return np.array(list(out), dtype=np.float32)


def unpack_cost(p: int) -> int:
    return load().orc_unpack_cost(p)


# ---- stereo ------------------------------------------------------------------------------------

def view_struct(width, height, fx, fy, cx, cy, R, t) -> OrcView:
    v = OrcView()
    v.width, v.height = width, height
    v.fx, v.fy, v.cx, v.cy = fx, fy, cx, cy
    v.R[:] = [float(x) for x in np.asarray(R, dtype=np.float64).reshape(9)]
    v.t[:] = [float(x) for x in np.asarray(t, dtype=np.float64).reshape(3)]
    return v


def views_array(views) -> "C.Array":
    """From objects with .width .height .intrinsics(.fx .fy .cx .cy) .pose(.R .t)."""
    structs = [view_struct(v.width, v.height, v.intrinsics.fx, v.intrinsics.fy, v.intrinsics.cx,
                           v.intrinsics.cy, v.pose.R, v.pose.t) for v in views]
    return (OrcView * len(structs))(*structs)


class Images:
    """Keeps the u8 planes alive and exposes the orc_images view of them."""

    def __init__(self, images: Sequence[np.ndarray]):
        self.arrays = [np.ascontiguousarray(im, dtype=np.uint8) for im in images]
        n = len(self.arrays)
        self._ptrs = (C.c_void_p * n)(*[a.ctypes.data for a in self.arrays])
        self._w = (C.c_int32 * n)(*[a.shape[1] for a in self.arrays])
        self._h = (C.c_int32 * n)(*[a.shape[0] for a in self.arrays])
        self.struct = OrcImages(n, self._ptrs, self._w, self._h)


def make_params(views, ref, sources, min_depth, max_depth, iteration=0, phase=0, step=1.0,
                flags=0, needed=0, topk=0) -> OrcParams:
    arr = views_array(views)
    src = (C.c_int32 * max(1, len(sources)))(*sources)
    p = OrcParams()
    load().orc_make_params(arr, len(views), ref, src, len(sources), min_depth, max_depth, iteration,
                           phase, step, flags, needed, topk, C.byref(p))
    return p


def window(p: OrcParams, images: Images, x: int, y: int):
    values = np.empty(36, dtype=np.float32)
    weights = np.empty(36, dtype=np.float32)
    stats = np.empty(3, dtype=np.float32)
    load().orc_window(C.byref(p), C.byref(images.struct), x, y, _p(values), _p(weights), _p(stats))
    return values, weights, stats


def homography(p: OrcParams, s: int, n, rho: float) -> np.ndarray:
    H = np.empty(9, dtype=np.float32)
    nn = (C.c_float * 3)(*n)
    load().orc_homography(C.byref(p), s, nn, rho, _p(H))
    return H


def view_cost(p, images, s, H, x, y, win) -> float:
    values, weights, stats = win
    H = np.ascontiguousarray(H, dtype=np.float32)
    return load().orc_view_cost(C.byref(p), C.byref(images.struct), s, _p(H), x, y, _p(values),
                                _p(weights), _p(stats))


def hypothesis_cost(p, images, x, y, win, n, d, depths=None) -> float:
    values, weights, stats = win
    nn = (C.c_float * 3)(*n)
    return load().orc_hypothesis_cost(C.byref(p), C.byref(images.struct), _p(depths), x, y,
                                      _p(values), _p(weights), _p(stats), nn, d)


def init(p, images, prior=None, depths=None, hyp=None, cost=None):
    n = p.width * p.height
    hyp = np.zeros((n, 4), dtype=np.float32) if hyp is None else hyp
    cost = np.zeros(n, dtype=np.float32) if cost is None else cost
    load().orc_init(C.byref(p), C.byref(images.struct), _p(prior), _p(depths), _p(hyp), _p(cost))
    return hyp, cost


def propagate(p, images, hyp, cost, depths=None):
    load().orc_propagate(C.byref(p), C.byref(images.struct), _p(depths), _p(hyp), _p(cost))
    return hyp, cost


def pack(p, hyp, cost) -> np.ndarray:
    out = np.empty(p.width * p.height, dtype=np.uint32)
    load().orc_pack(C.byref(p), _p(hyp), _p(cost), _p(out))
    return out


def map_offsets(views) -> np.ndarray:
    """Pixel offset of every layer's packed map in the map pool, plus the pool size at the end:
    layers one after another, each width*height (orc_map_offset)."""
    return np.concatenate([[0], np.cumsum([v.width * v.height for v in views])]).astype(np.int64)


def map_pool(views) -> np.ndarray:
    """An empty map pool for `views` (all maps 0)."""
    return np.zeros(int(map_offsets(views)[-1]), dtype=np.uint32)


def layer_map(views, depths: np.ndarray, layer: int) -> np.ndarray:
    """Layer `layer`'s packed map in the pool `depths` (a view, not a copy)."""
    off = map_offsets(views)
    return depths[off[layer]:off[layer + 1]]


def solve(views, images: Images, ref: int, sources: Sequence[int], min_depth: float,
          max_depth: float, prior: Optional[np.ndarray] = None, geometric: bool = False,
          iterations: int = 5, topk: int = 0, pack: bool = True,
          depths: Optional[np.ndarray] = None):
    """DS:185-239.  depths: the map pool of all layers (map_pool; layers may differ in size)."""
    arr = views_array(views)
    W, H = views[ref].width, views[ref].height
    if depths is None:
        depths = map_pool(views)
    assert depths.size == map_offsets(views)[-1], "depths must be the map pool of all views"
    hyp = np.zeros((W * H, 4), dtype=np.float32)
    cost = np.zeros(W * H, dtype=np.float32)
    src = (C.c_int32 * max(1, len(sources)))(*sources)
    if prior is not None:
        prior = np.ascontiguousarray(prior, dtype=np.float32).reshape(-1)
    load().orc_solve(arr, len(views), C.byref(images.struct), ref, src, len(sources), min_depth,
                     max_depth, _p(prior), 1 if geometric else 0, iterations, topk, 1 if pack else 0,
                     _p(depths), _p(hyp), _p(cost))
    return depths, hyp, cost


def filter_map(views, ref, checks, needed, depths) -> np.ndarray:
    """DS:250-263 over the map pool `depths`."""
    assert depths.size == map_offsets(views)[-1], "depths must be the map pool of all views"
    arr = views_array(views)
    W, H = views[ref].width, views[ref].height
    out = np.zeros(W * H, dtype=np.uint32)
    src = (C.c_int32 * max(1, len(checks)))(*checks)
    load().orc_filter(arr, len(views), ref, src, len(checks), needed, _p(depths), _p(out))
    return out


def reproject(p, depths, s: int, px: float, py: float, x) -> np.ndarray:
    """PM:166-190: (distance, depth, 0, 1) or all -1."""
    out = np.empty(4, dtype=np.float32)
    xx = (C.c_float * 3)(*[float(v) for v in x])
    load().orc_reproject(C.byref(p), _p(depths), s, px, py, xx, _p(out))
    return out


def init_coarse(p, images, ck, coarse_w, coarse_h, coarse_hyp):
    n = p.width * p.height
    hyp = np.zeros((n, 4), dtype=np.float32)
    cost = np.zeros(n, dtype=np.float32)
    k = (C.c_float * 4)(*ck)
    ch = np.ascontiguousarray(coarse_hyp, dtype=np.float32)
    load().orc_init_coarse(C.byref(p), C.byref(images.struct), k, coarse_w, coarse_h, _p(ch),
                           _p(hyp), _p(cost))
    return hyp, cost


# ---- TSDF --------------------------------------------------------------------------------------

class Grid:
    def __init__(self, origin, voxel, dims, coords, truncation):
        self.coords = np.ascontiguousarray(coords, dtype=np.int32).reshape(-1)
        g = OrcTsdfGrid()
        g.origin[:] = list(origin)
        g.voxel = voxel
        g.dims[:] = list(dims)
        g.n_bricks = len(self.coords) // 4
        g.coords = self.coords.ctypes.data_as(C.POINTER(C.c_int32))
        g.truncation = truncation
        self.struct = g
        self.n_bricks = g.n_bricks


def integration_view(width, height, fx, fy, cx, cy, R, t, max_depth, confidence=0.0):
    v = OrcIntegrationView()
    v.width, v.height = width, height
    v.fx, v.fy, v.cx, v.cy = fx, fy, cx, cy
    v.R[:] = [float(x) for x in np.asarray(R, dtype=np.float64).reshape(9)]
    v.t[:] = [float(x) for x in np.asarray(t, dtype=np.float64).reshape(3)]
    v.max_depth = max_depth
    v.confidence = confidence
    return v


def visible_bricks(grid: Grid, view) -> np.ndarray:
    flags = np.zeros(grid.n_bricks, dtype=np.uint8)
    load().orc_visible_bricks(C.byref(grid.struct), C.byref(view), _p(flags))
    return flags


def tsdf_integrate(grid: Grid, view, depth: np.ndarray, prior, tsdf, weight) -> int:
    depth = np.ascontiguousarray(depth, dtype=np.uint32).reshape(-1)
    if prior is not None:
        prior = np.ascontiguousarray(prior, dtype=np.float32).reshape(-1)
    return load().orc_tsdf_integrate(C.byref(grid.struct), C.byref(view), _p(depth), _p(prior),
                                     _p(tsdf), _p(weight))


# ---- FP64 cost and gradient (adjoint_oracle.c) -------------------------------------------------

class Orc64Src(C.Structure):
    _fields_ = [("k", C.c_double * 4), ("R", C.c_double * 9), ("t", C.c_double * 3),
                ("layer", C.c_int32), ("width", C.c_int32), ("height", C.c_int32)]


class Orc64Params(C.Structure):
    _fields_ = [("k", C.c_double * 4), ("inv_min", C.c_double), ("inv_max", C.c_double),
                ("n_src", C.c_int32), ("topk", C.c_int32), ("src", Orc64Src * MAX_SOURCES)]


class Orc64Result(C.Structure):
    _fields_ = [("cost", C.c_double), ("d_depth", C.c_double), ("d_normal", C.c_double * 3),
                ("d_R", (C.c_double * 9) * MAX_SOURCES), ("d_t", (C.c_double * 3) * MAX_SOURCES),
                ("selected", C.c_int32 * MAX_SOURCES), ("view_cost", C.c_double * MAX_SOURCES),
                ("min_kink", C.c_double)]


def params64(views, ref, sources, min_depth, max_depth, topk=0, rel=None) -> Orc64Params:
    """FP64 parameters; `rel` optionally overrides the relative poses [(R 3x3, t 3)] per source."""
    rv = views[ref]
    p = Orc64Params()
    p.k[:] = [rv.intrinsics.fx, rv.intrinsics.fy, rv.intrinsics.cx, rv.intrinsics.cy]
    p.inv_min, p.inv_max = 1.0 / min_depth, 1.0 / max_depth
    p.n_src, p.topk = len(sources), topk
    Ra, ta = np.asarray(rv.pose.R, float), np.asarray(rv.pose.t, float)
    for s, layer in enumerate(sources):
        sv = views[layer]
        if rel is None:
            R = np.asarray(sv.pose.R, float) @ Ra.T
            t = np.asarray(sv.pose.t, float) - R @ ta
        else:
            R, t = rel[s]
        o = p.src[s]
        o.k[:] = [sv.intrinsics.fx, sv.intrinsics.fy, sv.intrinsics.cx, sv.intrinsics.cy]
        o.R[:] = [float(x) for x in np.asarray(R).reshape(9)]
        o.t[:] = [float(x) for x in np.asarray(t).reshape(3)]
        o.layer, o.width, o.height = layer, sv.width, sv.height
    return p


def cost_grad64(p: Orc64Params, images: Images, ref: int, x: int, y: int, n, d: float) -> Orc64Result:
    lib = load()
    if not hasattr(lib, "_cg64"):
        lib.orc64_cost_grad.restype = C.c_int32
        lib.orc64_cost_grad.argtypes = [C.POINTER(Orc64Params), C.POINTER(OrcImages), C.c_int32, C.c_int32,
                                        C.c_int32, C.POINTER(C.c_double), C.c_double, C.POINTER(Orc64Result)]
        lib._cg64 = True
    r = Orc64Result()
    nn = (C.c_double * 3)(*[float(v) for v in n])
    lib.orc64_cost_grad(C.byref(p), C.byref(images.struct), ref, x, y, nn, float(d), C.byref(r))
    return r


# ---- FP64 TSDF integration adjoint (tsdf_adjoint_oracle.c) -------------------------------------

class Orc64DepthInput(C.Structure):
    _fields_ = [("depth", C.c_void_p), ("normal", C.c_void_p), ("cost", C.c_void_p), ("prior", C.c_void_p)]


def upsample_depth(meters, gw, gh, width, height) -> np.ndarray:
    lib = load()
    if not hasattr(lib, "_ud"):
        lib.orc_upsample_depth.restype = None
        lib.orc_upsample_depth.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        lib._ud = True
    m = np.ascontiguousarray(meters, dtype=np.float32).reshape(-1)
    out = np.zeros(int(width) * int(height), dtype=np.float32)
    lib.orc_upsample_depth(m.ctypes.data, gw, gh, width, height, out.ctypes.data)
    return out


def decode_maps(packed_maps):
    """Packed u32 maps -> (depth f64, normal f64 (W*H, 3), cost code u8) with the decoders of PK."""
    out = []
    lib = load()
    if not hasattr(lib, "_dm"):
        lib.orc_decode_map.restype = None
        lib.orc_decode_map.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        lib._dm = True
    for m in packed_maps:
        m = np.ascontiguousarray(m, dtype=np.uint32).reshape(-1)
        depth = np.zeros(m.size, dtype=np.float64)
        normal = np.zeros((m.size, 3), dtype=np.float64)
        cost = np.zeros(m.size, dtype=np.uint8)
        lib.orc_decode_map(m.ctypes.data, m.size, depth.ctypes.data, normal.ctypes.data, cost.ctypes.data)
        out.append((depth, normal, cost))
    return out


def _depth_inputs(decoded, priors):
    n = len(decoded)
    keep = []
    inputs = (Orc64DepthInput * n)()
    for i, (dep, nor, cost) in enumerate(decoded):
        pr = None if priors is None or priors[i] is None else np.ascontiguousarray(priors[i], dtype=np.float64).reshape(-1)
        dep, nor, cost = (np.ascontiguousarray(dep, dtype=np.float64), np.ascontiguousarray(nor, dtype=np.float64),
                          np.ascontiguousarray(cost, dtype=np.uint8))
        keep += [dep, nor, cost, pr]
        inputs[i] = Orc64DepthInput(dep.ctypes.data, nor.ctypes.data, cost.ctypes.data,
                                    pr.ctypes.data if pr is not None else None)
    return inputs, keep


def tsdf_backward64(grid: Grid, views, decoded, priors, voxels, targets):
    """FP64 replay + adjoint; returns dict(tsdf, kink, grad_depth (list per view), grad_pose (n, 12), loss)."""
    lib = load()
    if not hasattr(lib, "_tb64"):
        lib.orc64_tsdf_backward.restype = C.c_int32
        lib.orc64_tsdf_backward.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                            C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        lib._tb64 = True
    n = len(views)
    varr = (OrcIntegrationView * n)(*views)
    inputs, keep = _depth_inputs(decoded, priors)
    vox = np.ascontiguousarray(voxels, dtype=np.uint32)
    tgt = np.ascontiguousarray(targets, dtype=np.float64)
    sizes = [v.width * v.height for v in views]
    tsdf = np.zeros(vox.size)
    kink = np.zeros(vox.size)
    gd = np.zeros(sum(sizes))
    gp = np.zeros((n, 12))
    loss = np.zeros(1)
    rc = lib.orc64_tsdf_backward(C.addressof(grid.struct), C.addressof(varr), C.addressof(inputs), n,
                                 vox.ctypes.data, tgt.ctypes.data, vox.size, tsdf.ctypes.data, gd.ctypes.data,
                                 gp.ctypes.data, kink.ctypes.data, loss.ctypes.data)
    assert rc == 0
    splits = np.split(gd, np.cumsum(sizes)[:-1])
    return {"tsdf": tsdf, "kink": kink, "grad_depth": splits, "grad_pose": gp, "loss": float(loss[0])}


def visibility_flags(grid: Grid, views) -> np.ndarray:
    """Culling flags of every view, [view][brick] (uint8)."""
    return np.stack([visible_bricks(grid, v) for v in views]).astype(np.uint8)


def tsdf_loss64(grid: Grid, views, decoded, priors, voxels, targets, vis=None, per_voxel=False):
    """The FP64 loss of tsdf_backward64 alone (OpenMP over voxels, summed in voxel order); vis from
    visibility_flags or None.  per_voxel=True returns each voxel's (tsdf - target)^2 instead."""
    lib = load()
    if not hasattr(lib, "_tl64"):
        lib.orc64_tsdf_loss.restype = C.c_int32
        lib.orc64_tsdf_loss.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                        C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]
        lib._tl64 = True
    n = len(views)
    varr = (OrcIntegrationView * n)(*views)
    inputs, keep = _depth_inputs(decoded, priors)
    vox = np.ascontiguousarray(voxels, dtype=np.uint32)
    tgt = np.ascontiguousarray(targets, dtype=np.float64)
    flags = None if vis is None else np.ascontiguousarray(vis, dtype=np.uint8)
    loss = np.zeros(1)
    terms = np.zeros(vox.size)
    rc = lib.orc64_tsdf_loss(C.addressof(grid.struct), C.addressof(varr), C.addressof(inputs), n, vox.ctypes.data,
                             tgt.ctypes.data, vox.size, flags.ctypes.data if flags is not None else None,
                             loss.ctypes.data, terms.ctypes.data)
    assert rc == 0
    return terms if per_voxel else float(loss[0])


# ---- surface nets and mesh post-processing (mesh_oracle.c) ----------------------------------------

class OrcSdfBricks(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("voxel", C.c_double), ("dims", C.c_int32 * 3), ("n_bricks", C.c_int32),
                ("lookup", C.c_void_p), ("tsdf", C.c_void_p), ("weight", C.c_void_p)]


class OrcMesh(C.Structure):
    _fields_ = [("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("indices", C.POINTER(C.c_uint32)), ("n_vertices", C.c_int64), ("n_indices", C.c_int64)]


def _mesh_lib():
    lib = load()
    if not hasattr(lib, "_mesh"):
        lib.orc_extract_surface.restype = C.c_int32
        lib.orc_extract_surface.argtypes = [C.POINTER(OrcSdfBricks), C.c_double, C.POINTER(OrcMesh)]
        lib.orc_remove_small_components.restype = C.c_int32
        lib.orc_remove_small_components.argtypes = [C.POINTER(OrcMesh), C.c_double, C.c_int32, C.POINTER(OrcMesh)]
        lib.orc_mesh_free.restype = None
        lib.orc_mesh_free.argtypes = [C.POINTER(OrcMesh)]
        lib._mesh = True
    return lib


def _take(lib, m):
    nv, ni = m.n_vertices, m.n_indices
    out = {"positions": np.ctypeslib.as_array(m.positions, (max(nv, 1) * 3,))[:nv * 3].reshape(-1, 3).copy(),
           "normals": np.ctypeslib.as_array(m.normals, (max(nv, 1) * 3,))[:nv * 3].reshape(-1, 3).copy(),
           "indices": np.ctypeslib.as_array(m.indices, (max(ni, 1),))[:ni].copy()}
    return out


def extract_surface(origin, voxel, dims, lookup, tsdf, weight, min_weight):
    """surface-nets.ts extractSurface: dict(positions (V,3) f32, normals (V,3) f32, indices u32)."""
    lib = _mesh_lib()
    lookup = np.ascontiguousarray(lookup, dtype=np.int32)
    tsdf = np.ascontiguousarray(tsdf, dtype=np.float32)
    weight = np.ascontiguousarray(weight, dtype=np.float32)
    v = OrcSdfBricks()
    v.origin[:] = list(origin)
    v.voxel = voxel
    v.dims[:] = list(dims)
    v.n_bricks = tsdf.size // 512
    v.lookup, v.tsdf, v.weight = lookup.ctypes.data, tsdf.ctypes.data, weight.ctypes.data
    m = OrcMesh()
    assert lib.orc_extract_surface(C.byref(v), float(min_weight), C.byref(m)) == 0
    out = _take(lib, m)
    lib.orc_mesh_free(C.byref(m))
    return out


def remove_small_components(mesh, fraction, min_faces):
    """mesh.ts removeSmallComponents on a dict mesh."""
    lib = _mesh_lib()
    p = np.ascontiguousarray(mesh["positions"], dtype=np.float32).reshape(-1)
    n = np.ascontiguousarray(mesh["normals"], dtype=np.float32).reshape(-1)
    i = np.ascontiguousarray(mesh["indices"], dtype=np.uint32).reshape(-1)
    m = OrcMesh(p.ctypes.data_as(C.POINTER(C.c_float)), n.ctypes.data_as(C.POINTER(C.c_float)),
                i.ctypes.data_as(C.POINTER(C.c_uint32)), p.size // 3, i.size)
    o = OrcMesh()
    assert lib.orc_remove_small_components(C.byref(m), float(fraction), int(min_faces), C.byref(o)) == 0
    out = _take(lib, o)
    lib.orc_mesh_free(C.byref(o))
    return out


# ---- dense-cloud fusion (fusion_oracle.c) -------------------------------------------------------

class OrcFusionView(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("fx", C.c_double), ("fy", C.c_double),
                ("cx", C.c_double), ("cy", C.c_double), ("R", C.c_double * 9), ("t", C.c_double * 3),
                ("packed", C.c_void_p), ("thumb_width", C.c_int32), ("thumb_height", C.c_int32), ("rgb", C.c_void_p)]


class OrcCloud(C.Structure):
    _fields_ = [("n", C.c_int64), ("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("colors", C.POINTER(C.c_uint8)), ("support", C.POINTER(C.c_uint16))]


def fuse_depth_maps(views, maps, thumbs, cell, cell_budget, min_support=1):
    """depth-fusion.ts fuseDepthMaps: views (StereoView), maps (packed u32 per view), thumbs (H, W, 3)
    u8; returns (dict(positions, normals, colors, support), final cell) in the accumulator's slot order."""
    lib = load()
    if not hasattr(lib, "_fuse"):
        lib.orc_fuse_depth_maps.restype = C.c_int32
        lib.orc_fuse_depth_maps.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_int64, C.c_int32,
                                            C.POINTER(OrcCloud), C.POINTER(C.c_double)]
        lib.orc_cloud_free.restype = None
        lib.orc_cloud_free.argtypes = [C.POINTER(OrcCloud)]
        lib._fuse = True
    keep = []
    arr = (OrcFusionView * len(views))()
    for i, (v, m, th) in enumerate(zip(views, maps, thumbs)):
        m = np.ascontiguousarray(m, dtype=np.uint32)
        th = np.ascontiguousarray(th, dtype=np.uint8)
        keep += [m, th]
        k = v.intrinsics
        arr[i] = OrcFusionView(v.width, v.height, k.fx, k.fy, k.cx, k.cy,
                               (C.c_double * 9)(*np.asarray(v.pose.R, float).reshape(9)),
                               (C.c_double * 3)(*np.asarray(v.pose.t, float).reshape(3)),
                               m.ctypes.data, th.shape[1], th.shape[0], th.ctypes.data)
    out = OrcCloud()
    final = C.c_double()
    assert lib.orc_fuse_depth_maps(arr, len(views), float(cell), int(cell_budget), int(min_support), C.byref(out),
                                   C.byref(final)) == 0
    n = out.n
    res = {"positions": np.ctypeslib.as_array(out.positions, (max(n, 1) * 3,))[:n * 3].reshape(-1, 3).copy(),
           "normals": np.ctypeslib.as_array(out.normals, (max(n, 1) * 3,))[:n * 3].reshape(-1, 3).copy(),
           "colors": np.ctypeslib.as_array(out.colors, (max(n, 1) * 3,))[:n * 3].reshape(-1, 3).copy(),
           "support": np.ctypeslib.as_array(out.support, (max(n, 1),))[:n].copy()}
    lib.orc_cloud_free(C.byref(out))
    return res, final.value
