"""HipDenseStereo: the reference's DenseStereo (frontend/src/app/vision/mvs/dense-stereo.ts,
DS:61-371) on the lr3d C ABI.

Same members and argument meaning: ``set_image(layer, gray)``, ``solve(ref, options)``,
``filter(ref, sources, needed)``, ``filtered_map``, ``depth_map``, ``hypotheses``,
``destroy``; plus ``solve_batched`` (every view in shared launches, replayed from a HIP graph).
Per-view image sizes are allowed (mixed-resolution batches).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _native as N
from .camera import StereoView
from .runtime import Context


@dataclass
class SolveOptions:
    """DS:16-29 SolveOptions, plus the top-k override (0 = the reference rule)."""
    sources: Sequence[int]
    min_depth: float
    max_depth: float
    prior: Optional[np.ndarray] = None
    geometric: bool = False
    iterations: int = 5
    from_coarse: bool = False
    pack: bool = True
    topk: int = 0


def _view_struct(v: StereoView) -> N.View:
    s = N.View()
    s.width, s.height = int(v.width), int(v.height)
    k = v.intrinsics
    s.fx, s.fy, s.cx, s.cy = k.fx, k.fy, k.cx, k.cy
    s.R[:] = [float(x) for x in np.asarray(v.pose.R, dtype=np.float64).reshape(9)]
    s.t[:] = [float(x) for x in np.asarray(v.pose.t, dtype=np.float64).reshape(3)]
    return s


class HipDenseStereo:
    """GPU PatchMatch for every keyframe (DS:56-61): one device pool per buffer kind holds every
    view, so the geometric pass and the consistency filter read any neighbour without copies."""

    def __init__(self, ctx: Context, views: Sequence[StereoView], prior_pool: bool = False):
        """prior_pool: allocate the prior pool now (needed by solve_batched_noalloc for priors)."""
        self._lib = N.load()
        self.ctx = ctx
        # own copies: set_pose replaces poses here, never in the caller's objects
        self.views = [StereoView(v.intrinsics, v.pose, v.width, v.height) for v in views]
        arr = (N.View * len(self.views))(*[_view_struct(v) for v in self.views])
        handle = C.c_void_p()
        flags = N.STEREO_PRIOR_POOL if prior_pool else 0
        N.check(self._lib.lr3d_stereo_create(ctx.handle, arr, len(self.views), flags, C.byref(handle)))
        self.handle = handle
        self._keep = []

    # DS:80-87 exposes width/height of the (single-size) reference set.
    @property
    def width(self) -> int:
        return self.views[0].width

    @property
    def height(self) -> int:
        return self.views[0].height

    def size(self, layer: int) -> tuple[int, int]:
        return self.views[layer].width, self.views[layer].height

    def set_image(self, layer: int, gray: np.ndarray) -> None:
        """DS:175-182"""
        w, h = self.size(layer)
        g = np.ascontiguousarray(gray, dtype=np.uint8).reshape(h, w)
        N.check(self._lib.lr3d_stereo_set_image(self.handle, layer, N.ptr(g), w, h))

    def _desc(self, ref: int, o: SolveOptions):
        d = N.SolveDesc()
        d.ref = ref
        sources = list(o.sources)[: N.MAX_SOURCES]  # DS:186
        d.n_sources = len(sources)
        for i, s in enumerate(sources):
            d.sources[i] = int(s)
        d.min_depth, d.max_depth = float(o.min_depth), float(o.max_depth)
        prior = None
        if o.prior is not None:
            w, h = self.size(ref)
            prior = np.ascontiguousarray(o.prior, dtype=np.float32).reshape(h * w)
        d.prior = N.float_ptr(prior)
        d.geometric = 1 if o.geometric else 0
        d.iterations = int(o.iterations)
        d.from_coarse = 1 if o.from_coarse else 0
        d.pack = 1 if o.pack else 0
        d.topk = int(o.topk)
        return d, prior

    def solve(self, ref: int, options: SolveOptions) -> None:
        """DS:185-239: one launch per kernel per red-black phase."""
        if len(options.sources) == 0:
            return  # DS:187
        d, keep = self._desc(ref, options)
        N.check(self._lib.lr3d_stereo_solve(self.handle, C.byref(d)))

    def solve_batched(self, solves: Sequence[tuple[int, SolveOptions]]) -> None:
        """All (ref, options) pairs in shared launches; equals the solves one by one, in order
        (a geometric pass reading an earlier solve's map runs after that solve's pack)."""
        descs, keep = [], []
        for ref, o in solves:
            d, k = self._desc(ref, o)
            descs.append(d)
            keep.append(k)
        arr = (N.SolveDesc * len(descs))(*descs)
        N.check(self._lib.lr3d_stereo_solve_batched(self.handle, arr, len(descs)))

    def solve_batched_noalloc(self, solves: Sequence[tuple[int, SolveOptions]]) -> None:
        """solve_batched without device allocation: the launch plan in the library's fixed
        per-GPU scratch array, eager launches (at most NOALLOC_MAX_SOLVES solves)."""
        descs, keep = [], []
        for ref, o in solves:
            d, k = self._desc(ref, o)
            descs.append(d)
            keep.append(k)
        arr = (N.SolveDesc * max(1, len(descs)))(*descs)
        N.check(self._lib.lr3d_stereo_solve_batched_noalloc(self.handle, arr, len(descs)))

    def autotune(self, solves: Sequence[tuple[int, SolveOptions]], repeats: int = 3) -> dict:
        """Times the candidate launch shapes on this batch and keeps the fastest; returns them."""
        descs = [self._desc(r, o)[0] for r, o in solves]
        arr = (N.SolveDesc * len(descs))(*descs)
        N.check(self._lib.lr3d_stereo_autotune(self.handle, arr, len(descs), int(repeats)))
        return self.tiles()

    def tiles(self) -> dict:
        out = {}
        for k in ("init", "propagate", "pack"):
            tx, ty = C.c_int32(), C.c_int32()
            N.check(self._lib.lr3d_stereo_get_tile(self.handle, k.encode(), C.byref(tx), C.byref(ty)))
            out[k] = (tx.value, ty.value)
        return out

    def set_tile(self, kernel: str, tx: int, ty: int) -> None:
        N.check(self._lib.lr3d_stereo_set_tile(self.handle, kernel.encode(), int(tx), int(ty)))

    def filter(self, ref: int, sources: Sequence[int], needed: int) -> None:
        """DS:250-263: consistency check of `ref` into its filtered map (stays on the GPU)."""
        checked = list(sources)[: N.MAX_SOURCES]
        arr = (C.c_int32 * max(1, len(checked)))(*checked)
        N.check(self._lib.lr3d_stereo_filter(self.handle, ref, arr, len(checked), int(needed)))

    def filtered_map(self, ref: int, sources: Sequence[int], needed: int) -> np.ndarray:
        """DS:301-306"""
        self.filter(ref, sources, needed)
        return self.read_filtered(ref)

    def read_filtered(self, ref: int) -> np.ndarray:
        w, h = self.size(ref)
        out = np.empty(w * h, dtype=np.uint32)
        N.check(self._lib.lr3d_stereo_read_filtered(self.handle, ref, N.ptr(out)))
        return out

    def depth_map(self, ref: int) -> np.ndarray:
        """DS:309-312: packed u32 map of a layer."""
        w, h = self.size(ref)
        out = np.empty(w * h, dtype=np.uint32)
        N.check(self._lib.lr3d_stereo_read_depth(self.handle, ref, N.ptr(out)))
        return out

    def write_depth(self, ref: int, packed: np.ndarray) -> None:
        w, h = self.size(ref)
        p = np.ascontiguousarray(packed, dtype=np.uint32).reshape(w * h)
        N.check(self._lib.lr3d_stereo_write_depth(self.handle, ref, N.ptr(p)))

    def hypotheses(self, ref: int) -> tuple[np.ndarray, np.ndarray]:
        """DS:242-244: planes (n.xyz, d) and costs of a layer's last solve."""
        w, h = self.size(ref)
        hyp = np.empty((w * h, 4), dtype=np.float32)
        cost = np.empty(w * h, dtype=np.float32)
        N.check(self._lib.lr3d_stereo_read_hyp(self.handle, ref, N.ptr(hyp), N.ptr(cost)))
        return hyp, cost

    def set_pose(self, layer: int, pose) -> None:
        """Replace a layer's world -> camera pose (pose refinement)."""
        R = (C.c_double * 9)(*[float(v) for v in np.asarray(pose.R, dtype=np.float64).reshape(9)])
        t = (C.c_double * 3)(*[float(v) for v in np.asarray(pose.t, dtype=np.float64).reshape(3)])
        N.check(self._lib.lr3d_stereo_set_pose(self.handle, layer, R, t))
        self.views[layer].pose = pose

    def write_hyp(self, layer: int, hyp: np.ndarray) -> None:
        w, h = self.size(layer)
        a = np.ascontiguousarray(hyp, dtype=np.float32).reshape(w * h * 4)
        N.check(self._lib.lr3d_stereo_write_hyp(self.handle, layer, N.ptr(a)))

    def ncc_backward(self, ref: int, options: SolveOptions, weights: Optional[np.ndarray] = None) -> dict:
        """Forward cost and adjoint of L = sum_p weights[p] * C_p for the current planes of `ref`."""
        w, h = self.size(ref)
        d, _ = self._desc(ref, options)
        wts = None if weights is None else np.ascontiguousarray(weights, dtype=np.float32).reshape(w * h)
        cost = np.empty(w * h, dtype=np.float32)
        gd = np.empty(w * h, dtype=np.float32)
        gn = np.empty((w * h, 3), dtype=np.float32)
        gp = np.zeros((N.MAX_SOURCES, 12), dtype=np.float64)
        loss = np.zeros(1, dtype=np.float64)
        N.check(self._lib.lr3d_ncc_backward(self.handle, C.byref(d), N.ptr(wts) if wts is not None else None,
                                            N.ptr(cost), N.ptr(gd), N.ptr(gn), N.ptr(gp), N.ptr(loss)))
        return {"cost": cost, "grad_depth": gd, "grad_normal": gn, "grad_pose": gp[:d.n_sources],
                "loss": float(loss[0])}

    def refine_step(self, solves, step_depth: float, step_normal: float):
        """One gradient step on the planes of every (ref, options), kept per pixel where it lowers
        the cost; returns the relative-pose gradients (n, 8, 12) at the planes before the step and
        the losses before and after it (n,) each.  Steps of 0 only evaluate."""
        descs = [self._desc(r, o)[0] for r, o in solves]
        arr = (N.SolveDesc * len(descs))(*descs)
        gp = np.zeros((len(descs), N.MAX_SOURCES, 12), dtype=np.float64)
        before = np.zeros(len(descs), dtype=np.float64)
        after = np.zeros(len(descs), dtype=np.float64)
        N.check(self._lib.lr3d_ncc_refine_step(self.handle, arr, len(descs), float(step_depth),
                                               float(step_normal), N.ptr(gp), N.ptr(before), N.ptr(after)))
        return gp, before, after

    def set_coarse(self, coarse: Optional["HipDenseStereo"]) -> None:
        N.check(self._lib.lr3d_stereo_set_coarse(self.handle, coarse.handle if coarse else None))

    def destroy(self) -> None:
        """DS:357-370"""
        if self.handle:
            self._lib.lr3d_stereo_destroy(self.handle)
            self.handle = None
