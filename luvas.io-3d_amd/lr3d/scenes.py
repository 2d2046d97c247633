"""Seeded synthetic workloads with analytic ground truth (the scene specification of every
benchmark and test configuration).

Scenes are rendered by libscenegen (csrc/scenegen.c): u8 gray frames, ground-truth camera-z depth,
and exact signed distances.  Seeds drive mulberry32 (frontend/src/app/vision/math/dense.ts:92-101).

Scene choices that the workload descriptions leave open, fixed here:
  * sphere scene: a sphere of radius 0.5 m (unit diameter, so that it fits the 1.024 m TSDF cubes
    the same workloads fix) centred at the origin, resting on a 2 x 2 m ground rectangle at
    y = -0.5; cameras look at the origin from an orbit at 0.4 m height, world +y up;
  * the TSDF cubes are centred on the origin;
  * room: the inside of a 4 x 2.5 x 3 m box (x, y, z), floor at y = 0, with 12 seeded
    axis-aligned boxes (0.3-0.8 m wide, 0.3-1.0 m tall) standing on the floor; 64 cameras on a
    seeded loop (radii 1.1 x 0.8 m around the room centre, 1.2-1.8 m height) look inward across
    the room, 10-25 degrees down and up to 20 degrees off the radial direction, with a horizontal field of view
    of 70 degrees; resolutions cycle 640x480, 960x540, 1280x720, 1920x1080 along the loop, so
    neighbouring views differ in size; neighbour counts are drawn from 4-10 per view and sliced
    to 8 (DS:186); depth 0.3-6 m; the 384^3 x 1 cm TSDF cube is centred on the room centre (the
    two x walls, 2 m from the centre, fall 8 cm outside it).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from .camera import Intrinsics, Pose, StereoView, look_at, scale_intrinsics

LIB_PATH = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib",
                        "libscenegen.so")

SG_SPHERE, SG_BOX, SG_RECT_Y, SG_ROOM = 0, 1, 2, 3


class _Prim(C.Structure):
    _fields_ = [("kind", C.c_int32), ("a", C.c_double * 3), ("b", C.c_double * 3), ("r", C.c_double)]


class _Scene(C.Structure):
    _fields_ = [("n_prims", C.c_int32), ("prims", _Prim * 64), ("noise_offset", C.c_double * 3),
                ("freq", C.c_double), ("sky", C.c_double)]


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: run `make -C luvas.io-3d_amd/csrc`")
        lib = C.CDLL(LIB_PATH)
        lib.sg_render.restype = C.c_int32
        lib.sg_render.argtypes = [C.POINTER(_Scene), C.c_int32, C.c_int32, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.c_void_p, C.c_void_p]
        lib.sg_signed_distance.restype = C.c_int32
        lib.sg_signed_distance.argtypes = [C.POINTER(_Scene), C.c_void_p, C.c_int64, C.c_void_p]
        _lib = lib
    return _lib


def mulberry32(seed: int):
    """frontend/src/app/vision/math/dense.ts:92-101: deterministic PRNG returning [0, 1)."""
    state = seed & 0xFFFFFFFF

    def imul(a, b):
        return (a * b) & 0xFFFFFFFF

    def next_value() -> float:
        nonlocal state
        state = (state + 0x6D2B79F5) & 0xFFFFFFFF
        t = state
        t = imul(t ^ (t >> 15), t | 1)
        t ^= (t + imul(t ^ (t >> 7), t | 61)) & 0xFFFFFFFF
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0

    return next_value


@dataclass
class Primitive:
    kind: int
    a: tuple = (0.0, 0.0, 0.0)
    b: tuple = (0.0, 0.0, 0.0)
    r: float = 0.0


@dataclass
class SceneSpec:
    prims: List[Primitive]
    seed: int
    freq: float = 1.0
    sky: float = 200.0

    def _struct(self) -> _Scene:
        s = _Scene()
        s.n_prims = len(self.prims)
        for i, p in enumerate(self.prims):
            s.prims[i].kind = p.kind
            s.prims[i].a[:] = list(p.a)
            s.prims[i].b[:] = list(p.b)
            s.prims[i].r = p.r
        rnd = mulberry32(self.seed)
        s.noise_offset[:] = [rnd() * 100.0, rnd() * 100.0, rnd() * 100.0]
        s.freq = self.freq
        s.sky = self.sky
        return s

    def render(self, view: StereoView, gray: bool = True, depth: bool = True):
        lib = _load()
        s = self._struct()
        k = view.intrinsics
        R = (C.c_double * 9)(*np.asarray(view.pose.R, dtype=np.float64).reshape(9))
        t = (C.c_double * 3)(*np.asarray(view.pose.t, dtype=np.float64).reshape(3))
        g = np.empty(view.height * view.width, dtype=np.uint8) if gray else None
        d = np.empty(view.height * view.width, dtype=np.float32) if depth else None
        lib.sg_render(C.byref(s), view.width, view.height, k.fx, k.fy, k.cx, k.cy, R, t,
                      C.c_void_p(g.ctypes.data) if g is not None else None,
                      C.c_void_p(d.ctypes.data) if d is not None else None)
        return g, d

    def signed_distance(self, points: np.ndarray) -> np.ndarray:
        lib = _load()
        s = self._struct()
        p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
        out = np.empty(len(p), dtype=np.float64)
        lib.sg_signed_distance(C.byref(s), C.c_void_p(p.ctypes.data), len(p), C.c_void_p(out.ctypes.data))
        return out


@dataclass
class TsdfSpec:
    origin: tuple
    voxel: float
    voxels: int           # cube side in voxels
    truncation_voxels: float


@dataclass
class Workload:
    """One benchmark / test configuration: cameras, frames, truth, PatchMatch and TSDF settings."""
    name: str
    scene: SceneSpec
    views: List[StereoView]
    sources: List[List[int]]
    ranges: List[tuple]
    iterations: int
    topk: int
    tsdf: TsdfSpec
    images: List[np.ndarray] = field(default_factory=list)
    truth: List[np.ndarray] = field(default_factory=list)

    def render(self, layers: Optional[List[int]] = None) -> "Workload":
        """Render frames and ground-truth depth (all layers, or the given ones)."""
        layers = range(len(self.views)) if layers is None else layers
        if not self.images:
            self.images = [None] * len(self.views)
            self.truth = [None] * len(self.views)
        for i in layers:
            g, d = self.scene.render(self.views[i])
            self.images[i] = g.reshape(self.views[i].height, self.views[i].width)
            self.truth[i] = d.reshape(self.views[i].height, self.views[i].width)
        return self

    @property
    def megapixels(self) -> float:
        return sum(v.width * v.height for v in self.views) / 1e6

    def mesh_needed(self, ref: int, agreement: float = 0.25) -> int:
        """Views of the consistency check that must agree before a depth is fused: the
        reference's agreeing() (frontend/src/app/vision/pipeline/stages/stereo-plan.ts:99-101)
        with the High profile's mesh agreement 0.25 (core/models/quality.ts:133)."""
        count = min(len(self.sources[ref]), 8)
        return min(count, max(2, int(math.floor(agreement * count + 0.5))))

    def hypotheses_per_pixel(self) -> int:
        """Nominal hypotheses per pixel: 1 (init) + 14 per iteration (8 propagated + 6 refined)."""
        return 1 + 14 * self.iterations


def sphere_scene(seed: int, freq: float = 1.0) -> SceneSpec:
    """Sphere on a ground rectangle.  `freq` scales the noise frequencies so that the finest
    octave stays above ~2 pixels at the workload's resolution (no aliasing across views)."""
    return SceneSpec([Primitive(SG_SPHERE, (0.0, 0.0, 0.0), r=0.5),
                      Primitive(SG_RECT_Y, (-1.0, -0.5, -1.0), (1.0, 0.0, 1.0))], seed=seed,
                     freq=freq)


def _ring_sources(n_views: int, n_sources: int) -> List[List[int]]:
    """Nearest views on a ring: +1, -1, +2, -2, ..."""
    out = []
    for i in range(n_views):
        src = []
        step = 1
        while len(src) < n_sources:
            for s in (step, -step):
                if len(src) < n_sources:
                    src.append((i + s) % n_views)
            step += 1
        out.append(src)
    return out


def sphere_small(seed: int = 1) -> Workload:
    """Plain method, small, CPU-runnable: 8 views 320x240 on a 1.5 m circle at 0.4 m height,
    f = 288, depth 0.5-3 m; 4 neighbours, top-2, 4 iterations; TSDF 128^3 x 8 mm, truncation
    3 voxels."""
    n, W, H = 8, 320, 240
    k = Intrinsics(288.0, 288.0, (W - 1) / 2, (H - 1) / 2)
    views = []
    for i in range(n):
        a = 2 * math.pi * i / n
        pose = look_at((1.5 * math.sin(a), 0.4, 1.5 * math.cos(a)))
        views.append(StereoView(k, pose, W, H))
    return Workload("sphere-8x320x240", sphere_scene(seed, freq=0.35), views, _ring_sources(n, 4),
                    [(0.5, 3.0)] * n, iterations=4, topk=2,
                    tsdf=TsdfSpec((-0.512, -0.512, -0.512), 0.008, 128, 3.0))


def sphere_phone(seed: int = 2, n: int = 32) -> Workload:
    """Plain method at phone scale: 32 views 1280x720 on a seeded jittered orbit (radius
    1.2-1.8 m), f = 1000, depth 0.3-4 m; 8 neighbours, top-3, 8 iterations; TSDF 256^3 x 4 mm,
    truncation 4 voxels."""
    W, H = 1280, 720
    k = Intrinsics(1000.0, 1000.0, (W - 1) / 2, (H - 1) / 2)
    rnd = mulberry32(seed * 7919 + 13)
    views = []
    for i in range(n):
        radius = 1.2 + 0.6 * rnd()
        a = 2 * math.pi * (i + 0.6 * (rnd() - 0.5)) / n
        height = 0.4 + 0.2 * (rnd() - 0.5)
        pose = look_at((radius * math.sin(a), height, radius * math.cos(a)))
        views.append(StereoView(k, pose, W, H))
    return Workload(f"sphere-{n}x1280x720", sphere_scene(seed), views, _ring_sources(n, 8),
                    [(0.3, 4.0)] * n, iterations=8, topk=3,
                    tsdf=TsdfSpec((-0.512, -0.512, -0.512), 0.004, 256, 4.0))


def sphere_diff(seed: int = 4, n: int = 16) -> Workload:
    """Differentiation of the NCC cost: 16 views 960x540 of the sphere scene on a seeded jittered
    orbit (radius 1.2-1.8 m), f = 750 (the phone field of view), depth 0.3-4 m; 4 neighbours
    (+-45 degrees on the ring), top-2, 6 iterations; TSDF 256^3 x 4 mm, truncation 4 voxels.
    Poses as rendered; `noisy_views` perturbs them."""
    W, H = 960, 540
    k = Intrinsics(750.0, 750.0, (W - 1) / 2, (H - 1) / 2)
    rnd = mulberry32(seed * 7919 + 41)
    views = []
    for i in range(n):
        radius = 1.2 + 0.6 * rnd()
        a = 2 * math.pi * (i + 0.6 * (rnd() - 0.5)) / n
        height = 0.4 + 0.2 * (rnd() - 0.5)
        pose = look_at((radius * math.sin(a), height, radius * math.cos(a)))
        views.append(StereoView(k, pose, W, H))
    return Workload(f"sphere-diff-{n}x960x540", sphere_scene(seed, freq=0.75), views, _ring_sources(n, 4),
                    [(0.3, 4.0)] * n, iterations=6, topk=2,
                    tsdf=TsdfSpec((-0.512, -0.512, -0.512), 0.004, 256, 4.0))


def noisy_views(views: List[StereoView], seed: int = 4, rotation_deg: float = 0.5,
                translation: float = 0.005, fixed: int = 0) -> List[StereoView]:
    """Poses perturbed by seeded noise: rotation about a random axis with per-axis sigma
    rotation_deg / sqrt(3) (RMS angle rotation_deg), translation with per-axis sigma
    translation / sqrt(3) (RMS translation); view `fixed` is left exact (the gauge)."""
    from .refine import perturb
    rng = np.random.default_rng(seed)
    out = []
    for i, v in enumerate(views):
        w = rng.normal(0, math.radians(rotation_deg) / math.sqrt(3), 3)
        tau = rng.normal(0, translation / math.sqrt(3), 3)
        out.append(v if i == fixed else StereoView(v.intrinsics, perturb(v.pose, w, tau), v.width, v.height))
    return out


def room_scene(seed: int, freq: float = 0.3) -> SceneSpec:
    """The room and its 12 boxes (mulberry32 seeded)."""
    rnd = mulberry32(seed * 104729 + 7)
    prims = [Primitive(SG_ROOM, (-2.0, 0.0, -1.5), (2.0, 2.5, 1.5))]
    for _ in range(12):
        sx, sz = 0.3 + 0.5 * rnd(), 0.3 + 0.5 * rnd()
        h = 0.3 + 0.7 * rnd()
        x = -2.0 + sx / 2 + (4.0 - sx) * rnd()
        z = -1.5 + sz / 2 + (3.0 - sz) * rnd()
        prims.append(Primitive(SG_BOX, (x - sx / 2, 0.0, z - sz / 2), (x + sx / 2, h, z + sz / 2)))
    return SceneSpec(prims, seed=seed, freq=freq)


ROOM_SIZES = [(640, 480), (960, 540), (1280, 720), (1920, 1080)]


def room_batched(seed: int = 3, n: int = 64, scale: float = 1.0, iterations: int = 8) -> Workload:
    """Batched configuration: 64 views of the procedural room at mixed resolutions (16 each at
    640x480, 960x540, 1280x720, 1920x1080), neighbour counts drawn from 4-10 (sliced to 8),
    8 iterations, top-k by the reference rule; TSDF 384^3 x 1 cm, truncation 4 voxels.
    `scale` shrinks every image (intrinsics scaled alike) for CPU-checkable variants."""
    rnd = mulberry32(seed * 7919 + 29)
    views = []
    counts = []
    for i in range(n):
        W0, H0 = ROOM_SIZES[i % 4]
        W, H = max(16, int(round(W0 * scale))), max(16, int(round(H0 * scale)))
        f = (W / 2) / math.tan(math.radians(35.0))
        k = Intrinsics(f, f, (W - 1) / 2, (H - 1) / 2)
        a = 2 * math.pi * (i + 0.4 * (rnd() - 0.5)) / n
        height = 1.2 + 0.6 * rnd()
        c = (1.1 * math.cos(a), height, 0.8 * math.sin(a))
        yaw = a + math.pi + math.radians(20.0) * (2 * rnd() - 1)
        pitch = math.radians(10.0 + 15.0 * rnd())
        d = (math.cos(yaw) * math.cos(pitch), -math.sin(pitch), math.sin(yaw) * math.cos(pitch))
        pose = look_at(c, (c[0] + d[0], c[1] + d[1], c[2] + d[2]))
        views.append(StereoView(k, pose, W, H))
        counts.append(4 + int(7 * rnd()))
    sources = [ring[:cnt] for ring, cnt in zip(_ring_sources(n, 10), counts)]
    name = f"room-{n}-mixed" + ("" if scale == 1.0 else f"-x{scale:g}")
    return Workload(name, room_scene(seed), views, sources, [(0.3, 6.0)] * n,
                    iterations=iterations, topk=0,
                    tsdf=TsdfSpec((-1.92, 1.25 - 1.92, -1.92), 0.01, 384, 4.0))


def room_large(seed: int = 5, n: int = 200, iterations: int = 8) -> Workload:
    """Room-scale forward + TSDF adjoint: 200 views 1280x720 of the procedural room (the batched
    configuration's trajectory and neighbour rule), 8 iterations, top-k by the reference rule;
    TSDF 512^3 x 8 mm (1.07 GB of tsdf + weight), truncation 4 voxels, centred on the room."""
    rnd = mulberry32(seed * 7919 + 29)
    W, H = 1280, 720
    f = (W / 2) / math.tan(math.radians(35.0))
    k = Intrinsics(f, f, (W - 1) / 2, (H - 1) / 2)
    views, counts = [], []
    for i in range(n):
        a = 2 * math.pi * (i + 0.4 * (rnd() - 0.5)) / n
        height = 1.2 + 0.6 * rnd()
        c = (1.1 * math.cos(a), height, 0.8 * math.sin(a))
        yaw = a + math.pi + math.radians(20.0) * (2 * rnd() - 1)
        pitch = math.radians(10.0 + 15.0 * rnd())
        d = (math.cos(yaw) * math.cos(pitch), -math.sin(pitch), math.sin(yaw) * math.cos(pitch))
        views.append(StereoView(k, look_at(c, (c[0] + d[0], c[1] + d[1], c[2] + d[2])), W, H))
        counts.append(4 + int(7 * rnd()))
    sources = [ring[:cnt] for ring, cnt in zip(_ring_sources(n, 10), counts)]
    return Workload(f"room-{n}x1280x720", room_scene(seed), views, sources, [(0.3, 6.0)] * n,
                    iterations=iterations, topk=0,
                    tsdf=TsdfSpec((-2.048, 1.25 - 2.048, -2.048), 0.008, 512, 4.0))


def table_scene(seed: int = 6) -> SceneSpec:
    """A textured box (0.14 x 0.12 x 0.10 m) and a sphere (radius 5 cm) on a 1 m table top, at the
    reference's texture frequencies (frontend/tools/synthetic-scene.ts:8-13 describes the same
    layout)."""
    return SceneSpec([Primitive(SG_BOX, (-0.07, 0.0, -0.05), (0.07, 0.12, 0.05)),
                      Primitive(SG_SPHERE, (0.13, 0.05, 0.07), r=0.05),
                      Primitive(SG_RECT_Y, (-0.5, 0.0, -0.5), (0.5, 0.0, 0.5))], seed=seed, freq=1.0)


def reference_check(seed: int = 6) -> Workload:
    """The reference's PatchMatch known-answer configuration (frontend/tools/verify/
    stereo-check.ts:11-49): 5 views at 640x360 (f = 500, the 1280x720 f = 1000 camera scaled),
    10 degrees apart on a 0.42 m circle at 0.2 m height looking at (0.03, 0.05, 0.02); view 2 is
    solved from views 0, 1, 3, 4 over depths 0.1-2 m with 5 iterations, top-k by the reference
    rule.  The other views take their ring neighbours."""
    W, H = 640, 360
    k = scale_intrinsics(Intrinsics(1000.0, 1000.0, 639.5, 359.5), W / 1280, H / 720)
    views = []
    for step in (-2, -1, 0, 1, 2):
        a = step / 36 * 2 * math.pi
        views.append(StereoView(k, look_at((0.42 * math.sin(a), 0.2, 0.42 * math.cos(a)), (0.03, 0.05, 0.02)),
                                W, H))
    sources = [[i for i in range(5) if i != r] for r in range(5)]
    return Workload("reference-check-5x640x360", table_scene(seed), views, sources, [(0.1, 2.0)] * 5,
                    iterations=5, topk=0, tsdf=TsdfSpec((-0.256, -0.05, -0.256), 0.004, 128, 4.0))


def table_rings(seed: int = 6, iterations: int = 6) -> Workload:
    """The reference's verification capture around the table scene (frontend/tools/verify/
    trajectory.ts:5-12): two rings of keyframes looking at (0.03, 0.05, 0.02), 36 at 0.2 m height
    and 0.42 m radius and 24 at 0.36 m and 0.36 m, here at 640x360 (f = 500); each view takes its
    8 nearest cameras as sources; depth 0.1-1.2 m; TSDF 256^3 x 2 mm over the table centre,
    truncation 4 voxels."""
    W, H = 640, 360
    k = scale_intrinsics(Intrinsics(1000.0, 1000.0, 639.5, 359.5), W / 1280, H / 720)
    views = []
    for height, radius, count in ((0.2, 0.42, 36), (0.36, 0.36, 24)):
        for i in range(count):
            a = 2 * math.pi * i / count
            views.append(StereoView(k, look_at((radius * math.sin(a), height, radius * math.cos(a)),
                                               (0.03, 0.05, 0.02)), W, H))
    centres = np.array([-(v.pose.R.T @ v.pose.t) for v in views])
    sources = []
    for r in range(len(views)):
        d = np.linalg.norm(centres - centres[r], axis=1)
        d[r] = np.inf
        sources.append([int(j) for j in np.argsort(d, kind="stable")[:8]])
    return Workload("table-rings-60x640x360", table_scene(seed), views, sources, [(0.1, 1.2)] * len(views),
                    iterations=iterations, topk=0, tsdf=TsdfSpec((-0.256, -0.05, -0.256), 0.002, 256, 4.0))


WORKLOADS = {
    "sphere-small": sphere_small,
    "sphere-phone": sphere_phone,
    "room-batched": room_batched,
    "sphere-diff": sphere_diff,
    "room-large": room_large,
    "reference-check": reference_check,
    "table-rings": table_rings,
}
