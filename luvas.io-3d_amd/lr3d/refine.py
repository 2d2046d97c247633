"""Gradient refinement of PatchMatch planes and camera poses with the NCC adjoint.

The reference has no refinement of this kind; this module drives lr3d_ncc_refine_step (planes,
on the GPU) and turns the relative-pose gradients it returns into gradients of the world -> camera
poses (6-DoF per view), which are updated on the host.

Pose perturbation (left, in the camera frame): R' = exp([w]x) R, t' = exp([w]x) t + tau.
For the relative pose of reference r and source s (R = R_s R_r^T, t = t_s - R t_r,
frontend/src/app/vision/geometry/camera.ts:73-77), first order:
  source:    dR = [w]x R,  dt = [w]x t + tau   ->  dL/dw = sum_k R[:,k] x G[:,k] + t x g,  dL/dtau = g
  reference: dR = -R [w]x, dt = -R tau         ->  dL/dw = -vee(B - B^T) with B = R^T G,  dL/dtau = -R^T g
where G = dL/dR and g = dL/dt.
"""
from __future__ import annotations

import numpy as np

from .camera import Pose, relative_pose


def _vee_skew(B):
    return np.array([B[2, 1] - B[1, 2], B[0, 2] - B[2, 0], B[1, 0] - B[0, 1]])


def world_pose_gradients(views, ref: int, sources, grad_rel) -> dict:
    """{layer: (dL/dw, dL/dtau)} from relative-pose gradients grad_rel[s] = (dL/dR (9), dL/dt (3))."""
    out = {}
    for s, layer in enumerate(sources):
        G = np.asarray(grad_rel[s][:9], dtype=np.float64).reshape(3, 3)
        g = np.asarray(grad_rel[s][9:12], dtype=np.float64)
        rel = relative_pose(views[ref].pose, views[layer].pose)
        R, t = rel.R, rel.t
        w_src = sum(np.cross(R[:, k], G[:, k]) for k in range(3)) + np.cross(t, g)
        w_ref = -_vee_skew(R.T @ G)
        for lay, w, tau in ((layer, w_src, g), (ref, w_ref, -R.T @ g)):
            a, b = out.get(lay, (np.zeros(3), np.zeros(3)))
            out[lay] = (a + w, b + tau)
    return out


def exp_so3(w) -> np.ndarray:
    w = np.asarray(w, dtype=np.float64)
    theta = np.linalg.norm(w)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    if theta < 1e-12:
        return np.eye(3) + K
    return np.eye(3) + np.sin(theta) / theta * K + (1 - np.cos(theta)) / theta ** 2 * (K @ K)


def perturb(pose: Pose, w, tau) -> Pose:
    """R' = exp([w]x) R, t' = exp([w]x) t + tau."""
    E = exp_so3(w)
    return Pose(E @ pose.R, E @ pose.t + np.asarray(tau, dtype=np.float64))


def _pose_totals(stereo, solves, grads) -> dict:
    total = {}
    for (ref, o), g in zip(solves, grads):
        srcs = list(o.sources)[:8]
        for lay, (w, tau) in world_pose_gradients(stereo.views, ref, srcs, g[:len(srcs)]).items():
            a, b = total.get(lay, (np.zeros(3), np.zeros(3)))
            total[lay] = (a + w, b + tau)
    return total


def refine(stereo, solves, steps: int = 50, step_depth: float = 2e-4, step_normal: float = 2e-3,
           step_rot: float = 1e-3, step_trans: float = 1e-3, fixed=(0,), resolve: bool = True,
           momentum: float = 0.9, log=None):
    """`steps` joint gradient steps on the planes of every (ref, options) and on the poses of every
    view not in `fixed` (gauge).  The planes must be solved at the current poses beforehand.

    Every step takes one plane step on the GPU (lr3d_ncc_refine_step: a gradient step on each
    pixel's plane, kept where it lowers the pixel's cost, with the pose gradients at the planes
    before it) and one pose step: each view turns by step_rot radians and moves by step_trans
    metres along its negative gradient direction, averaged with the previous steps' directions
    (heavy-ball momentum, reset when a step is rejected).

    resolve=True: after the pose step every view's planes are solved again at the new poses
    (PatchMatch with the solves' options), followed by a plane step, and the pose step is kept if
    that loss is lower (the steps then grow by 1.5); otherwise the poses are restored and the
    planes solved again, which reproduces them (the solve is deterministic), and the steps are
    halved.  Planes fitted at wrong poses hold those poses in place: with the planes fixed, a move
    toward the true poses raises the loss; re-solved planes follow the poses, so the comparison is
    between planes fitted at each.
    resolve=False: the planes are kept; the pose step is kept if the loss of the current planes
    does not rise (steps grow by 1.25), else undone and halved.

    The loss is non-increasing.  Returns the history: the loss before the first step, then the
    loss after each step."""
    history = []
    rot, trans = step_rot, step_trans
    grads, before, after = stereo.refine_step(solves, step_depth, step_normal)
    history.append(float(before.sum()))
    loss = float(after.sum())
    velocity = {}
    for it in range(steps):
        total = _pose_totals(stereo, solves, grads)
        saved = {lay: stereo.views[lay].pose for lay in total if lay not in fixed}
        for lay, (w, tau) in total.items():
            if lay in fixed:
                continue
            vw, vt = velocity.get(lay, (np.zeros(3), np.zeros(3)))
            vw = momentum * vw + w / (np.linalg.norm(w) + 1e-30)
            vt = momentum * vt + tau / (np.linalg.norm(tau) + 1e-30)
            velocity[lay] = (vw, vt)
            dw = -rot * vw / (np.linalg.norm(vw) + 1e-30)
            dt = -trans * vt / (np.linalg.norm(vt) + 1e-30)
            stereo.set_pose(lay, perturb(stereo.views[lay].pose, dw, dt))
        if resolve:
            stereo.solve_batched(solves)
            g2, _, a2 = stereo.refine_step(solves, step_depth, step_normal)
            if float(a2.sum()) < loss:
                loss, grads = float(a2.sum()), g2
                rot, trans = rot * 1.5, trans * 1.5
            else:
                for lay, pose in saved.items():
                    stereo.set_pose(lay, pose)
                stereo.solve_batched(solves)
                grads, _, a3 = stereo.refine_step(solves, step_depth, step_normal)
                loss = min(loss, float(a3.sum()))
                rot, trans = rot * 0.5, trans * 0.5
                velocity = {}
        else:
            _, trial, _ = stereo.refine_step(solves, 0.0, 0.0)
            if float(trial.sum()) <= loss:
                loss = float(trial.sum())
                rot, trans = rot * 1.25, trans * 1.25
            else:
                for lay, pose in saved.items():
                    stereo.set_pose(lay, pose)
                rot, trans = rot * 0.5, trans * 0.5
                velocity = {}
            if it + 1 < steps:
                grads, _, a2 = stereo.refine_step(solves, step_depth, step_normal)
                loss = min(loss, float(a2.sum()))
        history.append(loss)
        if log:
            log(it, loss)
    return history


def _half_frames(views, images):
    """Every view's frame box-filtered 2 x 2 (rounded) and its intrinsics scaled by 0.5 keeping pixel
    centres aligned (camera.ts:28-36); the same poses."""
    from .camera import StereoView, scale_intrinsics
    out_views, frames = [], []
    for v, im in zip(views, images):
        im = np.asarray(im, dtype=np.uint16)
        h, w = im.shape[0] // 2, im.shape[1] // 2
        im = im[:2 * h, :2 * w]
        frames.append(((im[0::2, 0::2] + im[1::2, 0::2] + im[0::2, 1::2] + im[1::2, 1::2] + 2) // 4).astype(np.uint8))
        out_views.append(StereoView(scale_intrinsics(v.intrinsics, w / v.width, h / v.height), v.pose, w, h))
    return out_views, frames


def _stereo_of(ctx, views, frames):
    from .stereo import HipDenseStereo
    st = HipDenseStereo(ctx, views)
    for i, f in enumerate(frames):
        st.set_image(i, f)
    return st


def half_resolution(stereo, images):
    """A half-resolution copy of a stereo set for the coarse stage of refine_pyramid: every view's
    frame box-filtered 2 x 2 (rounded), intrinsics scaled by 0.5 keeping pixel centres aligned
    (camera.ts:28-36), the same poses."""
    return _stereo_of(stereo.ctx, *_half_frames(stereo.views, images))


def pyramid_levels(stereo, images, levels: int = 3):
    """`levels` coarse copies of a stereo set, each half the resolution of the one before (half,
    quarter, eighth, ...; finest first), for refine_pyramid."""
    views, frames, out = stereo.views, images, []
    for _ in range(levels):
        views, frames = _half_frames(views, frames)
        out.append(_stereo_of(stereo.ctx, views, frames))
    return out


def refine_pyramid(stereo, coarse, solves, coarse_steps=(50, 50, 50), steps: int = 25, fixed=(0,), log=None, **kw):
    """Pose refinement coarse to fine.  `coarse`: one coarse copy of the set (half_resolution) or a
    list of them, finest first (pyramid_levels); `coarse_steps`: the steps of refine() on each,
    an int or one per copy.  From the coarsest copy up, each level starts from the poses the level
    below it reached, solves its planes there and runs its steps; then the poses are copied to
    `stereo`, its planes solved and `steps` steps of refine() run on it.  At a level of half the
    resolution a pose error moves the windows half as many pixels, so more of them stay inside the
    NCC basin; measured on the 16 x 960x540 differentiation set, three coarse levels of 50 steps and
    25 full-resolution steps reach about half the rotation error of one half-resolution level of 25
    and 25 full steps.  Returns (coarse histories, coarsest first, fine history); with a single
    coarse copy, (its history, fine history)."""
    single = not isinstance(coarse, (list, tuple))
    levels = [coarse] if single else list(coarse)
    counts = [coarse_steps] * len(levels) if isinstance(coarse_steps, int) else list(coarse_steps)
    if len(counts) != len(levels):
        raise ValueError("refine_pyramid: one coarse_steps entry per coarse level")
    poses = [v.pose for v in stereo.views]
    histories = []
    for level, count in zip(reversed(levels), reversed(counts)):
        for lay in range(len(stereo.views)):
            level.set_pose(lay, poses[lay])
        level.solve_batched(solves)
        histories.append(refine(level, solves, steps=count, fixed=fixed, log=log, **kw))
        poses = [level.views[lay].pose for lay in range(len(stereo.views))]
    for lay in range(len(stereo.views)):
        if lay not in fixed:
            stereo.set_pose(lay, poses[lay])
    stereo.solve_batched(solves)
    hf = refine(stereo, solves, steps=steps, fixed=fixed, log=log, **kw)
    return (histories[0] if single else histories), hf
