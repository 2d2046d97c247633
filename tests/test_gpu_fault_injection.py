"""Error paths on the device: with LR3D_FAULT_INJECT=k the k-th checked HIP call of a process
reports a failure.  Each case runs the sphere pipeline (context, stereo, two solves, filter, TSDF
pass) in a fresh process until the injected failure surfaces as an Lr3dError with code
LR3D_ERR_DEVICE, keeps the failed objects aside, then runs the whole pipeline again on new objects
in the same process: its maps and volume must equal those of a process with no injection.

The failure is injected after the real call has returned and the device has drained, so no kernel
is cut short; what is exercised is the library's own unwinding (partially built contexts, streams,
buffers and plans) and that a later context is unaffected by it.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PIPELINE = r'''
import hashlib, sys
import numpy as np
import lr3d
from lr3d import HipDenseStereo, HipTsdfVolume, IntegrationView, SolveOptions, dense_grid, scenes

work = scenes.sphere_small().render()
keep = []

def run():
    ctx = lr3d.Context(0); keep.append(ctx)
    st = HipDenseStereo(ctx, work.views); keep.append(st)
    for i, im in enumerate(work.images):
        st.set_image(i, im)
    for ref in (0, 1):
        lo, hi = work.ranges[ref]
        st.solve(ref, SolveOptions(sources=work.sources[ref], min_depth=lo, max_depth=hi,
                                   iterations=work.iterations, topk=work.topk))
    st.filter(0, [1], 1)
    grid, layout = dense_grid(work.tsdf.origin, work.tsdf.voxel, work.tsdf.voxels)
    vol = HipTsdfVolume(ctx, grid, layout, work.tsdf.truncation_voxels); keep.append(vol)
    views = [IntegrationView(v.intrinsics, v.pose, work.ranges[i][1], v.width, v.height)
             for i, v in enumerate(work.views[:2])]
    vol.integrate_batched(st, [0, 1], views)
    h = hashlib.sha256()
    for a in (st.depth_map(0), st.depth_map(1), st.read_filtered(0), *vol.read()):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()

fault = "none"
try:
    digest = run()
except lr3d.Lr3dError as e:
    fault = f"{e.code} {e}"
print("FAULT", fault.replace("\n", " "), flush=True)
if fault == "none":  # the pipeline made fewer than k checked calls: nothing to recover from
    print("HASH", digest, flush=True)
    sys.exit(0)
for obj in reversed(keep):  # the failed objects are destroyed before the clean run
    obj.destroy()
keep.clear()
print("HASH", run(), flush=True)
'''


def run_case(k):
    env = dict(os.environ, LR3D_FAULT_INJECT=str(k))
    env["PYTHONPATH"] = os.path.join(ROOT, "luvas.io-3d_amd") + os.pathsep + env.get("PYTHONPATH", "")
    print(f"case: LR3D_FAULT_INJECT={k}", flush=True)
    out = subprocess.run([sys.executable, "-u", "-c", PIPELINE], env=env, capture_output=True, text=True,
                         timeout=240, cwd=ROOT)
    assert out.returncode == 0, out.stderr[-2000:]
    lines = dict(line.split(" ", 1) for line in out.stdout.splitlines() if line[:5] in ("FAULT", "HASH "))
    print(f"  fault: {lines['FAULT'][:160]}", flush=True)
    return lines["FAULT"], lines["HASH"]


def test_injected_faults_unwind_and_later_work_is_unaffected():
    from lr3d import _native as N
    fault, clean = run_case(0)
    assert fault == "none"
    landed = 0
    for k in (1, 3, 8, 20, 32, 45, 60, 75):
        fault, digest = run_case(k)
        assert digest == clean
        if fault == "none":
            continue  # the pipeline makes fewer than k checked calls
        landed += 1
        code, msg = fault.split(" ", 1)
        assert int(code) == N.LR3D_ERR_DEVICE
        assert f"injected fault (LR3D_FAULT_INJECT={k})" in msg
        assert digest == clean, f"results after the fault at call {k} differ from a clean run"
    assert landed >= 6
