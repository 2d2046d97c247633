"""Device context: the runtime the reference gets from requestGpu()
(frontend/src/app/vision/gpu/gpu-device.ts:29-61), one HIP device and stream."""
from __future__ import annotations

import ctypes as C

from . import _native as N


class Context:
    """An lr3d context on one HIP device.  Errors raise Lr3dError (the reference surfaces device
    errors as exceptions at stage end, stage-runner.ts:40-50)."""

    def __init__(self, device: int = 0):
        self._lib = N.load()
        handle = C.c_void_p()
        N.check(self._lib.lr3d_ctx_create(device, C.byref(handle)))
        self.handle = handle
        self.device = device

    def synchronize(self) -> None:
        N.check(self._lib.lr3d_ctx_synchronize(self.handle))

    def set_profiling(self, enable: bool) -> None:
        N.check(self._lib.lr3d_ctx_set_profiling(self.handle, 1 if enable else 0))

    def kernel_stats(self, name: str) -> tuple[int, float]:
        """(launches, total milliseconds) of a kernel since profiling was enabled."""
        launches = C.c_int64()
        total = C.c_double()
        N.check(self._lib.lr3d_ctx_kernel_stats(self.handle, name.encode(), C.byref(launches),
                                                C.byref(total)))
        return launches.value, total.value

    def destroy(self) -> None:
        if self.handle:
            self._lib.lr3d_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()
