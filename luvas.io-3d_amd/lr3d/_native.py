"""ctypes binding of liblr3d.so (the C ABI declared in include/lr3d.h).

The library is loaded from the package's ``lib/`` directory (built in-tree by
``luvas.io-3d_amd/csrc/Makefile``).  There is no fallback: if the library is missing the import of
the stereo / TSDF classes raises, so nothing silently runs a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "lib")
LIB_PATH = os.environ.get("LR3D_LIBRARY") or os.path.join(LIB_DIR, "liblr3d.so")

LR3D_OK = 0
LR3D_ERR_INVALID = -1
LR3D_ERR_DEVICE = -2
LR3D_ERR_NOMEM = -3
LR3D_ERR_STATE = -4
MAX_SOURCES = 8
STEREO_PRIOR_POOL = 1
NOALLOC_MAX_SOLVES = 1024
MAP_RAW = 0
MAP_FILTERED = 1

# Every symbol include/lr3d.h declares (checked by the CPU test suite).
EXPORTED = (
    "lr3d_last_error", "lr3d_ctx_create", "lr3d_ctx_destroy", "lr3d_ctx_synchronize",
    "lr3d_ctx_set_profiling", "lr3d_ctx_kernel_stats",
    "lr3d_stereo_create", "lr3d_stereo_destroy", "lr3d_stereo_set_image", "lr3d_stereo_solve",
    "lr3d_stereo_solve_batched", "lr3d_stereo_solve_batched_noalloc", "lr3d_stereo_filter", "lr3d_stereo_read_depth",
    "lr3d_stereo_read_filtered", "lr3d_stereo_read_hyp", "lr3d_stereo_write_depth",
    "lr3d_stereo_set_coarse", "lr3d_stereo_set_pose", "lr3d_stereo_write_hyp",
    "lr3d_ncc_backward", "lr3d_ncc_refine_step",
    "lr3d_tsdf_create", "lr3d_tsdf_destroy", "lr3d_tsdf_integrate", "lr3d_tsdf_integrate_host",
    "lr3d_tsdf_integrate_batched", "lr3d_tsdf_read", "lr3d_tsdf_set_counting",
    "lr3d_tsdf_update_count", "lr3d_tsdf_pass_counts", "lr3d_tsdf_backward", "lr3d_stereo_autotune", "lr3d_stereo_get_tile",
    "lr3d_stereo_set_tile", "lr3d_tsdf_autotune", "lr3d_tsdf_get_bricks_per_block",
    "lr3d_tsdf_set_bricks_per_block", "lr3d_tsdf_get_xcd_chunk", "lr3d_tsdf_set_xcd_chunk",
    "lr3d_tsdf_set_depth_culling", "lr3d_tsdf_write", "lr3d_tsdf_extract_surface",
    "lr3d_mesh_remove_small_components", "lr3d_mesh_size", "lr3d_mesh_read", "lr3d_mesh_destroy",
    "lr3d_fuse_depth_maps", "lr3d_cloud_size", "lr3d_cloud_read", "lr3d_cloud_destroy",
)


class Lr3dError(RuntimeError):
    """A non-zero status from the library, with lr3d_last_error()'s message."""

    def __init__(self, code: int, message: str):
        super().__init__(f"lr3d error {code}: {message}")
        self.code = code


class View(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32),
                ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("R", C.c_double * 9), ("t", C.c_double * 3)]


class SolveDesc(C.Structure):
    _fields_ = [("ref", C.c_int32), ("n_sources", C.c_int32), ("sources", C.c_int32 * MAX_SOURCES),
                ("min_depth", C.c_double), ("max_depth", C.c_double),
                ("prior", C.POINTER(C.c_float)), ("geometric", C.c_int32),
                ("iterations", C.c_int32), ("from_coarse", C.c_int32), ("pack", C.c_int32),
                ("topk", C.c_int32)]


class IntegrationView(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32),
                ("fx", C.c_double), ("fy", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
                ("R", C.c_double * 9), ("t", C.c_double * 3), ("max_depth", C.c_double),
                ("prior", C.POINTER(C.c_float)), ("confidence", C.c_double)]


_lib = None


def load() -> C.CDLL:
    """Load liblr3d.so once; raises FileNotFoundError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `make -C luvas.io-3d_amd/csrc` "
            "(or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    i32, i64, u32, dbl = C.c_int32, C.c_int64, C.c_uint32, C.c_double
    sig = {
        "lr3d_last_error": (C.c_char_p, []),
        "lr3d_ctx_create": (i32, [i32, C.POINTER(P)]),
        "lr3d_ctx_destroy": (i32, [P]),
        "lr3d_ctx_synchronize": (i32, [P]),
        "lr3d_ctx_set_profiling": (i32, [P, i32]),
        "lr3d_ctx_kernel_stats": (i32, [P, C.c_char_p, C.POINTER(i64), C.POINTER(dbl)]),
        "lr3d_stereo_create": (i32, [P, C.POINTER(View), i32, u32, C.POINTER(P)]),
        "lr3d_stereo_destroy": (i32, [P]),
        "lr3d_stereo_set_image": (i32, [P, i32, C.c_void_p, i32, i32]),
        "lr3d_stereo_solve": (i32, [P, C.POINTER(SolveDesc)]),
        "lr3d_stereo_solve_batched": (i32, [P, C.POINTER(SolveDesc), i32]),
        "lr3d_stereo_solve_batched_noalloc": (i32, [P, C.POINTER(SolveDesc), i32]),
        "lr3d_stereo_filter": (i32, [P, i32, C.POINTER(i32), i32, i32]),
        "lr3d_stereo_read_depth": (i32, [P, i32, C.c_void_p]),
        "lr3d_stereo_read_filtered": (i32, [P, i32, C.c_void_p]),
        "lr3d_stereo_read_hyp": (i32, [P, i32, C.c_void_p, C.c_void_p]),
        "lr3d_stereo_write_depth": (i32, [P, i32, C.c_void_p]),
        "lr3d_stereo_set_coarse": (i32, [P, P]),
        "lr3d_stereo_set_pose": (i32, [P, i32, C.POINTER(dbl), C.POINTER(dbl)]),
        "lr3d_stereo_write_hyp": (i32, [P, i32, C.c_void_p]),
        "lr3d_ncc_backward": (i32, [P, C.POINTER(SolveDesc), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p]),
        "lr3d_ncc_refine_step": (i32, [P, C.POINTER(SolveDesc), i32, C.c_float, C.c_float, C.c_void_p,
                                       C.c_void_p, C.c_void_p]),
        "lr3d_tsdf_create": (i32, [P, C.POINTER(dbl), dbl, C.POINTER(i32), C.c_void_p, i32, dbl,
                                   C.POINTER(P)]),
        "lr3d_tsdf_destroy": (i32, [P]),
        "lr3d_stereo_autotune": (i32, [P, C.POINTER(SolveDesc), i32, i32]),
        "lr3d_stereo_get_tile": (i32, [P, C.c_char_p, C.POINTER(i32), C.POINTER(i32)]),
        "lr3d_stereo_set_tile": (i32, [P, C.c_char_p, i32, i32]),
        "lr3d_tsdf_autotune": (i32, [P, P, C.c_void_p, i32, C.c_void_p, i32, i32]),
        "lr3d_tsdf_get_bricks_per_block": (i32, [P, C.POINTER(i32)]),
        "lr3d_tsdf_set_bricks_per_block": (i32, [P, i32]),
        "lr3d_tsdf_get_xcd_chunk": (i32, [P, C.POINTER(i32)]),
        "lr3d_tsdf_set_xcd_chunk": (i32, [P, i32]),
        "lr3d_tsdf_set_depth_culling": (i32, [P, i32]),
        "lr3d_tsdf_write": (i32, [P, C.c_void_p, C.c_void_p]),
        "lr3d_tsdf_extract_surface": (i32, [P, dbl, C.POINTER(P)]),
        "lr3d_mesh_remove_small_components": (i32, [P, dbl, i32, C.POINTER(P)]),
        "lr3d_mesh_size": (i32, [P, C.POINTER(i64), C.POINTER(i64)]),
        "lr3d_mesh_read": (i32, [P, C.c_void_p, C.c_void_p, C.c_void_p]),
        "lr3d_mesh_destroy": (i32, [P]),
        "lr3d_fuse_depth_maps": (i32, [P, C.c_void_p, i32, i32, C.c_void_p, dbl, i64, i32, C.POINTER(P),
                                       C.POINTER(dbl)]),
        "lr3d_cloud_size": (i32, [P, C.POINTER(i64)]),
        "lr3d_cloud_read": (i32, [P, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "lr3d_cloud_destroy": (i32, [P]),
        "lr3d_tsdf_backward": (i32, [P, P, C.c_void_p, i32, C.c_void_p, i32, C.c_void_p, C.c_void_p, C.c_int64,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "lr3d_tsdf_integrate": (i32, [P, P, i32, i32, C.POINTER(IntegrationView)]),
        "lr3d_tsdf_integrate_host": (i32, [P, C.c_void_p, C.POINTER(IntegrationView)]),
        "lr3d_tsdf_integrate_batched": (i32, [P, P, C.POINTER(i32), i32,
                                              C.POINTER(IntegrationView), i32]),
        "lr3d_tsdf_read": (i32, [P, C.c_void_p, C.c_void_p]),
        "lr3d_tsdf_set_counting": (i32, [P, i32]),
        "lr3d_tsdf_update_count": (i32, [P, C.POINTER(i64)]),
        "lr3d_tsdf_pass_counts": (i32, [P, C.POINTER(i64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int) -> None:
    """Raise Lr3dError for a non-zero status."""
    if status != LR3D_OK:
        msg = load().lr3d_last_error()
        raise Lr3dError(status, msg.decode() if msg else "")


def ptr(array) -> C.c_void_p:
    """Data pointer of a C-contiguous numpy array."""
    assert array.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return C.c_void_p(array.ctypes.data)


def float_ptr(array):
    if array is None:
        return C.POINTER(C.c_float)()
    assert array.dtype.name == "float32" and array.flags["C_CONTIGUOUS"]
    return array.ctypes.data_as(C.POINTER(C.c_float))
