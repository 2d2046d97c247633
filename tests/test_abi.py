"""The C ABI: liblr3d.so loads and exports exactly what include/lr3d.h declares, and reports
errors through status codes and lr3d_last_error() (no device needed)."""
import ctypes as C
import os
import re

import pytest

from lr3d import _native as N

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "lr3d.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lr3d_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(N.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = N.load()
    for name in declared_functions():
        assert hasattr(lib, name), name


def _no_device():
    lib = N.load()
    h = C.c_void_p()
    rc = lib.lr3d_ctx_create(0, C.byref(h))
    if rc == N.LR3D_OK:
        lib.lr3d_ctx_destroy(h)
        return False
    return True


def test_null_arguments_are_rejected():
    lib = N.load()
    assert lib.lr3d_ctx_create(0, None) == N.LR3D_ERR_INVALID
    assert b"NULL" in lib.lr3d_last_error()
    assert lib.lr3d_stereo_solve(None, None) == N.LR3D_ERR_INVALID
    assert lib.lr3d_tsdf_read(None, None, None) == N.LR3D_ERR_INVALID
    assert lib.lr3d_ctx_destroy(None) == N.LR3D_OK


def test_context_without_device_fails_loudly():
    if not _no_device():
        pytest.skip("a HIP device is present")
    import lr3d
    with pytest.raises(lr3d.Lr3dError) as err:
        lr3d.Context(0)
    assert err.value.code == N.LR3D_ERR_DEVICE
    assert "no HIP device" in str(err.value)


def _run_with_fault(k, code):
    """Runs `code` in a fresh interpreter with LR3D_FAULT_INJECT=k (the variable is read once per
    process, so each setting needs its own process)."""
    import subprocess
    import sys
    env = dict(os.environ, LR3D_FAULT_INJECT=str(k))
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "luvas.io-3d_amd")
    env["PYTHONPATH"] = pkg + os.pathsep + env.get("PYTHONPATH", "")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    return out.stdout


def test_injected_fault_is_reported_as_device_error():
    # The first checked HIP call of a process is lr3d_ctx_create's device count, with or without a
    # device: the injected failure surfaces as LR3D_ERR_DEVICE with its own message, and only once.
    code = (
        "import ctypes as C\n"
        "from lr3d import _native as N\n"
        "lib = N.load(); h = C.c_void_p()\n"
        "print(lib.lr3d_ctx_create(0, C.byref(h)), lib.lr3d_last_error().decode())\n"
        "rc = lib.lr3d_ctx_create(0, C.byref(h))\n"
        "print(rc, lib.lr3d_last_error().decode() if rc else '')\n"
        "if rc == 0: lib.lr3d_ctx_destroy(h)\n")
    first, second = _run_with_fault(1, code).splitlines()
    rc, msg = first.split(" ", 1)
    assert int(rc) == N.LR3D_ERR_DEVICE
    assert "injected fault (LR3D_FAULT_INJECT=1)" in msg and "hipGetDeviceCount" in msg
    rc, msg = second.split(" ", 1) if " " in second else (second, "")
    assert "injected" not in msg
    assert int(rc) in (N.LR3D_OK, N.LR3D_ERR_DEVICE)


def test_fault_injection_off_by_default():
    code = (
        "import ctypes as C\n"
        "from lr3d import _native as N\n"
        "lib = N.load(); h = C.c_void_p()\n"
        "rc = lib.lr3d_ctx_create(0, C.byref(h))\n"
        "print(rc, lib.lr3d_last_error().decode() if rc else '')\n"
        "if rc == 0: lib.lr3d_ctx_destroy(h)\n")
    for k in (0, -3):
        out = _run_with_fault(k, code)
        assert "injected" not in out
