"""The oracle's own pins, run again against an AddressSanitizer + UndefinedBehaviorSanitizer build
of the same C sources (`make -C oracle asan`): any out-of-bounds access, use after free, signed
overflow, misaligned load or bad shift in the oracle aborts the run with the sanitizer's report.

The sanitized library is loaded into an uninstrumented Python, so the sanitizer runtimes are
preloaded and leak checking (which would report the interpreter's own allocations) is off.  The
two slowest whole-schedule tests are left out to keep the CPU suite short; every other oracle test
file runs in full.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
ASAN_LIB = os.path.join(ORACLE, "liblr3d_oracle_asan.so")


def runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.fixture(scope="module")
def sanitized_env():
    asan, ubsan = runtime("libasan.so"), runtime("libubsan.so")
    if not (asan and ubsan):
        pytest.skip("gcc's sanitizer runtimes are not installed")
    build = subprocess.run(["make", "-C", ORACLE, "-s", "asan"], capture_output=True, text=True, timeout=300)
    assert build.returncode == 0, build.stderr
    env = dict(os.environ, LR3D_ORACLE_LIB=ASAN_LIB, LD_PRELOAD=f"{asan} {ubsan}",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="4")
    env["PYTHONPATH"] = os.pathsep.join([ORACLE, os.path.join(ROOT, "luvas.io-3d_amd"), env.get("PYTHONPATH", "")])
    return env


def test_sanitized_build_is_the_one_loaded(sanitized_env):
    probe = ("import oracle as O\n"
             "O.load()\n"
             "maps = open('/proc/self/maps').read()\n"
             "print('asan-oracle' if 'liblr3d_oracle_asan.so' in maps else 'plain-oracle',\n"
             "      'asan-runtime' if 'libasan' in maps else 'no-runtime')\n")
    out = subprocess.run([sys.executable, "-c", probe], env=sanitized_env, capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.split() == ["asan-oracle", "asan-runtime"]


def test_oracle_pins_pass_under_address_and_ub_sanitizers(sanitized_env):
    files = sorted(f for f in os.listdir(os.path.join(ROOT, "tests"))
                   if f.startswith("test_oracle_") and f.endswith(".py") and f != os.path.basename(__file__))
    skip = "not test_known_answer_check_in_the_reference_configuration and not with_the_geometric_term"
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-k", skip,
           *[os.path.join("tests", f) for f in files]]
    out = subprocess.run(cmd, env=sanitized_env, capture_output=True, text=True, timeout=900, cwd=ROOT)
    report = (out.stdout + out.stderr)[-4000:]
    assert out.returncode == 0, report
    assert "ERROR: AddressSanitizer" not in report and "runtime error:" not in report, report
