"""CPU pins of the correctly rounded short forms in luvas.io-3d_amd/csrc/fp_exact.h.

The TSDF pass divides with `div_rn` (a reciprocal RN(1/b), then two FMA corrections) and takes
square roots with `sqrt_normal` (the hardware's 1-ulp root, then a residual selection); the
PatchMatch NCC finish uses `sqrt_normal` too.  Both must equal the IEEE operations bit for bit
for the kernels to keep the oracle's bits.  `exact_arith_check.c` runs the same operation
sequences on the CPU (gcc, -ffp-contract=off, C's correctly rounded fmaf, division and sqrtf)
over seeded operands, including mantissas near all-ones, and every square-root seed within 1 ulp.
`rcp_exact` (v_rcp_f32 plus one Newton step) depends on the GPU instruction's own results and is
not pinned here: the same Newton step from an arbitrary 1-ulp seed is not always exact.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not available")
    exe = str(tmp_path_factory.mktemp("exact") / "exact_arith_check")
    subprocess.run([gcc, "-O2", "-std=c11", "-ffp-contract=off", "-o", exe,
                    os.path.join(HERE, "exact_arith_check.c"), "-lm"], check=True)
    return exe


@pytest.mark.parametrize("seed", [1, 2024, 88172645463325252])
def test_div_rn_and_sqrt_select_are_correctly_rounded(checker, seed):
    out = subprocess.run([checker, "4000000", str(seed)], check=True, capture_output=True,
                         text=True).stdout.split()
    fields = dict(zip(out[0::2], map(int, out[1::2])))
    assert fields["samples"] == 4000000
    assert fields["quotients"] > 3000000
    assert fields["div_rn_mismatches"] == 0
    assert fields["sqrt_mismatches"] == 0


def test_a_sign_error_in_the_last_residual_is_caught(checker, tmp_path):
    """The check is able to fail: with the last residual's sign flipped the quotient is wrong."""
    src = open(os.path.join(HERE, "exact_arith_check.c")).read()
    planted = src.replace("return fmaf(fmaf(-b, q1, a), y, q1);", "return fmaf(fmaf(b, q1, a), y, q1);")
    assert planted != src
    path = tmp_path / "planted.c"
    path.write_text(planted)
    exe = str(tmp_path / "planted")
    subprocess.run([shutil.which("gcc"), "-O2", "-std=c11", "-ffp-contract=off", "-o", exe, str(path),
                    "-lm"], check=True)
    out = subprocess.run([exe, "1000000", "7"], check=True, capture_output=True, text=True).stdout.split()
    fields = dict(zip(out[0::2], map(int, out[1::2])))
    assert fields["div_rn_mismatches"] > 0
