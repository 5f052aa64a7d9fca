"""Host check of the fused sweep's division of |p| and |q| by alpha1 and
alpha0 (csrc/div_rn.hpp, restated in tests/compat/div_rn_host_check.c with
libm's fmaf) against double-precision division: 5e6 random and up to 1.5e7
near-midpoint or exact quotients over divisors in [2^-60, 2^36], zero, and
tiny numerators compared with 1."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_division_by_reciprocal_and_correction_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "div_rn_host_check")
    src = os.path.join(HERE, "compat", "div_rn_host_check.c")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", src, "-o", exe, "-lm"], check=True)
    r = subprocess.run([exe, "5000000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "mismatches 0; zero mismatches 0; tiny comparisons mismatches 0" in r.stdout
