"""The fused sweep divides |p| by alpha1 and |q| by alpha0 through the host's
reciprocal and one fused correction (csrc/div_rn.hpp) instead of the IEEE
operator.  The check program compares, on the device, 8 numerators for every
divisor significand at eight exponents with the IEEE division bit for bit,
plus zero and tiny numerators.  The fused sweep's own bitwise parity tests
against the two-sweep loop and the oracle cover the same code on the
iteration's values."""
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu


def test_division_by_constant_equals_ieee_division():
    from ambient_image_processor_amd import _build
    print("division check", _build.DIV_RN_CHECK, flush=True)
    r = subprocess.run([_build.DIV_RN_CHECK], capture_output=True, text=True, timeout=120)
    sys.stdout.write(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
