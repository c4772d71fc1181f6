"""The momentum kernel's division by a fixed divisor from its correctly
rounded reciprocal (bio-cfd_amd/csrc/rdiv.hpp), run on the host against the
division: tests/rdiv_host.cpp draws random numerators and divisors over the
range the kernel uses it in, quotients next to rounding midpoints, divisors
with significands of all ones or nearly 1, zeros of both signs and the
kernel's own kind of denominators, and requires every quotient to equal
x / b bit for bit."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")


@pytest.mark.skipif(HIPCC is None, reason="hipcc not available")
def test_corrected_reciprocal_division_equals_division(tmp_path):
    exe = str(tmp_path / "rdiv_host")
    cmd = [HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "bio-cfd_amd", "csrc"), os.path.join(ROOT, "tests", "rdiv_host.cpp"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe, "20000000"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "every quotient equals the division" in r.stdout
