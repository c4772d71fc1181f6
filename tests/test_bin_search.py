"""The binned nearest-element search (bio-cfd_amd/csrc/bin_search.hpp), run
on the host.

tests/bin_search_host.cpp builds the centroid bins as the GPU does, runs the
64 lane bodies of a wavefront per query with the shell-by-shell reduction,
and requires every nearest element to equal the reference's scan over all
elements (first minimum, strict <): sphere meshes, a flat plate, duplicated
centroids (exact ties), one element, queries near, inside and far outside
the body, under AddressSanitizer.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")


@pytest.mark.skipif(HIPCC is None, reason="hipcc not available")
def test_binned_search_equals_reference_scan(tmp_path):
    exe = str(tmp_path / "bin_search_host")
    cmd = [HIPCC, "-x", "hip", "--cuda-host-only", "-O1", "-g", "-std=c++17", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "bio-cfd_amd", "csrc"),
           os.path.join(ROOT, "tests", "bin_search_host.cpp"), "-o", exe,
           "-Xarch_host", "-fsanitize=address"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "all queries equal the reference scan" in r.stdout
