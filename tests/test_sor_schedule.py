"""The fused red-black SOR tile schedule, run on the host.

tests/sor_tile_host.cpp executes the per-thread phase bodies of the gfx950
kernel (bio-cfd_amd/csrc/sor_fused.hpp) workgroup by workgroup with
AddressSanitizer, for every compiled tile shape, on grids whose sides are not
multiples of the tile, and requires every iteration to equal the in-place
red-black sweep bit for bit (values and the update norm).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")


@pytest.mark.skipif(HIPCC is None, reason="hipcc not available")
def test_fused_sor_schedule_matches_inplace_sweep(tmp_path):
    exe = str(tmp_path / "sor_tile_host")
    cmd = [HIPCC, "-x", "hip", "--cuda-host-only", "-O1", "-g", "-std=c++17", "-ffp-contract=off",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "bio-cfd_amd", "csrc"),
           os.path.join(ROOT, "tests", "sor_tile_host.cpp"), "-o", exe,
           "-Xarch_host", "-fsanitize=address"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-3000:])
    assert "all cases bit-identical" in r.stdout


def test_every_compiled_shape_is_emulated():
    """The host emulation runs every (type, tile, threads, depth, iterations
    per pass) the library instantiates, including the one-iteration kernel of
    every multi-iteration variant."""
    import re

    csrc = os.path.join(ROOT, "bio-cfd_amd", "csrc")
    emu = open(os.path.join(ROOT, "tests", "sor_tile_host.cpp")).read()
    run = set()
    for t in re.findall(r"run_shape<(\w+), (\d+), (\d+), (\d+), (\d+)(?:, (\d+))?>\(\)", emu):
        run.add(("cells",) + t[:5] + (t[5] or "1",))
    for t in re.findall(r"run_pair<(\w+), (\d+), (\d+), (\d+), (\d+), (\d)>\(\)", emu):
        run.add(("pair",) + t)
    lib = set()
    for real, src, macro, pmacro in (("double", "k_pcor.hip", "FV", "PV"), ("float", "k_mixed.hip", "SV", "PS")):
        text = open(os.path.join(csrc, src)).read()
        for ti, tj, nt, d, nit in re.findall(macro + r"\((\d+), (\d+), (\d+), (\d+), \d+, (\d+)\),", text):
            lib.add(("cells", real, ti, tj, nt, d, nit))
            lib.add(("cells", real, ti, tj, nt, d, "1"))
        for ti, tj, nt, nit, creg in re.findall(pmacro + r"\((\d+), (\d+), (\d+), \d+, (\d+), (\d)\),", text):
            lib.add(("pair", real, ti, tj, nt, nit, creg))
            lib.add(("pair", real, ti, tj, nt, "1", creg))
    assert len(lib) > 15 and lib <= run, sorted(lib - run)
