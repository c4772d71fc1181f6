"""The Fortran host driver (bio-cfd_amd/fortran/biocfd_run.f90), calling the
C ABI through the iso_c_binding module, reproduces the oracle's time step on
the four parity cases: integers exactly, FP64 fields within 1e-12 relative."""
import os
import subprocess
import tempfile

import pytest

from biocfd_amd import cases as CS
from biocfd_amd.dumps import read_dump
from parity import compare, run_oracle, small_cases

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUN = os.path.join(ROOT, "bio-cfd_amd", "lib", "biocfd_run")
STEPS = (0, 1, 3)


@pytest.mark.parametrize("name", ["channel", "sphere", "nested", "wing", "move", "batched"])
def test_fortran_driver_parity(name):
    assert os.path.exists(RUN), "Fortran driver not built (bio-cfd_amd/build.py)"
    case = small_cases()[name]
    ref = run_oracle(case, STEPS)
    wd = tempfile.mkdtemp(prefix="frun_")
    case.dump_steps = STEPS
    path = os.path.join(wd, "case.bin")
    CS.write_case(case, path)
    print(f"fortran driver: case {name}", flush=True)
    r = subprocess.run([RUN, path, wd], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    bad = []
    for s in STEPS:
        for g in range(len(case.blocks)):
            got = read_dump(os.path.join(wd, f"dump_s{s}_b{g + 1}.bin"))
            assert len(got) > 20
            bad += compare(ref[s][g], got, label=f"{name} step {s} block {g + 1}: ")
    assert not bad, "\n".join(bad[:40])
