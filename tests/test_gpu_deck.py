"""The Fortran host driver runs the reference's input deck on the GPU
(biocfd_run --deck) and reproduces the oracle run of the same deck as the
reference's readers load it (tests/golden/deck_wing.case): integers exactly,
FP64 fields within 1e-12 relative, after the set-up and after the last step."""
import os
import shutil
import subprocess
import tempfile

import pytest

from biocfd_amd.dumps import read_dump
from parity import ORC_RUN, ROOT, compare

pytestmark = pytest.mark.gpu

RUN = os.path.join(ROOT, "bio-cfd_amd", "lib", "biocfd_run")
GOLDEN = os.path.join(ROOT, "tests", "golden")
DECK_DIR = os.path.join(GOLDEN, "deck_wing")
DECK_CASE = os.path.join(GOLDEN, "deck_wing.case")
STEPS = (0, 2)


def test_fortran_driver_runs_the_deck():
    assert os.path.exists(RUN), "Fortran driver not built (bio-cfd_amd/build.py)"
    wd = tempfile.mkdtemp(prefix="gdeck_")
    deck = os.path.join(wd, "deck")
    shutil.copytree(DECK_DIR, deck)
    ref_dir = os.path.join(wd, "oracle")
    got_dir = os.path.join(wd, "gpu")
    os.makedirs(ref_dir)
    os.makedirs(got_dir)
    subprocess.run([ORC_RUN, DECK_CASE, ref_dir], check=True, capture_output=True)
    print("fortran driver: deck_wing", flush=True)
    r = subprocess.run([RUN, "--deck", deck, got_dir], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    bad = []
    for s in STEPS:
        for g in (1, 2):
            ref = read_dump(os.path.join(ref_dir, f"dump_s{s}_b{g}.bin"))
            got = read_dump(os.path.join(got_dir, f"dump_s{s}_b{g}.bin"))
            assert len(got) > 20
            bad += compare(ref, got, label=f"deck step {s} block {g}: ")
    assert not bad, "\n".join(bad[:40])
