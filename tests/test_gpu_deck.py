"""The Fortran host driver runs the reference's input deck on the GPU
(biocfd_run --deck) and reproduces the oracle run of the same deck as read
(tests/golden/deck_wing.case, whose bytes the reference's own readers give,
or biocfd_run --deck-to-case): integers exactly, FP64 fields within 1e-12
relative, after the set-up and after the last step.  The second deck is the
moving case: its fine block moves in both steps."""
import os
import shutil
import subprocess
import tempfile

import pytest

from biocfd_amd import cases as CS
from biocfd_amd import deck as DK
from biocfd_amd.dumps import read_dump
from parity import ORC_RUN, ROOT, compare

pytestmark = pytest.mark.gpu

RUN = os.path.join(ROOT, "bio-cfd_amd", "lib", "biocfd_run")
GOLDEN = os.path.join(ROOT, "tests", "golden")
STEPS = (0, 2)


@pytest.mark.parametrize("name", ["deck_wing", "move"])
def test_fortran_driver_runs_the_deck(name):
    assert os.path.exists(RUN), "Fortran driver not built (bio-cfd_amd/build.py)"
    wd = tempfile.mkdtemp(prefix="gdeck_")
    deck = os.path.join(wd, "deck")
    if name == "deck_wing":
        shutil.copytree(os.path.join(GOLDEN, "deck_wing"), deck)
        case_file = os.path.join(GOLDEN, "deck_wing.case")
    else:
        DK.write_deck(CS.config_nested_move(), deck)
        case_file = os.path.join(wd, "deck.case")
        r = subprocess.run([RUN, "--deck-to-case", deck, case_file], capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    ref_dir = os.path.join(wd, "oracle")
    got_dir = os.path.join(wd, "gpu")
    os.makedirs(ref_dir)
    os.makedirs(got_dir)
    subprocess.run([ORC_RUN, case_file, ref_dir], check=True, capture_output=True)
    print(f"fortran driver: deck {name}", flush=True)
    r = subprocess.run([RUN, "--deck", deck, got_dir], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    bad = []
    for s in STEPS:
        for g in (1, 2):
            ref = read_dump(os.path.join(ref_dir, f"dump_s{s}_b{g}.bin"))
            got = read_dump(os.path.join(got_dir, f"dump_s{s}_b{g}.bin"))
            assert len(got) > 20
            bad += compare(ref, got, label=f"{name} step {s} block {g}: ")
    if name == "move":
        moved = read_dump(os.path.join(got_dir, "dump_s2_b2.bin"))["x1"]
        start = read_dump(os.path.join(got_dir, "dump_s0_b2.bin"))["x1"]
        assert moved[0] < start[0]          # the fine block moved in -x
    assert not bad, "\n".join(bad[:40])
