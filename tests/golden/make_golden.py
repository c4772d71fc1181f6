"""Regenerate the golden fixtures from the reference's own CPU build.

Needs the reference sources (/root/reference) and the reference driver binary
built from them (make -C oracle ref).  Each fixture holds, for one small case
and a few steps, the arrays the reference driver dumps after that step.  The
case itself is rebuilt from the parameters in GOLDEN_CASES, so the tests can
run the oracle (or the GPU path) on identical inputs.

The input-deck fixture: deck_wing/ is the reference input deck that
biocfd_amd.deck writes for config_deck_wing(), deck_wing.case is that deck
as the reference's own readers (readInput, readBlockInterface,
readSurfaceMeshGmsh) load it, in the case-file layout, and
deck_wing_digests.json holds SHA-256 digests of the arrays the reference
driver dumps when it runs the deck (the arrays themselves are too large to
keep).

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "bio-cfd_amd", "python"))

from biocfd_amd import cases as CS  # noqa: E402
from biocfd_amd import deck as DK  # noqa: E402
from biocfd_amd.dumps import read_dump  # noqa: E402

REF_BIN = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
STEPS = (0, 2)

# arrays kept in the fixtures (everything the parity bar covers)
KEEP = ("u", "v", "w", "p", "ut", "vt", "wt", "pc", "b", "x1", "y1", "z1", "cell", "cell2", "nodeIdTag", "cell_n", "nIterPcor",
        "derr2", "derrStdSt", "fluidCellCount", "solidCellCount", "ibCellCount", "redCellCount",
        "blackCellCount", "TSCellCount", "fluidIndexPtr", "redCellIndexPtr", "blackCellIndexPtr", "Acx", "Acy",
        "Acz", "ca_uu", "ck_uu", "ca_vu", "ck_vu", "ca_uw", "ck_uw", "interceptedIndexPtr", "solidIndexPtr",
        "TSIndexPtr", "index_ts", "nelp", "nelu1", "nelu2", "nelv1", "nelv2", "nelw1", "nelw2", "pNormDis",
        "u1NormDis", "u2NormDis", "v2NormDis", "w1NormDis", "xcent", "cosAlpha", "element_length", "ynode1",
        "u2_ghost", "u2t_ghost", "p_ghost", "v1_ghost", "w2_ghost", "xdot", "ydot", "thetaDot1", "piv_y",
        "viscous_drag", "pressure_drag", "viscous_lift", "pressure_lift")


def golden_cases():
    """name -> Case; small variants of the four case families."""
    out = {}
    out["sphere"] = CS.config_sphere_64(n=16, markers=256, nsteps=max(STEPS), pc_ita_max=12, epsi=0.0)
    out["channel"] = CS.config_channel(n=(12, 10, 8), nsteps=max(STEPS), pc_ita_max=12)
    out["nested"] = CS.config_nested(nc=16, nf=16, markers=128, nsteps=max(STEPS), pc_ita_max=10, epsi=0.0)
    body = CS.BodyPlacement(CS.ellipsoid_body(256, (0.25, 0.07, 0.2), surf_id=51), (0.45, 0.5, 0.5),
                            a0_deg=20.0)
    blk = CS.BlockSpec((0, 0, 0), (1, 1, 1), (16, 16, 16), omega=1.7, body=body, moving=True)
    out["wing"] = CS.Case("wing", [blk], dt=0.002, re=300.0, nsteps=max(STEPS), pc_ita_max=10, epsi=0.0,
                          freq=0.5, phase_angle_deg=90.0)
    out["move"] = CS.config_nested_move(nsteps=max(STEPS))
    return out


def fixture_path(name):
    return os.path.join(HERE, f"{name}.npz")


DECK_DIR = os.path.join(HERE, "deck_wing")
DECK_CASE = os.path.join(HERE, "deck_wing.case")
DECK_DIGESTS = os.path.join(HERE, "deck_wing_digests.json")
DECK_STEPS = (0, 2)   # the reference driver's deck dumps: after the set-up and the last step


def array_digest(a) -> str:
    """SHA-256 of an array's dtype, shape and Fortran-order bytes."""
    a = np.asarray(a)
    h = hashlib.sha256(f"{a.dtype.str}{a.shape}".encode())
    h.update(np.asfortranarray(a).tobytes(order="F"))
    return h.hexdigest()


BATCHED_DIGESTS = os.path.join(HERE, "batched_digests.json")
BATCHED_STEPS = (0, 3)


def batched_case():
    """The uneven multi-block case (a coarse block, a fine block with a
    heaving, pitching plate and three body-free fine blocks)."""
    case = CS.config_batched(nsteps=max(BATCHED_STEPS))
    case.dump_steps = BATCHED_STEPS
    return case


def make_batched_fixture():
    case = batched_case()
    wd = tempfile.mkdtemp(prefix="golden_batched_")
    cpath = os.path.join(wd, "case.bin")
    CS.write_case(case, cpath)
    subprocess.run([REF_BIN, cpath, wd], check=True, capture_output=True, cwd=wd)
    digests = {}
    for s in BATCHED_STEPS:
        for g in range(len(case.blocks)):
            d = read_dump(os.path.join(wd, f"dump_s{s}_b{g + 1}.bin"))
            for k in KEEP:
                if k in d:
                    digests[f"s{s}_b{g + 1}_{k}"] = array_digest(d[k])
    with open(BATCHED_DIGESTS, "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)
    print("batched", len(digests), "digests")


def make_deck_fixture():
    if os.path.isdir(DECK_DIR):
        shutil.rmtree(DECK_DIR)
    case = DK.config_deck_wing()
    DK.write_deck(case, DECK_DIR)
    wd = tempfile.mkdtemp(prefix="golden_deck_")
    run_dir = os.path.join(wd, "deck")
    shutil.copytree(DECK_DIR, run_dir)
    subprocess.run([REF_BIN, "--deck-to-case", DECK_CASE], check=True, capture_output=True, cwd=run_dir)
    out = os.path.join(wd, "out")
    os.makedirs(out)
    subprocess.run([REF_BIN, "--deck", out], check=True, capture_output=True, cwd=run_dir)
    digests = {}
    for s in DECK_STEPS:
        for g in range(len(case.blocks)):
            d = read_dump(os.path.join(out, f"dump_s{s}_b{g + 1}.bin"))
            for k in KEEP:
                if k in d:
                    digests[f"s{s}_b{g + 1}_{k}"] = array_digest(d[k])
    with open(DECK_DIGESTS, "w") as f:
        json.dump(digests, f, indent=1, sort_keys=True)
    print("deck_wing", len(digests), "digests")


def main():
    if not os.path.exists(REF_BIN):
        sys.exit(f"{REF_BIN} missing: run `make -C oracle ref` with the reference present")
    for name, case in golden_cases().items():
        case.dump_steps = STEPS
        wd = tempfile.mkdtemp(prefix=f"golden_{name}_")
        cpath = os.path.join(wd, "case.bin")
        CS.write_case(case, cpath)
        subprocess.run([REF_BIN, cpath, wd], check=True, capture_output=True, cwd=wd)
        arrays = {}
        for s in STEPS:
            for g in range(len(case.blocks)):
                d = read_dump(os.path.join(wd, f"dump_s{s}_b{g + 1}.bin"))
                for k in KEEP:
                    if k in d:
                        arrays[f"s{s}_b{g + 1}_{k}"] = np.asarray(d[k])
        np.savez_compressed(fixture_path(name), **arrays)
        print(name, os.path.getsize(fixture_path(name)), "bytes")
    make_deck_fixture()
    make_batched_fixture()


if __name__ == "__main__":
    main()
