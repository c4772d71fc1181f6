"""The drop-in: the reference's own app/main.F90, unchanged, linked with the
replacement hot-path modules (bio-cfd_amd/fortran/dropin/, the reference's
module names and argument lists, work on the GPU) and the reference's
set-up and I/O modules (oracle/_ref/dropin_main, built by
oracle/ref/Makefile).  On the committed wing deck, on the moving deck (its
fine block moves every step) and on a restart from a checkpoint, every file
the reference driver writes (field files after the set-up and at steps 1
and 2, body plots, the SOR iteration log, the per-step body index files, the
tag and cell-count files of the set-up) equals the file the unchanged
reference build (oracle/_ref/ref_main, a CPU program) writes: text and
integers exactly, reals within 1e-12 relative."""
import glob
import os
import shutil
import subprocess
import tempfile

import pytest

from biocfd_amd import cases as CS
from biocfd_amd import deck as DK
from parity import ROOT
from test_gpu_output import same_numbers, set_namelist, tokens

pytestmark = pytest.mark.gpu

REF_MAIN = os.path.join(ROOT, "oracle", "_ref", "ref_main")
DROPIN = os.path.join(ROOT, "oracle", "_ref", "dropin_main")
RUN = os.path.join(ROOT, "bio-cfd_amd", "lib", "biocfd_run")
GOLDEN_DECK = os.path.join(ROOT, "tests", "golden", "deck_wing")

needs_bins = pytest.mark.skipif(not (os.path.exists(REF_MAIN) and os.path.exists(DROPIN)),
                                reason="reference driver or drop-in binary not built (make -C oracle ref)")


def run(exe, deck_dir):
    r = subprocess.run([exe], cwd=deck_dir, capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, OMP_NUM_THREADS="4"))
    assert r.returncode == 0, (exe, r.stdout[-2000:] + r.stderr[-2000:])
    return r.stdout


def written_files(d):
    """Every .dat file a run wrote: under out/ and in the working directory
    (the deck's own input .dat files excluded)."""
    inputs = {os.path.basename(p) for p in glob.glob(os.path.join(GOLDEN_DECK, "*.dat"))}
    out = sorted(os.path.join("out", os.path.basename(p)) for p in glob.glob(os.path.join(d, "out", "*.dat")))
    cwd = sorted(os.path.basename(p) for p in glob.glob(os.path.join(d, "*.dat"))
                 if os.path.basename(p) not in inputs)
    return out + cwd


def compare_runs(ref_dir, our_dir):
    names, ours = written_files(ref_dir), written_files(our_dir)
    assert names == ours, (names, ours)
    bad, identical = [], 0
    for n in names:
        a, b = os.path.join(ref_dir, n), os.path.join(our_dir, n)
        identical += open(a, "rb").read() == open(b, "rb").read()
        msg = same_numbers(a, b)
        if msg:
            bad.append(f"{n}: {msg}")
    print(f"{identical} of {len(names)} files byte-identical: {names}", flush=True)
    return names, bad


def two_copies(wd, make_deck):
    dirs = []
    for name in ("ref", "dropin"):
        d = os.path.join(wd, name)
        make_deck(d)
        os.makedirs(os.path.join(d, "out", "Chkpt"), exist_ok=True)
        dirs.append(d)
    return dirs


@needs_bins
@pytest.mark.parametrize("deck", ["deck_wing", "move"])
def test_unchanged_main_with_gpu_modules(deck):
    wd = tempfile.mkdtemp(prefix="dropin_")
    if deck == "deck_wing":
        make = lambda d: shutil.copytree(GOLDEN_DECK, d)  # noqa: E731
    else:
        make = lambda d: DK.write_deck(CS.config_nested_move(), d)  # noqa: E731
    ref_dir, our_dir = two_copies(wd, make)
    run(REF_MAIN, ref_dir)
    print(f"drop-in: unchanged main.F90 with the GPU modules on deck {deck}", flush=True)
    run(DROPIN, our_dir)
    names, bad = compare_runs(ref_dir, our_dir)
    fields = [n for n in names if "fielddata" in n]
    assert len(fields) == 6, names                      # 2 blocks x (set-up, step 1, step 2)
    for must in ("sphere_iter.dat", "out/butterfly_002.dat", "butter_f.002.dat", "butter_cellcount_f.002.dat"):
        assert must in names, (must, names)
    assert any(n.startswith("d") and "_index." in n for n in names), names
    assert not bad, bad


@needs_bins
def test_unchanged_main_restart_with_gpu_modules():
    """istart = 1: the reference's lastConditions reads a checkpoint into
    the host arrays, and the drop-in takes the fields to the GPU."""
    assert os.path.exists(RUN), "Fortran driver not built"
    wd = tempfile.mkdtemp(prefix="dropin_rst_")
    first = os.path.join(wd, "first")
    shutil.copytree(GOLDEN_DECK, first)
    os.makedirs(os.path.join(first, "out", "Chkpt"))
    r = subprocess.run([RUN, "--deck", first, first, "--output", "--checkpoint-every", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    chk = sorted(glob.glob(os.path.join(first, "out", "Chkpt", "*_butter_chkpt.*.dat")))
    assert len(chk) == 2, chk
    ref_dir, our_dir = two_copies(wd, lambda d: shutil.copytree(GOLDEN_DECK, d))
    for d in (ref_dir, our_dir):
        set_namelist(os.path.join(d, "input_data.nml"), "istart", "1")
        for path in chk:
            t = tokens(path)
            rows = []
            for q in range(0, len(t), 7):
                rec = t[q:q + 7]
                rec[5] = "0"
                rows.append(" ".join(rec))
            name = os.path.basename(path)
            g = int(name.split(".")[1])
            re_txt = name.split(".")[2] + "." + name.split(".")[3]
            with open(os.path.join(d, "out", f"aorta_chkpt.{g:03d}.{re_txt}.dat"), "w") as f:
                f.write("\n".join(rows) + "\n")
    run(REF_MAIN, ref_dir)
    print("drop-in: restart", flush=True)
    run(DROPIN, our_dir)
    names, bad = compare_runs(ref_dir, our_dir)
    assert len([n for n in names if "fielddata" in n]) == 6, names
    assert not bad, bad
