"""Shared helpers of the parity tests: run the C oracle on a case file and
compare record dictionaries under the stated tolerances."""
import os
import subprocess
import tempfile

import numpy as np

from biocfd_amd import cases as CS
from biocfd_amd.dumps import read_dump

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_RUN = os.path.join(ROOT, "oracle", "orc_run")
REF_RUN = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

# columns 2..n of the upwind tables are the only ones the reference defines
TABLE_PREFIX = ("ca_", "ck_")
FIELD_RTOL = 1e-12


def run_oracle(case, steps, workdir=None):
    """Dumps of the C oracle at the requested steps: {step: [block dicts]}."""
    wd = workdir or tempfile.mkdtemp(prefix="orc_")
    case.dump_steps = tuple(steps)
    path = os.path.join(wd, "case.bin")
    CS.write_case(case, path)
    subprocess.run([ORC_RUN, path, wd], check=True, capture_output=True)
    out = {}
    for s in steps:
        out[s] = [read_dump(os.path.join(wd, f"dump_s{s}_b{g + 1}.bin")) for g in range(len(case.blocks))]
    return out


def compare(ref, got, label="", field_rtol=FIELD_RTOL, skip=()):
    """List of mismatch descriptions (empty = parity)."""
    bad = []
    for k, x in ref.items():
        if k in skip or k not in got:
            continue
        y = got[k]
        x = np.asarray(x)
        y = np.asarray(y)
        if k.startswith(TABLE_PREFIX):
            x = x[:, 1:-1]
            y = y[:, 1:-1]
            x = x[:, : max(0, x.shape[1] - 1)]
            y = y[:, : max(0, y.shape[1] - 1)]
        if x.shape != y.shape:
            bad.append(f"{label}{k}: shape {x.shape} vs {y.shape}")
            continue
        if x.size == 0:
            continue
        if x.dtype.kind in "iu":
            if not np.array_equal(x, y):
                bad.append(f"{label}{k}: {np.count_nonzero(x != y)} integer mismatches")
        else:
            d = np.abs(x - y)
            scale = max(float(np.max(np.abs(x))), 1e-300)
            rel = float(np.max(d)) / scale
            if not (rel <= field_rtol):
                bad.append(f"{label}{k}: max rel diff {rel:.3e}")
    return bad


def small_cases():
    """The parity cases: single sphere, channel, nested heaving sphere, flapping body, a
    nested case whose fine block moves every step, and five uneven blocks (a
    pitching, heaving plate in one fine block, three body-free fine blocks)."""
    sphere = CS.config_sphere_64(n=24, markers=256, nsteps=3, pc_ita_max=20, epsi=0.0)
    channel = CS.config_channel(n=(16, 12, 10), nsteps=3)
    nested = CS.config_nested(nc=24, nf=24, markers=256, nsteps=3, pc_ita_max=15, epsi=0.0)
    body = CS.BodyPlacement(CS.ellipsoid_body(512, (0.25, 0.06, 0.18), surf_id=51), (0.45, 0.5, 0.5), a0_deg=20.0)
    blk = CS.BlockSpec((0, 0, 0), (1, 1, 1), (24, 24, 24), omega=1.7, body=body, moving=True)
    wing = CS.Case("wing", [blk], dt=0.002, re=300.0, nsteps=3, pc_ita_max=15, epsi=0.0, freq=0.5,
                   phase_angle_deg=90.0)
    move = CS.config_nested_move(nsteps=3)
    batched = CS.config_batched(nsteps=3)
    # power-of-two spacings (1/32 on every axis): the momentum kernels' exact-reciprocal path
    sphere32 = CS.config_sphere_64(n=32, markers=256, nsteps=3, pc_ita_max=20, epsi=0.0)
    blk32 = CS.BlockSpec((0, 0, 0), (1, 1, 1), (32, 32, 32), omega=1.7, body=body, moving=True)
    wing32 = CS.Case("wing32", [blk32], dt=0.002, re=300.0, nsteps=3, pc_ita_max=15, epsi=0.0, freq=0.5,
                     phase_angle_deg=90.0)
    return {"sphere": sphere, "channel": channel, "nested": nested, "wing": wing, "move": move, "batched": batched,
            "sphere32": sphere32, "wing32": wing32}
