"""Mixed precision (FP32 red-black SOR smoothing inside FP64 defect
correction) against the FP64 oracle on the parity cases.

The pressure solve keeps the FP64 method's sweep budget (the mixed solve is
given no residual tolerance here, so both run pcItaMax + 1 sweeps), so the
FP32 path follows the FP64 iterates up to the FP32 rounding of each
correction.  Integer data (tags, index lists, nearest elements, counts,
nIterPcor) must match exactly; fields within the per-field relative
tolerances below (max |diff| / max |oracle|)."""
import numpy as np
import pytest

from parity import run_oracle, small_cases

pytestmark = pytest.mark.gpu

STEPS = 3
# the library's mixed-precision defaults (variant, planes, sweeps per correction, residual tolerance),
# restored after each test
DEFAULT = (26, 32, 128, 1e-8)
# (variant, planes, sweeps per correction): the per-phase cell-list FP32 kernel at 4 sweeps and the
# pair-owned FP32 kernels (fused multiply-adds, hardware reciprocal; one and two sweeps per pass) at
# 32, 64 and the default 128 (these cases' SOR calls are shorter than 64 sweeps)
CONFIGS = [(1, 32, 4), (15, 16, 32), (21, 32, 32), (21, 32, 64), (21, 32, 128), (25, 32, 64), (26, 32, 32),
           (26, 32, 128)]
# max |mixed - fp64| / max |fp64| allowed per field.  The FP32 rounding the corrections carry grows
# with the sweeps per correction: at 4 sweeps the fields stay within 1e-8 of the FP64 run, at 32 within
# 2e-7 (velocities, p) on these cases, so the set for 32 and 64 is the accuracy contract of the mixed
# path (1e-6 relative velocity error) with the derived arrays a decade looser.
CONTRACT = {"u": 1e-6, "v": 1e-6, "w": 1e-6, "ut": 1e-6, "vt": 1e-6, "wt": 1e-6, "p": 1e-6, "pc": 1e-5, "b": 1e-5}
TOL = {4: {"u": 1e-7, "v": 1e-7, "w": 1e-7, "ut": 1e-7, "vt": 1e-7, "wt": 1e-7, "p": 1e-7, "pc": 1e-6, "b": 1e-6},
       32: CONTRACT, 64: CONTRACT, 128: CONTRACT}
TOL_OTHER = {4: 1e-6, 32: 1e-5, 64: 1e-5, 128: 1e-5}  # every other FP64 array (ghosts, normal distances, forces ...)
SKIP = ("derr2", "pco")   # derr2 is the residual norm in the mixed solve, pco is not kept


def rel_err(ref, got):
    ref = np.asarray(ref, dtype=np.float64)
    got = np.asarray(got, dtype=np.float64)
    scale = max(float(np.max(np.abs(ref))), 1e-300) if ref.size else 1.0
    return float(np.max(np.abs(got - ref))) / scale if ref.size else 0.0


@pytest.mark.parametrize("config", CONFIGS)
@pytest.mark.parametrize("name", ["channel", "sphere", "nested", "wing"])
def test_mixed_step_against_fp64_oracle(name, config):
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = small_cases()[name]
    ref = run_oracle(case, (STEPS,))[STEPS]
    sim = Simulation(case, precision="mixed")
    try:
        H.call("biocfd_mixed_config", *config, 0.0)
        for _ in range(STEPS):
            sim.step()
        got = sim.dump()
    finally:
        sim.close()
        H.call("biocfd_mixed_config", *DEFAULT)
    bad, report = [], []
    for g, (r, o) in enumerate(zip(ref, got)):
        for k, x in r.items():
            if k in SKIP or k not in o or k.startswith(("ca_", "ck_")):
                continue
            x = np.asarray(x)
            y = np.asarray(o[k])
            if x.shape != y.shape:
                bad.append(f"block {g + 1} {k}: shape {x.shape} vs {y.shape}")
            elif x.dtype.kind in "iu":
                if not np.array_equal(x, y):
                    bad.append(f"block {g + 1} {k}: {np.count_nonzero(x != y)} integer mismatches")
            else:
                e = rel_err(x, y)
                tol = TOL[config[2]].get(k, TOL_OTHER[config[2]])
                report.append(f"{k}={e:.2e}")
                if not e <= tol:
                    bad.append(f"block {g + 1} {k}: relative error {e:.3e} > {tol:.0e}")
    print(f"{name} {config}: " + " ".join(report))
    assert not bad, "\n".join(bad[:40])


@pytest.mark.parametrize("config", CONFIGS)
def test_mixed_stops_at_residual_tolerance(config):
    """With a residual tolerance the correction loop stops early (at the
    end of a correction) and the final FP64 residual is below it."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = CS.config_sphere_64(n=24, markers=256, nsteps=2, pc_ita_max=400, epsi=0.0)
    sim = Simulation(case, precision="mixed")
    try:
        H.call("biocfd_mixed_config", *config, 1e-6)
        sim.step()
        d = sim.dump()[0]
        r = sim.residual_norm(1)
    finally:
        sim.close()
        H.call("biocfd_mixed_config", *DEFAULT)
    assert int(d["nIterPcor"]) < 401
    assert int(d["nIterPcor"]) % config[2] == 0
    assert r < 1e-6
    assert abs(float(d["derr2"]) - r) <= 1e-12 * max(1.0, r)


def test_mixed_runs_at_least_one_correction():
    """A residual tolerance the starting pc already meets does not end the
    call before its first correction: the reference's loop runs at least one
    iteration per call (PcorVcor.F90:309), so nIterPcor advances by one
    correction's sweeps."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = CS.config_sphere_64(n=24, markers=256, nsteps=1, pc_ita_max=400, epsi=0.0)
    sim = Simulation(case, precision="mixed")
    try:
        H.call("biocfd_mixed_config", 26, 32, 32, 1e30)
        sim.step()
        d = sim.dump()[0]
    finally:
        sim.close()
        H.call("biocfd_mixed_config", *DEFAULT)
    assert int(d["nIterPcor"]) == 32


@pytest.mark.parametrize("tol", [0.0, 1e-7])
def test_mixed_batched_fine_blocks_bit_identical(tol):
    """The fine blocks' mixed solves batched (one FP32 sweep launch per pass
    over every fine block whose call starts at the same iteration) give the
    bits of one block after the other.  With a residual tolerance the blocks
    stop after different numbers of corrections, so a stopped block's tiles
    are skipped while the others go on."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = CS.config_batched(nsteps=2, pc_ita_max=300, epsi=0.0)
    dumps = {}
    for batching in (1, 0):
        sim = Simulation(case, precision="mixed")
        try:
            H.call("biocfd_sor_batching", batching)
            H.call("biocfd_mixed_config", 26, 32, 16, tol)
            for _ in range(2):
                sim.step()
            dumps[batching] = sim.dump()
        finally:
            sim.close()
            H.call("biocfd_sor_batching", 1)
            H.call("biocfd_mixed_config", *DEFAULT)
    iters = [int(b["nIterPcor"]) for b in dumps[1]]
    print(f"tol {tol}: nIterPcor per block {iters}")
    if tol > 0:
        assert len(set(iters[1:])) > 1, iters
    bad = []
    for g, (a, b) in enumerate(zip(dumps[1], dumps[0])):
        for k, x in a.items():
            x = np.asarray(x)
            y = np.asarray(b[k])
            if x.shape != y.shape or x.dtype != y.dtype or x.tobytes() != y.tobytes():
                bad.append(f"block {g + 1} {k} differs")
    assert not bad, "\n".join(bad[:40])


# FP32 momentum tendency (biocfd_momentum_precision(1)) on top of the mixed
# solve: the dense pass's tendency is evaluated in FP32, whose third-order
# upwind sums cancel terms of order |u|^2, so each step adds an absolute
# velocity error of order 1e-8 (dt = 2e-3, unit inflow) to every component.
# The velocity contract (1e-6) is relative to the velocity scale (the largest
# |u|, |v|, |w| of the step): on the channel case v and w are pure 1e-3 noise,
# and the same absolute error is 6e-6 of their own scale.  p, pc and b are
# relative to their own scale.
VEL = ("u", "v", "w", "ut", "vt", "wt")
FP32_MOM_TOL = {"vel": 1e-6, "p": 1e-4, "pc": 1e-4, "b": 1e-4}


@pytest.mark.parametrize("pow2", [1, 0])
@pytest.mark.parametrize("name", ["channel", "sphere", "nested", "wing"])
def test_fp32_momentum_step_against_fp64_oracle(name, pow2):
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = small_cases()[name]
    ref = run_oracle(case, (STEPS,))[STEPS]
    sim = Simulation(case, precision="mixed")
    try:
        H.call("biocfd_mixed_config", 26, 32, 32, 0.0)
        H.call("biocfd_momentum_config", 4, pow2)
        H.call("biocfd_momentum_precision", 1)
        for _ in range(STEPS):
            sim.step()
        got = sim.dump()
    finally:
        H.call("biocfd_momentum_precision", 0)
        H.call("biocfd_momentum_config", 4, 1)
        sim.close()
        H.call("biocfd_mixed_config", *DEFAULT)
    bad, report = [], []
    for g, (r, o) in enumerate(zip(ref, got)):
        vmax = max(float(np.max(np.abs(np.asarray(r[f])))) for f in ("u", "v", "w"))
        for k, x in r.items():
            if k in SKIP or k not in o or k.startswith(("ca_", "ck_")):
                continue
            x = np.asarray(x)
            y = np.asarray(o[k])
            if x.dtype.kind in "iu":
                if not np.array_equal(x, y):
                    bad.append(f"block {g + 1} {k}: {np.count_nonzero(x != y)} integer mismatches")
            elif k in VEL:
                e = float(np.max(np.abs(y - x))) / vmax
                tol = FP32_MOM_TOL["vel"]
                report.append(f"{k}={e:.2e}")
                if not e <= tol:
                    bad.append(f"block {g + 1} {k}: error {e:.3e} of the velocity scale > {tol:.0e}")
            else:
                e = rel_err(x, y)
                tol = FP32_MOM_TOL.get(k, 1e-4)
                report.append(f"{k}={e:.2e}")
                if not e <= tol:
                    bad.append(f"block {g + 1} {k}: relative error {e:.3e} > {tol:.0e}")
    print(f"{name} pow2={pow2}: " + " ".join(report))
    assert not bad, "\n".join(bad[:40])


def test_fp32_momentum_tendency_on_the_256_workload():
    """One momentum call on the reduced 256^3 sphere workload (64x32x32 grid,
    one step in): the FP32 tendency against the FP64 one, relative to the
    tendency's own scale, per component."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = CS.config_sphere_256(n=(64, 32, 32), markers=512, nsteps=2, pc_ita_max=40)
    sim = Simulation(case)
    try:
        sim.step()
        u = {f: sim.field(1, f) for f in ("u", "v", "w")}
        out = {}
        for fp32 in (0, 1):
            H.call("biocfd_momentum_precision", fp32)
            H.call("biocfd_ns_momentum", 1, case.dt, case.re)
            H.call("biocfd_synchronize")
            out[fp32] = {f: sim.field(1, f + "t") for f in ("u", "v", "w")}
    finally:
        H.call("biocfd_momentum_precision", 0)
        sim.close()
    for f in ("u", "v", "w"):
        tend = out[0][f] - u[f]
        scale = float(np.max(np.abs(tend)))
        err = float(np.max(np.abs(out[1][f] - out[0][f])))
        print(f"{f}: tendency scale {scale:.3e}, FP32 - FP64 {err:.3e} ({err / scale:.2e})")
        assert scale > 0
        assert err <= 1e-5 * scale


def test_momentum_reciprocal_rows_fall_back_to_divisions():
    """Velocities with subnormal, tiny (1e-300) and huge (1e200) entries: the
    cells whose numerators leave the range where the corrected products
    with the reciprocals are exact are evaluated again with divisions, so
    the march gives the bits of the divisions (biocfd_momentum_divisions 0)
    everywhere, these cells included."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = CS.config_sphere_256(n=(64, 32, 32), markers=512, nsteps=2, pc_ita_max=20)
    sim = Simulation(case)
    try:
        sim.step()
        rng = np.random.default_rng(7)
        for f in ("u", "v", "w"):
            a = sim.field(1, f)
            idx = rng.integers(0, a.size, 300)
            vals = np.array([5e-324, 1e-310, 1e-300, 3e-290, 1e200, -1e-305])
            a.reshape(-1, order="F")[idx] = vals[rng.integers(0, len(vals), idx.size)]
            a = np.asfortranarray(a)
            H.call("biocfd_field_upload", 1, H.BF[f], H.ptr(a))
        out = {}
        for mode in (0, 1):
            H.call("biocfd_momentum_divisions", mode)
            H.call("biocfd_ns_momentum", 1, case.dt, case.re)
            H.call("biocfd_synchronize")
            out[mode] = {f: sim.field(1, f + "t") for f in ("u", "v", "w")}
    finally:
        H.call("biocfd_momentum_divisions", 1)
        sim.close()
    for f in ("u", "v", "w"):
        a, b = out[0][f], out[1][f]
        same = (a.view(np.int64) == b.view(np.int64)) | (np.isnan(a) & np.isnan(b))
        assert same.all(), (f, int((~same).sum()))
