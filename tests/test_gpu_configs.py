"""GPU parity on the benchmark workloads at their configured sizes (or, where
the oracle would take hours, with a short fixed SOR budget at full size).

* sphere 256^3 (the headline workload): one time step at full size with a
  20-iteration SOR budget, FP64 against the oracle and mixed precision
  against the FP64 oracle; and a reduced grid of the same domain (64x32x32,
  512 markers, power-of-two spacing 1/8) for three steps.
* sphere 64^3 with 1024 markers, 20 steps, the reference's rule (update norm
  < 1e-8 every fifth iteration, pcItaMax = 500) as configured.
* nested two-block 2 x 128^3 with 4096 markers, heaving sphere, 2 steps.
* the 17-block uneven workload, 2 steps, with the fine blocks' SOR calls
  batched and one after the other.
* five uneven blocks whose SOR calls stop at different iterations (update
  norm tolerance 1e-6), batched and one after the other.
* the FP64 residual stopping rule: the GPU stops at the first test
  iteration whose residual is below the tolerance, with the oracle's
  iterates at that iteration.

Integers exactly, FP64 fields within 1e-12 relative (parity.compare)."""
import os
import shutil
import tempfile

import numpy as np
import pytest

from parity import compare, run_oracle

pytestmark = pytest.mark.gpu

# mixed precision: library defaults restored after each test
MIXED_DEFAULT = (26, 32, 128, 1e-8)
MIXED_CONTRACT = {"u": 1e-6, "v": 1e-6, "w": 1e-6, "ut": 1e-6, "vt": 1e-6, "wt": 1e-6, "p": 1e-6, "pc": 1e-5,
                  "b": 1e-5}
MIXED_OTHER = 1e-5
MIXED_SKIP = ("derr2", "pco")


def _oracle(case, steps):
    wd = tempfile.mkdtemp(prefix="orc_cfg_")
    try:
        return run_oracle(case, steps, wd)
    finally:
        shutil.rmtree(wd, ignore_errors=True)


def _run_gpu(case, steps, precision="fp64", setup=None):
    """{step: dump} of the GPU run at `steps` (step 0 = after the set-up)."""
    from biocfd_amd.driver import Simulation

    sim = Simulation(case, precision=precision)
    out = {}
    try:
        if setup:
            setup()
        for s in range(max(steps) + 1):
            if s > 0:
                sim.step()
            if s in steps:
                out[s] = sim.dump()
    finally:
        sim.close()
    return out


def _fp64_bad(name, ref, got):
    bad = []
    for s in ref:
        for g, (r, o) in enumerate(zip(ref[s], got[s])):
            bad += compare(r, o, label=f"{name} step {s} block {g + 1}: ")
    return bad


def _mixed_bad(name, ref, got):
    bad, report = [], []
    for s in ref:
        for g, (r, o) in enumerate(zip(ref[s], got[s])):
            for k, x in r.items():
                if k in MIXED_SKIP or k not in o or k.startswith(("ca_", "ck_")):
                    continue
                x = np.asarray(x)
                y = np.asarray(o[k])
                if x.shape != y.shape:
                    bad.append(f"{name} step {s} block {g + 1} {k}: shape {x.shape} vs {y.shape}")
                elif x.dtype.kind in "iu":
                    if not np.array_equal(x, y):
                        bad.append(f"{name} step {s} block {g + 1} {k}: {np.count_nonzero(x != y)} integer mismatches")
                elif x.size:
                    e = float(np.max(np.abs(y - x))) / max(float(np.max(np.abs(x))), 1e-300)
                    tol = MIXED_CONTRACT.get(k, MIXED_OTHER)
                    report.append(f"{k}={e:.1e}")
                    if not e <= tol:
                        bad.append(f"{name} step {s} block {g + 1} {k}: relative error {e:.3e} > {tol:.0e}")
    print(f"{name}: " + " ".join(report))
    return bad


def _mixed_setup(sweeps):
    from biocfd_amd import hip as H

    return lambda: H.call("biocfd_mixed_config", MIXED_DEFAULT[0], MIXED_DEFAULT[1], sweeps, 0.0)


# ---------------------------------------------------------------- sphere 256^3


@pytest.fixture(scope="module")
def sphere256_oracle():
    from biocfd_amd import cases as CS

    case = CS.config_sphere_256(nsteps=1, pc_ita_max=20, epsi=0.0)
    print("oracle: sphere 256^3, 1 step, 21 SOR iterations", flush=True)
    return case, _oracle(case, (1,))


@pytest.mark.timeout(900)
def test_sphere256_full_size_fp64(sphere256_oracle):
    case, ref = sphere256_oracle
    print("gpu: sphere 256^3 fp64", flush=True)
    got = _run_gpu(case, (1,))
    bad = _fp64_bad("sphere256", ref, got)
    assert not bad, "\n".join(bad[:40])


@pytest.mark.timeout(900)
def test_sphere256_full_size_mixed(sphere256_oracle):
    from biocfd_amd import hip as H

    case, ref = sphere256_oracle
    print("gpu: sphere 256^3 mixed, 16 sweeps per correction", flush=True)
    try:
        got = _run_gpu(case, (1,), precision="mixed", setup=_mixed_setup(16))
    finally:
        H.call("biocfd_mixed_config", *MIXED_DEFAULT)
    bad = _mixed_bad("sphere256 mixed", ref, got)
    assert not bad, "\n".join(bad[:40])


def _sphere256_reduced():
    from biocfd_amd import cases as CS

    return CS.config_sphere_256(n=(64, 32, 32), markers=512, nsteps=3, pc_ita_max=20, epsi=0.0)


@pytest.mark.parametrize("mom,pow2", [(4, 1), (4, 0), (0, 1)])
def test_sphere256_reduced_fp64(mom, pow2):
    """The 256^3 workload's domain, body and parameters on a 64x32x32 grid
    (spacing 1/8 on every axis: the power-of-two reciprocal path, and with
    pow2=0 the division path), with the default plane-march momentum kernel
    (4) and the per-plane one (0)."""
    from biocfd_amd import hip as H

    case = _sphere256_reduced()
    ref = _oracle(case, (0, 1, 3))
    try:
        got = _run_gpu(case, (0, 1, 3), setup=lambda: H.call("biocfd_momentum_config", mom, pow2))
    finally:
        H.call("biocfd_momentum_config", 4, 1)
    bad = _fp64_bad("sphere256 reduced", ref, got)
    assert not bad, "\n".join(bad[:40])


@pytest.mark.parametrize("sweeps", [4, 16])
def test_sphere256_reduced_mixed(sweeps):
    """Mixed precision on the reduced 256^3 workload against the FP64 oracle
    (the workload's 4 sweeps per correction, and 16)."""
    from biocfd_amd import hip as H

    case = _sphere256_reduced()
    ref = _oracle(case, (3,))
    try:
        got = _run_gpu(case, (3,), precision="mixed", setup=_mixed_setup(sweeps))
    finally:
        H.call("biocfd_mixed_config", *MIXED_DEFAULT)
    bad = _mixed_bad(f"sphere256 reduced mixed {sweeps}", ref, got)
    assert not bad, "\n".join(bad[:40])


# ---------------------------------------------------------------- sphere 64^3, nested 2 x 128^3


@pytest.mark.timeout(600)
def test_sphere64_configured():
    """64^3, 1024 markers, 20 steps, epsi 1e-8 and pcItaMax 500 as configured."""
    from biocfd_amd import cases as CS

    case = CS.config_sphere_64(n=64, markers=1024, nsteps=20, pc_ita_max=500, epsi=1e-8)
    steps = (0, 1, 20)
    print("oracle: sphere 64^3, 20 steps", flush=True)
    ref = _oracle(case, steps)
    print("gpu: sphere 64^3, 20 steps", flush=True)
    got = _run_gpu(case, steps)
    bad = _fp64_bad("sphere64", ref, got)
    assert not bad, "\n".join(bad[:40])


@pytest.mark.timeout(600)
def test_nested128_configured():
    """Coarse 128^3 on [0,4]^3 plus a 128^3 fine block on [1,3]^3 with a
    4096-marker sphere heaving in y, 2 steps, 21 SOR iterations per call."""
    from biocfd_amd import cases as CS

    case = CS.config_nested(nc=128, nf=128, markers=4096, nsteps=2, pc_ita_max=20, epsi=0.0)
    steps = (0, 2)
    print("oracle: nested 2 x 128^3, 2 steps", flush=True)
    ref = _oracle(case, steps)
    print("gpu: nested 2 x 128^3, 2 steps", flush=True)
    got = _run_gpu(case, steps)
    bad = _fp64_bad("nested128", ref, got)
    assert not bad, "\n".join(bad[:40])


# ---------------------------------------------------------------- uneven multi-block


@pytest.fixture(scope="module")
def uneven_oracle():
    from biocfd_amd import cases as CS

    case = CS.config_uneven(nsteps=2, pc_ita_max=20, epsi=0.0)
    print(f"oracle: uneven workload ({len(case.blocks)} blocks), 2 steps", flush=True)
    return case, _oracle(case, (0, 2))


@pytest.mark.timeout(600)
@pytest.mark.parametrize("batching", [1, 0])
def test_uneven_workload(uneven_oracle, batching):
    from biocfd_amd import hip as H

    case, ref = uneven_oracle
    print(f"gpu: uneven workload, batching {batching}", flush=True)
    try:
        got = _run_gpu(case, (0, 2), setup=lambda: H.call("biocfd_sor_batching", batching))
    finally:
        H.call("biocfd_sor_batching", 1)
    bad = _fp64_bad(f"uneven batching {batching}", ref, got)
    assert not bad, "\n".join(bad[:40])


@pytest.mark.parametrize("batching", [1, 0])
def test_batched_calls_stop_at_different_iterations(batching):
    """Five uneven blocks with the update-norm tolerance 1e-6: the fine
    blocks' SOR calls end at different iterations (the per-block done flag,
    the skipped tiles of stopped blocks and the pass parity per block of the
    batched launch), nIterPcor and the fields as the oracle's."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H

    case = CS.config_batched(nsteps=2, pc_ita_max=300, epsi=1e-6)
    ref = _oracle(case, (1, 2))
    iters = sorted({int(d["nIterPcor"]) for d in ref[1][1:]})
    assert len(iters) > 1, f"the fine blocks' calls all end at {iters}"
    try:
        got = _run_gpu(case, (1, 2), setup=lambda: H.call("biocfd_sor_batching", batching))
    finally:
        H.call("biocfd_sor_batching", 1)
    bad = _fp64_bad(f"batched epsi 1e-6 batching {batching}", ref, got)
    assert not bad, "\n".join(bad[:40])


# ---------------------------------------------------------------- residual stopping rule


def residual_max(d, dt):
    """max over the fluid cells of |b/dt - A pc| from a dump (b, pc, Acx,
    Acy, Acz, cell), the quantity the residual rule tests."""
    pc, b, cell = d["pc"], d["b"], d["cell"]
    ax, ay, az = d["Acx"], d["Acy"], d["Acz"]
    c = pc[1:-1, 1:-1, 1:-1]
    a1 = ax[:, 0][:, None, None]
    a2 = ax[:, 1][:, None, None]
    a3 = ax[:, 2][:, None, None]
    b1 = ay[:, 0][None, :, None]
    b2 = ay[:, 1][None, :, None]
    b3 = ay[:, 2][None, :, None]
    c1 = az[:, 0][None, None, :]
    c2 = az[:, 1][None, None, :]
    c3 = az[:, 2][None, None, :]
    ap = (a1 * pc[:-2, 1:-1, 1:-1] + a3 * pc[2:, 1:-1, 1:-1] + b1 * pc[1:-1, :-2, 1:-1] + b3 * pc[1:-1, 2:, 1:-1]
          + c1 * pc[1:-1, 1:-1, :-2] + c3 * pc[1:-1, 1:-1, 2:] + (a2 + b2 + c2) * c)
    r = b[1:-1, 1:-1, 1:-1] / dt - ap
    fluid = cell[1:-1, 1:-1, 1:-1] == 0
    return float(np.max(np.abs(r[fluid])))


@pytest.mark.parametrize("cadence", [10, 7])
def test_residual_rule_stops_at_first_test_below_tolerance(cadence):
    """Rule 1 on the 24^3 sphere: the call ends at iteration N, a multiple of
    the cadence; the oracle's iterates after N iterations are the GPU's bit
    for bit, their residual is below the tolerance, and the residual after
    N - cadence iterations is not."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H

    tol = 1e-5
    case = CS.config_sphere_64(n=24, markers=256, nsteps=1, pc_ita_max=5000, epsi=tol)
    try:
        got = _run_gpu(case, (1,), setup=lambda: H.call("biocfd_sor_stop_rule", 1, cadence))
    finally:
        H.call("biocfd_sor_stop_rule", 0, 5)
    n = int(got[1][0]["nIterPcor"])
    assert n % cadence == 0 and n < 5000, n
    assert float(got[1][0]["derr2"]) < tol
    # the oracle with a fixed budget of n and n - cadence iterations
    fixed = CS.config_sphere_64(n=24, markers=256, nsteps=1, pc_ita_max=n - 1, epsi=0.0)
    ref = _oracle(fixed, (1,))
    bad = compare(ref[1][0], got[1][0], label="residual rule vs oracle: ", skip=("derr2",))
    assert not bad, "\n".join(bad[:40])
    r_n = residual_max(ref[1][0], case.dt)
    assert r_n < tol, r_n
    # the same residual up to the rounding of a different summation order
    assert abs(r_n - float(got[1][0]["derr2"])) <= 1e-3 * tol
    before = CS.config_sphere_64(n=24, markers=256, nsteps=1, pc_ita_max=n - cadence - 1, epsi=0.0)
    r_before = residual_max(_oracle(before, (1,))[1][0], case.dt)
    assert r_before >= tol, r_before


@pytest.mark.parametrize("rule", [(0, 5), (1, 10), "batched"])
def test_sor_graphs_give_the_same_bits(rule):
    """The SOR calls with their pass chunks replayed from HIP graphs
    (biocfd_sor_graphs(1), the default) and launched pass by pass give the
    same fields, iteration counts and norms over four steps of the 32^3
    sphere, under the reference's rule (tolerance 1e-8, pcItaMax 500: calls
    of a few hundred iterations, some ending inside a graph chunk) and under
    the residual rule (tests every 10th iteration); and on the five uneven
    blocks with the fine blocks' calls batched (tolerance 1e-6, pcItaMax
    300, blocks stopping at different iterations)."""
    from biocfd_amd import cases as CS
    from biocfd_amd import hip as H

    if rule == "batched":
        case = CS.config_batched(nsteps=4, pc_ita_max=300, epsi=1e-6)
        rule = (0, 5)
    else:
        case = CS.config_sphere_64(n=32, markers=512, nsteps=4, pc_ita_max=500,
                                   epsi=1e-8 if rule[0] == 0 else 1e-6)
    runs = {}
    stats = {}
    try:
        for graphs in (0, 1):
            def setup(g=graphs):
                H.call("biocfd_sor_graphs", g)
                H.call("biocfd_sor_stop_rule", *rule)
            c0, l0 = np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64)
            H.call("biocfd_sor_graph_stats", H.ptr(c0), H.ptr(l0))
            runs[graphs] = _run_gpu(case, (1, 4), setup=setup)
            c1, l1 = np.zeros(1, dtype=np.int64), np.zeros(1, dtype=np.int64)
            H.call("biocfd_sor_graph_stats", H.ptr(c1), H.ptr(l1))
            stats[graphs] = (int(c1[0] - c0[0]), int(l1[0] - l0[0]))
    finally:
        H.call("biocfd_sor_graphs", 1)
        H.call("biocfd_sor_stop_rule", 0, 5)
    print("graphs captured / launched:", stats)
    assert stats[0] == (0, 0)
    # a few patterns captured, many more chunks replayed
    assert 0 < stats[1][0] < stats[1][1], stats
    for s in (1, 4):
        assert int(runs[0][s][0]["nIterPcor"]) > 50, runs[0][s][0]["nIterPcor"]
        for g, (a, b) in enumerate(zip(runs[0][s], runs[1][s])):
            for k, x in a.items():
                assert np.array_equal(np.asarray(x), np.asarray(b[k])), (s, g + 1, k)
