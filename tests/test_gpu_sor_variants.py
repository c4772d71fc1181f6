"""Every SOR kernel variant (colour-split sweeps, fused red+black tiles of
each compiled shape, several planes-per-workgroup counts) gives the oracle's
time step: the SOR iteration count exactly and the fields within 1e-12
relative (bit-identical in practice)."""
import pytest

from parity import compare, run_oracle, small_cases

pytestmark = pytest.mark.gpu

# the library's default SOR variant, restored after each test
DEFAULT_VARIANT = 26


def n_variants():
    from biocfd_amd import tune

    return len(tune.variants())


def cases():
    from biocfd_amd import cases as CS

    c = dict(small_cases())
    # convergence-terminated SOR calls (220, 200 and 185 iterations)
    c["sphere_conv"] = CS.config_sphere_64(n=24, markers=256, nsteps=3, pc_ita_max=300, epsi=1e-6)
    return c


@pytest.mark.parametrize("name", ["sphere", "nested", "sphere_conv"])
def test_sor_variants_parity(name):
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = cases()[name]
    ref = run_oracle(case, (3,))[3]
    bad = []
    try:
        for v in range(n_variants()):
            for planes in ((1, 5, 32) if v else (32,)):
                print(f"case {name} variant {v} planes {planes}", flush=True)
                H.call("biocfd_sor_config", v, planes)
                sim = Simulation(case)
                try:
                    for _ in range(3):
                        sim.step()
                    got = sim.dump()
                finally:
                    sim.close()
                for g, (r, o) in enumerate(zip(ref, got)):
                    bad += compare(r, o, label=f"{name} variant {v} planes {planes} block {g + 1}: ")
    finally:
        H.call("biocfd_sor_config", DEFAULT_VARIANT, 32)
    assert not bad, "\n".join(bad[:40])


def test_sor_time_leaves_state_unchanged():
    """The autotuning probe restores pc and the block's SOR state, so tuning
    inside a run does not change its results."""
    import numpy as np

    from biocfd_amd import hip as H
    from biocfd_amd import tune
    from biocfd_amd.driver import Simulation

    case = small_cases()["sphere"]
    ref = run_oracle(case, (2,))[2]
    sim = Simulation(case)
    try:
        sim.step()
        before = sim.dump()[0]
        v, planes, table = tune.autotune(block=1, dt=case.dt, iters=7, planes=(3, 32))
        assert len(table) == 1 + 2 * (len(tune.variants()) - 1)
        after = sim.dump()[0]
        for k in ("pc", "p", "u", "nIterPcor", "derr2"):
            assert np.array_equal(np.asarray(before[k]), np.asarray(after[k])), k
        sim.step()
        bad = compare(ref[0], sim.dump()[0], label=f"after tuning (variant {v}, planes {planes}): ")
        assert not bad, "\n".join(bad[:40])
    finally:
        sim.close()
        H.call("biocfd_sor_config", DEFAULT_VARIANT, 32)


def test_stop_on_first_iteration_of_a_pass():
    """Under the reference's rule every pass runs two iterations, so a check
    iteration (every fifth) is a pass's first every other time; a call that
    stops there redoes that iteration alone from the pass's source.  The
    convergence-terminated case stops at 185 (a pass's first iteration) and
    at 220 and 200 (a pass's last): every step matches the oracle."""
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = cases()["sphere_conv"]
    steps = (1, 2, 3)
    ref = run_oracle(case, steps)
    iters = [int(ref[s][0]["nIterPcor"]) for s in steps]
    print(f"nIterPcor per step {iters}")
    assert any(n % 10 == 5 for n in iters) and any(n % 10 == 0 for n in iters), iters
    H.call("biocfd_sor_config", DEFAULT_VARIANT, 32)
    sim = Simulation(case)
    bad = []
    try:
        for s in steps:
            sim.step()
            for g, (r, o) in enumerate(zip(ref[s], sim.dump())):
                bad += compare(r, o, label=f"step {s} block {g + 1}: ")
    finally:
        sim.close()
    assert not bad, "\n".join(bad[:40])
