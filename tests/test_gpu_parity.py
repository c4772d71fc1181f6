"""GPU parity: the HIP time step against the C oracle on the same seeded
inputs.  Integer data (tags, index lists, nearest-element indices, counts,
SOR iteration counts) must match exactly; FP64 fields within 1e-12 relative
(max |diff| / max |oracle|); in practice they are bit-identical."""
import pytest

from parity import compare, run_oracle, small_cases

pytestmark = pytest.mark.gpu

STEPS = (0, 1, 3)


@pytest.mark.parametrize("name", ["channel", "sphere", "nested", "wing", "move"])
def test_time_step_parity(name):
    from biocfd_amd.driver import Simulation

    case = small_cases()[name]
    ref = run_oracle(case, STEPS)
    sim = Simulation(case)
    try:
        bad = []
        for s in range(0, max(STEPS) + 1):
            if s > 0:
                sim.step()
            if s in STEPS:
                got = sim.dump()
                for g, (r, o) in enumerate(zip(ref[s], got)):
                    bad += compare(r, o, label=f"{name} step {s} block {g + 1}: ")
        assert not bad, "\n".join(bad[:40])
    finally:
        sim.close()
