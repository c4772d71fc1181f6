"""The time step with its per-block loops on several host threads (as the
reference's OpenMP loops over blocks call the boundary procedures) gives the
oracle's result: the library serialises the calls on its in-order stream."""
import pytest

from parity import compare, run_oracle, small_cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["nested", "sphere", "move"])
def test_threaded_blocks_match_oracle(name):
    from biocfd_amd.driver import Simulation

    case = small_cases()[name]
    ref = run_oracle(case, (3,))[3]
    sim = Simulation(case, threads=4)
    try:
        for _ in range(3):
            sim.step()
        got = sim.dump()
    finally:
        sim.close()
    bad = []
    for g, (r, o) in enumerate(zip(ref, got)):
        bad += compare(r, o, label=f"{name} threads=4 block {g + 1}: ")
    assert not bad, "\n".join(bad[:40])
