"""GPU parity: the HIP time step against the C oracle on the same seeded
inputs.  Integer data (tags, index lists, nearest-element indices, counts,
SOR iteration counts) must match exactly; FP64 fields within 1e-12 relative
(max |diff| / max |oracle|); in practice they are bit-identical."""
import pytest

from parity import compare, run_oracle, small_cases

pytestmark = pytest.mark.gpu

STEPS = (0, 1, 3)


MOM_DEFAULT = 4


@pytest.mark.parametrize("name,batching,mom,pow2", [
    ("channel", 1, 4, 1), ("sphere", 1, 4, 1), ("nested", 1, 4, 1), ("wing", 1, 4, 1), ("move", 1, 4, 1),
    ("batched", 1, 4, 1), ("batched", 0, 4, 1),
    ("sphere", 1, 0, 1), ("wing", 1, 0, 1), ("sphere", 1, 1, 1), ("wing", 1, 1, 1), ("sphere", 1, 2, 1),
    ("wing", 1, 2, 1), ("sphere", 1, 4, 0), ("wing", 1, 4, 0), ("nested", 1, 0, 0), ("wing", 1, 5, 1),
    ("wing", 1, 7, 0), ("sphere", 1, 8, 1), ("channel", 1, 3, 0), ("nested", 1, 6, 1),
    ("sphere32", 1, 4, 1), ("sphere32", 1, 0, 1), ("sphere32", 1, 1, 1), ("sphere32", 1, 2, 1), ("sphere32", 1, 4, 0),
    ("sphere32", 1, 5, 0), ("wing32", 1, 4, 1), ("wing32", 1, 2, 1), ("wing32", 1, 0, 0), ("wing32", 1, 4, 0),
    ("wing32", 1, 6, 1)])
def test_time_step_parity(name, batching, mom, pow2):
    """batching: the fine blocks' SOR calls in one launch per pass (1, the
    default) or one block after the other (0).  mom: the dense momentum
    kernel (biocfd_momentum_config): 0 = all three components per thread per
    plane, 1 = one launch per velocity component, 2 = 0 with four waves per
    SIMD, 3..8 = the plane march (4, the default: 64 x 4 column tiles of 32
    planes).  pow2: divisions by power-of-two grid spacings as products by
    the exact reciprocal (1, the default) or as divisions (0)."""
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = small_cases()[name]
    ref = run_oracle(case, STEPS)
    sim = Simulation(case)
    H.call("biocfd_sor_batching", batching)
    H.call("biocfd_momentum_config", mom, pow2)
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
        H.call("biocfd_sor_batching", 1)
        H.call("biocfd_momentum_config", MOM_DEFAULT, 1)
        sim.close()


@pytest.mark.parametrize("name,pow2", [("sphere", 1), ("wing", 0), ("sphere32", 1), ("wing32", 1), ("nested", 1)])
def test_momentum_with_divisions(name, pow2):
    """The plane march dividing by the upwind denominators and re
    (biocfd_momentum_divisions(0)) instead of correcting products with the
    reciprocals (the default): the same bits as the oracle either way."""
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = small_cases()[name]
    ref = run_oracle(case, STEPS)
    sim = Simulation(case)
    H.call("biocfd_momentum_config", MOM_DEFAULT, pow2)
    H.call("biocfd_momentum_divisions", 0)
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
        H.call("biocfd_momentum_divisions", 1)
        H.call("biocfd_momentum_config", MOM_DEFAULT, 1)
        sim.close()


@pytest.mark.parametrize("name", ["wing", "wing32"])
def test_ts_pass_in_list_order(name):
    """The thin-surface pass's deferred cells one after the other in list
    order (biocfd_ts_pass_config(0), the reference's loop order) against the
    oracle; the default evaluates them all at once."""
    from biocfd_amd import hip as H
    from biocfd_amd.driver import Simulation

    case = small_cases()[name]
    ref = run_oracle(case, STEPS)
    sim = Simulation(case)
    H.call("biocfd_ts_pass_config", 0)
    try:
        bad = []
        for s in range(0, max(STEPS) + 1):
            if s > 0:
                sim.step()
            if s in STEPS:
                for g, (r, o) in enumerate(zip(ref[s], sim.dump())):
                    bad += compare(r, o, label=f"{name} step {s} block {g + 1}: ")
        assert not bad, "\n".join(bad[:40])
    finally:
        H.call("biocfd_ts_pass_config", 1)
        sim.close()
