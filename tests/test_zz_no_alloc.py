"""biocfd_stress_calculation_noalloc (the allocation-free variant of
stress_calculation, whose scratch is a static device array) against
biocfd_stress_calculation, through the C ABI: bit-identical coefficients on
two calls in a row, the second on a body with fewer elements, and a clear
error for a body beyond the scratch's maximum."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_ELEMS = 65536


def _both(sim, case):
    from biocfd_amd import hip as H

    kv = sim.blocks[0].kin_vec()
    ref = np.zeros(4)
    got = np.zeros(4)
    H.call("biocfd_stress_calculation", 1, case.mu_f, case.rho_f, H.ptr(kv), H.ptr(ref))
    H.call("biocfd_stress_calculation_noalloc", 1, case.mu_f, case.rho_f, H.ptr(kv), H.ptr(got))
    return ref, got


def test_noalloc_matches_stress_calculation_on_two_calls():
    from biocfd_amd import cases as CS
    from biocfd_amd.driver import Simulation

    results = []
    for markers in (1024, 256):  # the second call on a smaller body
        case = CS.config_sphere_64(n=24, markers=markers, nsteps=1, pc_ita_max=10, epsi=0.0)
        sim = Simulation(case)
        try:
            sim.step()
            ref, got = _both(sim, case)
        finally:
            sim.close()
        print(f"{markers} markers: {ref} {got}")
        assert np.all(np.isfinite(ref)) and np.any(ref != 0.0)
        assert np.array_equal(ref.view(np.int64), got.view(np.int64)), (ref, got)
        results.append(ref)
    assert not np.array_equal(results[0], results[1])


def test_noalloc_refuses_a_body_beyond_its_scratch():
    from biocfd_amd import hip as H

    H.call("biocfd_init", 0)
    blk = 7
    n = 8
    H.call("biocfd_block_create", blk, n, n, n, 1.7, 3, 0.125, 0.125, 0.125)
    try:
        ne = MAX_ELEMS + 2
        el = np.ones(ne, dtype=np.int64)
        sid = np.full(ne, 50, dtype=np.int64)
        box = np.array([2, n + 1, 2, n + 1, 2, n + 1], dtype=np.int64)
        H.call("biocfd_body_create", blk, 3, ne, H.ptr(el), H.ptr(el), H.ptr(el), H.ptr(sid), H.ptr(box))
        kv = np.zeros(H.BK_N)
        out = np.zeros(4)
        rc = H.lib().biocfd_stress_calculation_noalloc(blk, 1.0, 1.0, H.ptr(kv), H.ptr(out))
        assert rc != 0
        msg = H.lib().biocfd_last_error().decode(errors="replace")
        assert "maximum of 65536" in msg, msg
    finally:
        H.call("biocfd_block_destroy", blk)
