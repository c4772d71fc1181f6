"""Failure detection through the C ABI, every condition forced from the host
so that each kernel still ends normally:

* SOR budget exhaustion: a pressure solve with a tolerance of 0 (never met)
  ends every SOR call on pcItaMax, and the count the end-of-call kernels
  keep is read back per block; a tolerance of 1e30 (met at the first test)
  counts nothing.  For the reference's update-norm rule, the FP64 residual
  rule and the mixed solve (whose tolerance is the defect correction's).
  The reference checks nothing here (PcorVcor.F90:181, 351-375), so the
  count has no reference value: it is pinned by the number of SOR calls of a
  poissonSolver call (PcorVcor.F90:86-129: the coarse block once, each fine
  block twice).
* non-finite values: a NaN and an infinity written into uploaded fields are
  counted and located by the scan; a NaN in ut before a pressure solve makes
  the solve return -3 once the checks are on.
* a velocity limit below / above the field's maximum (the maximum itself is
  compared with numpy's over the downloaded fields, exactly).
* an injected allocation failure: block creation returns -1 with the
  injected text and leaves no block; the next creation succeeds."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIXED_DEFAULT = (26, 32, 128, 1e-8)


@pytest.fixture
def H():
    from biocfd_amd import hip as H

    yield H
    H.call("biocfd_health_config", 0, 0.0, 0)
    H.call("biocfd_inject_alloc_failure", 0)
    H.call("biocfd_sor_stop_rule", 0, 5)
    H.call("biocfd_mixed_config", *MIXED_DEFAULT)


def _exhausted(H, g):
    c = np.zeros(1, dtype=np.int32)
    H.call("biocfd_sor_exhausted", g, H.ptr(c))
    return int(c[0])


def _scan(H, g):
    n = np.zeros(1, dtype=np.int64)
    vmax = np.zeros(1)
    where = np.zeros(4, dtype=np.int64)
    H.call("biocfd_health_check", g, H.ptr(n), H.ptr(vmax), H.ptr(where))
    return int(n[0]), float(vmax[0]), tuple(int(x) for x in where)


def _solve(H, sim, pc_ita_max, epsi):
    """poissonSolver over every block; in mixed mode epsi becomes the FP64
    residual tolerance of the defect correction."""
    c = sim.case
    if sim.precision == "mixed":
        H.call("biocfd_mixed_config", *MIXED_DEFAULT[:3], epsi)
    H.call("biocfd_poisson_solver", 1, len(sim.blocks), pc_ita_max, epsi, c.dt, c.uc)


@pytest.mark.parametrize("precision,rule", [("fp64", 0), ("fp64", 1), ("mixed", 0)])
def test_sor_exhaustion_counted_per_block(H, precision, rule):
    from biocfd_amd import cases as CS
    from biocfd_amd.driver import Simulation

    case = CS.config_nested(nc=24, nf=24, markers=256, nsteps=1, pc_ita_max=12, epsi=0.0)
    if rule:
        H.call("biocfd_sor_stop_rule", 1, 4)
    sim = Simulation(case, precision=precision)
    try:
        if precision == "mixed":
            H.call("biocfd_mixed_config", *MIXED_DEFAULT[:3], 0.0)
        sim.step()
        counts = [_exhausted(H, g) for g in (1, 2)]
        print(f"{precision} rule {rule}: exhausted SOR calls per block {counts}")
        assert counts == [1, 2]
        _solve(H, sim, 12, 1e30)
        assert [_exhausted(H, g) for g in (1, 2)] == [0, 0]
        _solve(H, sim, 12, 0.0)
        assert [_exhausted(H, g) for g in (1, 2)] == [1, 2]
        H.call("biocfd_health_config", 1, 0.0, 1)
        with pytest.raises(H.BiocfdError) as e:
            _solve(H, sim, 12, 0.0)
        print(e.value)
        assert e.value.code == -4 and "pcItaMax" in str(e.value) and "block 1" in str(e.value)
        _solve(H, sim, 12, 1e30)
    finally:
        sim.close()


def test_scan_counts_and_locates_nonfinite_values(H):
    from biocfd_amd import cases as CS
    from biocfd_amd.driver import Simulation

    sim = Simulation(CS.config_channel())
    try:
        sim.step()
        n, vmax, where = _scan(H, 1)
        fields = {k: sim.field(1, k) for k in ("u", "v", "w", "p")}
        ref = max(float(np.abs(fields[k]).max()) for k in ("u", "v", "w"))
        print(f"clean: {n} non-finite, max |velocity| {vmax!r} (numpy {ref!r})")
        assert n == 0 and where == (-1, -1, -1, -1)
        assert vmax == ref and vmax > 0.0
        p = fields["p"].copy()
        p[2, 3, 4] = np.inf
        sim.upload(1, "p", p)
        v = fields["v"].copy()
        v[7, 5, 3] = np.nan
        v[9, 1, 2] = -np.inf
        sim.upload(1, "v", v)
        n, vmax2, where = _scan(H, 1)
        print(f"planted: {n} non-finite, first {where}")
        assert n == 3
        # v precedes p in the field order; in v, (10,2,3) precedes (8,6,4) in (k,j,i) order
        assert where == (1, 10, 2, 3)
        v[9, 1, 2] = 0.0
        sim.upload(1, "v", v)
        assert _scan(H, 1)[2] == (1, 8, 6, 4)
    finally:
        sim.close()


def test_nonfinite_value_fails_the_pressure_solve(H):
    from biocfd_amd import cases as CS
    from biocfd_amd.driver import Simulation

    sim = Simulation(CS.config_channel())
    try:
        sim.step()
        H.call("biocfd_health_config", 1, 0.0, 0)
        _solve(H, sim, 40, 0.0)  # a clean field passes
        ut = sim.field(1, "ut")
        ut[6, 5, 4] = np.nan
        sim.upload(1, "ut", ut)
        with pytest.raises(H.BiocfdError) as e:
            _solve(H, sim, 40, 0.0)
        print(e.value)
        assert e.value.code == -3 and "non-finite" in str(e.value) and "block 1" in str(e.value)
        n, _, where = _scan(H, 1)
        assert n > 0 and where[0] in (0, 1, 2, 3)
    finally:
        sim.close()


def test_velocity_limit(H):
    from biocfd_amd import cases as CS
    from biocfd_amd.driver import Simulation

    sim = Simulation(CS.config_channel())
    try:
        sim.step()
        _, vmax, _ = _scan(H, 1)
        H.call("biocfd_health_config", 1, 0.5 * vmax, 0)
        with pytest.raises(H.BiocfdError) as e:
            _solve(H, sim, 40, 0.0)
        print(e.value)
        assert e.value.code == -3 and "over the limit" in str(e.value)
        H.call("biocfd_health_config", 1, 100.0 * vmax, 0)
        _solve(H, sim, 40, 0.0)
        for bad in (-1.0, float("inf"), float("nan")):
            with pytest.raises(H.BiocfdError):
                H.call("biocfd_health_config", 1, bad, 0)
    finally:
        sim.close()


def test_injected_allocation_failure_leaves_the_library_usable(H):
    H.call("biocfd_init", 0)
    try:
        args = (40, 8, 8, 8, 1.5, 3, 0.1, 0.1, 0.1)
        for nth in (1, 3, 12):
            H.call("biocfd_inject_alloc_failure", nth)
            with pytest.raises(H.BiocfdError) as e:
                H.call("biocfd_block_create", *args)
            print(f"allocation {nth}: {e.value}")
            assert e.value.code == -1 and "injected" in str(e.value)
            out = np.zeros(6, dtype=np.int64)
            with pytest.raises(H.BiocfdError):
                H.call("biocfd_counts", 40, H.ptr(out))
        H.call("biocfd_block_create", *args)
        u = np.asfortranarray(np.random.default_rng(3).standard_normal((10, 10, 10)))
        H.call("biocfd_field_upload", 40, H.BF["u"], H.ptr(u))
        back = np.zeros((10, 10, 10), order="F")
        H.call("biocfd_field_download", 40, H.BF["u"], H.ptr(back))
        assert np.array_equal(back, u)
        H.call("biocfd_block_destroy", 40)
    finally:
        H.call("biocfd_finalize")
