"""The tangent-linear and adjoint time step (include/biocfd_hip.h,
biocfd_tlm_* / biocfd_adj_*; bio-cfd_amd/python/biocfd_amd/adjoint.py) on the
reduced adjoint workload: a heaving, pitching ellipsoid (ID 51, so every IB
cell is a thin-surface cell) in [0,3]^3 at 32^3, geometry frozen (the moving
geometry's part: tests/test_gpu_adjoint_geometry.py).

* every piece: dot products <T x, y> = <x, A y> between its tangent-linear
  pass T and its adjoint pass A over all seven fields, the fourteen ghost
  arrays and the kinematic vector (the forward is linear in all of them but
  the momentum stencil, whose T and A are both taken at the same state);
* the tangent-linear pieces against central differences of the forward
  routines themselves (the momentum stencil in u, v, w, p; the forcing
  routines in the kinematic vector);
* a checkpoint replays the forward bit for bit;
* the whole run: dJ/dtheta of the time-averaged x-force by the adjoint
  with checkpointing against the tangent-linear runs (to rounding) and
  against central differences of J (h = 1e-6, within 1e-5 relative)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "ut", "vt", "wt", "p")
N_ITER = 20
DOT_RTOL = 1e-11
FD_H = 1e-6
FD_RTOL = 1e-5


def _case(nsteps=4):
    from biocfd_amd import cases as CS

    return CS.config_adjoint(n=32, markers=512, nsteps=nsteps, pc_ita_max=N_ITER - 1, epsi=0.0)


@pytest.fixture(scope="module")
def run():
    from biocfd_amd.adjoint import AdjointRun

    r = AdjointRun(_case(), every=2, avg_steps=2, frozen=True)
    yield r
    r.close()


class Sens:
    """Upload / download of every sensitivity array of block g."""

    def __init__(self, g, n2, n_ts):
        self.g, self.n2, self.n_ts = g, n2, n_ts

    def put(self, xs):
        from biocfd_amd import hip as H

        H.call("biocfd_ad_reset", self.g)
        for q, name in enumerate(FIELDS):
            H.call("biocfd_ad_upload", self.g, H.BF[name], H.ptr(xs[q]))
        for w in range(len(H.BGH_NAMES)):
            H.call("biocfd_ad_ghost_upload", self.g, w, H.ptr(xs[7 + w]), self.n_ts)

    def get(self):
        from biocfd_amd import hip as H

        out = []
        for name in FIELDS:
            a = np.zeros(self.n2)
            H.call("biocfd_ad_download", self.g, H.BF[name], H.ptr(a))
            out.append(a)
        for w in range(len(H.BGH_NAMES)):
            a = np.zeros(self.n_ts)
            H.call("biocfd_ad_ghost_download", self.g, w, H.ptr(a), self.n_ts)
            out.append(a)
        return out

    def kin(self):
        from biocfd_amd import hip as H

        kb = np.zeros(H.BK_N)
        H.call("biocfd_ad_kin", self.g, H.ptr(kb), 1)
        return kb

    def random(self, rng):
        from biocfd_amd import hip as H

        return [rng.standard_normal(self.n2) for _ in FIELDS] + [rng.standard_normal(self.n_ts)
                                                                 for _ in H.BGH_NAMES]


def _dot(a, b):
    return sum(float(np.dot(x, y)) for x, y in zip(a, b))


def _sens(run):
    c = run.sim.counts(run.g)
    st = run.st
    n2 = int(np.prod(st.shape))
    return Sens(run.g, n2, c["ts"])


def _at_step(run, n):
    run.theta = run.theta0.copy()
    run._restore(0)
    for _ in range(n):
        run.sim.step()


def test_workload_has_thin_surface_cells(run):
    _at_step(run, 1)
    c = run.sim.counts(run.g)
    assert c["ts"] > 0 and c["ts"] == c["ib"], c


PIECES = ["momentum", "velocity_bc", "solid_cell_bc", "forcing0", "forcing1", "forcing2", "forcing3", "forcing4",
          "forcing5", "poisson", "step_start", "step_end"]


def _tlm(run, piece, kv, kd):
    from biocfd_amd import hip as H

    c, g = run.case, run.g
    if piece == "momentum":
        H.call("biocfd_tlm_momentum", g, c.dt, c.re)
    elif piece == "velocity_bc":
        H.call("biocfd_tlm_velocity_bc", g, c.dt, c.uc)
    elif piece == "solid_cell_bc":
        H.call("biocfd_tlm_solid_cell_bc", g)
    elif piece.startswith("forcing"):
        H.call("biocfd_tlm_forcing", g, int(piece[-1]), H.ptr(kv), H.ptr(kd), -1)
    elif piece == "poisson":
        H.call("biocfd_tlm_poisson", g, run.n_iter, c.dt, c.uc)
    elif piece == "step_start":
        H.call("biocfd_tlm_step_start", g, run.n_iter, c.dt, c.re, c.uc, H.ptr(kv), H.ptr(kd), -1)
    elif piece == "step_end":
        H.call("biocfd_tlm_step_end", g, H.ptr(kv), H.ptr(kd), -1)


def _adj(run, piece, kv):
    from biocfd_amd import hip as H

    c, g = run.case, run.g
    if piece == "momentum":
        H.call("biocfd_adj_momentum", g, c.dt, c.re)
    elif piece == "velocity_bc":
        H.call("biocfd_adj_velocity_bc", g, c.dt, c.uc)
    elif piece == "solid_cell_bc":
        H.call("biocfd_adj_solid_cell_bc", g)
    elif piece.startswith("forcing"):
        H.call("biocfd_adj_forcing", g, int(piece[-1]), H.ptr(kv), -1)
    elif piece == "poisson":
        H.call("biocfd_adj_poisson", g, run.n_iter, c.dt, c.uc)
    elif piece == "step_start":
        H.call("biocfd_adj_step_start", g, run.n_iter, c.dt, c.re, c.uc, H.ptr(kv), -1)
    elif piece == "step_end":
        H.call("biocfd_adj_step_end", g, H.ptr(kv), -1)


@pytest.mark.parametrize("piece", PIECES)
def test_piece_dot_product(run, piece):
    from biocfd_amd import hip as H

    print(f"adjoint dot product: {piece}", flush=True)
    _at_step(run, 2)
    S = _sens(run)
    rng = np.random.default_rng(11)
    kv = run._kin()
    x = S.random(rng)
    kd = rng.standard_normal(H.BK_N)
    S.put(x)
    _tlm(run, piece, kv, kd)
    tx = S.get()
    y = S.random(rng)
    S.put(y)
    _adj(run, piece, kv)
    ay = S.get()
    kb = S.kin()
    lhs = _dot(tx, y)
    rhs = _dot(x, ay) + float(np.dot(kd, kb))
    assert abs(lhs - rhs) <= DOT_RTOL * max(abs(lhs), abs(rhs), 1.0), (piece, lhs, rhs)


def test_stress_dot_product(run):
    from biocfd_amd import hip as H

    _at_step(run, 2)
    S = _sens(run)
    rng = np.random.default_rng(12)
    c = run.case
    kv = run._kin()
    x = S.random(rng)
    kd = rng.standard_normal(H.BK_N)
    S.put(x)
    out = np.zeros(4)
    H.call("biocfd_tlm_stress", run.g, c.mu_f, c.rho_f, H.ptr(kv), H.ptr(kd), 0, H.ptr(out))
    sbar = rng.standard_normal(4)
    S.put([np.zeros_like(a) for a in x])
    H.call("biocfd_adj_stress", run.g, c.mu_f, c.rho_f, H.ptr(kv), H.ptr(sbar), 0)
    ay = S.get()
    kb = S.kin()
    lhs = float(np.dot(out, sbar))
    rhs = _dot(x, ay) + float(np.dot(kd, kb))
    assert abs(lhs - rhs) <= DOT_RTOL * max(abs(lhs), abs(rhs), 1.0), (lhs, rhs)


def _fields(run):
    return [run.sim.field(run.g, n).ravel(order="F") for n in FIELDS]


def _put_fields(run, xs):
    from biocfd_amd import hip as H

    for q, name in enumerate(FIELDS):
        a = np.ascontiguousarray(xs[q])
        H.call("biocfd_field_upload", run.g, H.BF[name], H.ptr(a))


def test_momentum_tangent_matches_central_differences(run):
    """d(ut, vt, wt, u, v, w, p)/d(u, v, w, p) of nsMomentum2order (with the
    thin-surface pass) along a random direction.  The upwind terms hold
    |velocity|: an operand within h of zero has a kink inside the difference
    interval, so a few entries (at most 1e-4 of them) may differ; every other
    entry agrees to 1e-6 of the largest tangent."""
    from biocfd_amd import hip as H

    _at_step(run, 2)
    S = _sens(run)
    c = run.case
    x0 = _fields(run)
    rng = np.random.default_rng(13)
    d = [rng.standard_normal(S.n2) * (1.0 if name in ("u", "v", "w", "p") else 0.0) for name in FIELDS]
    h = 1e-8
    outs = []
    for sgn in (1.0, -1.0):
        _put_fields(run, [a + sgn * h * b for a, b in zip(x0, d)])
        H.call("biocfd_ns_momentum", run.g, c.dt, c.re)
        outs.append(_fields(run))
    fd = [(a - b) / (2 * h) for a, b in zip(*outs)]
    _put_fields(run, x0)
    S.put(d + [np.zeros(S.n_ts) for _ in H.BGH_NAMES])
    H.call("biocfd_tlm_momentum", run.g, c.dt, c.re)
    t = S.get()[:7]
    for name, a, b in zip(FIELDS, t, fd):
        scale = max(float(np.max(np.abs(a))), 1e-300)
        bad = int(np.count_nonzero(np.abs(a - b) > 1e-6 * scale))
        assert bad <= 1e-4 * a.size, (name, bad, float(np.max(np.abs(a - b))) / scale)


@pytest.mark.parametrize("routine", [0, 1, 2, 3, 4, 5])
def test_forcing_tangent_matches_central_differences_in_kin(run, routine):
    """d(fields, ghosts)/d(kin) of each forcing routine (the pressure routines
    are quadratic in the pitch rate)."""
    from biocfd_amd import hip as H

    names = ["biocfd_velocity_forcing1", "biocfd_pressure_forcing1", "biocfd_velocity_forcing_field",
             "biocfd_pressure_forcing_field", "biocfd_velocity_forcing_ghost", "biocfd_pressure_forcing_ghost"]
    _at_step(run, 2)
    S = _sens(run)
    x0 = _fields(run)
    kv = run._kin()
    rng = np.random.default_rng(14 + routine)
    kd = np.zeros(H.BK_N)
    kd[:7] = rng.standard_normal(7)
    h = 1e-4
    outs = []
    for sgn in (1.0, -1.0):
        _put_fields(run, x0)
        k = kv + sgn * h * kd
        H.call(names[routine], run.g, H.ptr(k))
        f = _fields(run)
        gh = []
        for w in range(len(H.BGH_NAMES)):
            a = np.zeros(S.n_ts)
            H.call("biocfd_ghost_download", run.g, w, H.ptr(a), S.n_ts)
            gh.append(a)
        outs.append(f + gh)
    fd = [(a - b) / (2 * h) for a, b in zip(*outs)]
    _put_fields(run, x0)
    S.put([np.zeros(S.n2) for _ in FIELDS] + [np.zeros(S.n_ts) for _ in H.BGH_NAMES])
    H.call("biocfd_tlm_forcing", run.g, routine, H.ptr(kv), H.ptr(kd), -1)
    t = S.get()
    worst = 0.0
    scale = max(max(float(np.max(np.abs(a))) for a in t), 1e-300)
    for a, b in zip(t, fd):
        worst = max(worst, float(np.max(np.abs(a - b))) if a.size else 0.0)
    assert worst <= 1e-7 * scale, (routine, worst, scale)


def test_checkpoint_replays_the_forward_bit_for_bit(run):
    from biocfd_amd import hip as H

    run._restore(0)
    run.sim.step()
    run._save(40)
    run.sim.step()
    run.sim.step()
    ref = _fields(run) + [run.sim.tag(run.g, "cell").ravel()]
    run._restore(40)
    run.sim.step()
    run.sim.step()
    got = _fields(run) + [run.sim.tag(run.g, "cell").ravel()]
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    H.call("biocfd_snapshot_free", run.g, 40)


def test_gradient_adjoint_tangent_and_finite_differences(run):
    """dJ/dtheta (heave amplitude, pitch amplitude, phase, frequency) of the
    mean x-force over the last two of four steps: the adjoint with
    checkpoints every two steps, four tangent-linear runs, central
    differences of J."""
    from biocfd_amd.adjoint import N_THETA

    print("adjoint gradient: 4 steps, checkpoints every 2", flush=True)
    J, grad = run.gradient()
    tlm = np.zeros(N_THETA)
    fd = np.zeros(N_THETA)
    for q in range(N_THETA):
        e = np.zeros(N_THETA)
        e[q] = 1.0
        Jt, tlm[q] = run.tangent(e)
        assert Jt == J
        fd[q] = (run.forward(run.theta0 + FD_H * e) - run.forward(run.theta0 - FD_H * e)) / (2 * FD_H)
    scale = float(np.max(np.abs(tlm)))
    assert scale > 0
    print("J", J, "adjoint", grad, "tangent", tlm, "fd", fd, flush=True)
    assert np.max(np.abs(grad - tlm)) <= 1e-10 * scale, (grad, tlm)
    assert np.max(np.abs(tlm - fd)) <= FD_RTOL * scale, (tlm, fd)


def test_autograd_wrapper(run):
    import torch
    from adjoint_harness import BodyForce

    theta = torch.tensor(run.theta0, dtype=torch.float64, requires_grad=True)
    nodes = torch.from_numpy(np.stack(run.nodes0, axis=1))
    J = BodyForce.apply(theta, nodes, run)
    J.backward()
    _, ref = run.gradient()
    # a second backward pass: atomics add in a different order (last bits)
    assert np.max(np.abs(theta.grad.numpy() - ref)) <= 1e-12 * np.max(np.abs(ref))
