"""The adjoint time step with the moving geometry (include/biocfd_hip.h,
biocfd_ad_record, biocfd_ad_geometry_tangent / _adjoint and the tlm/adj
calls with a record; bio-cfd_amd/python/biocfd_amd/adjoint.py, frozen=False)
on the reduced adjoint workload (32^3, a heaving, pitching ID-51 ellipsoid):
the element centroids, normals and lengths move with the kinematic
parameters and the body-frame marker nodes, and enter every reconstruction
and the force sums.

* dot products <T x, y> = <x, A y> of the forcing routines, the step halves
  and the force sums over fields, ghosts, the kinematic vector and the node
  positions (the geometric part with the values the forward call read);
* the whole run: dJ/dtheta and dJ/d(nodes) . D by the adjoint with
  checkpoints against tangent-linear runs (to rounding) and against central
  differences of the run itself (h = 1e-6, within 1e-5 relative): the
  gradient of the run, not of a frozen-geometry model."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("u", "v", "w", "ut", "vt", "wt", "p")
N_ITER = 20
DOT_RTOL = 1e-11
FD_H = 1e-6
FD_RTOL = 1e-5


@pytest.fixture(scope="module")
def run():
    from biocfd_amd import cases as CS
    from biocfd_amd.adjoint import AdjointRun

    case = CS.config_adjoint(n=32, markers=512, nsteps=4, pc_ita_max=N_ITER - 1, epsi=0.0)
    r = AdjointRun(case, every=2, avg_steps=2)
    yield r
    r.close()


def _sens_io(run):
    from biocfd_amd import hip as H

    c = run.sim.counts(run.g)
    n2 = int(np.prod(run.st.shape))
    n_ts = c["ts"]
    g = run.g

    def put(xs):
        H.call("biocfd_ad_reset", g)
        for q, name in enumerate(FIELDS):
            H.call("biocfd_ad_upload", g, H.BF[name], H.ptr(xs[q]))
        for w in range(len(H.BGH_NAMES)):
            H.call("biocfd_ad_ghost_upload", g, w, H.ptr(xs[7 + w]), n_ts)

    def get():
        out = []
        for name in FIELDS:
            a = np.zeros(n2)
            H.call("biocfd_ad_download", g, H.BF[name], H.ptr(a))
            out.append(a)
        for w in range(len(H.BGH_NAMES)):
            a = np.zeros(n_ts)
            H.call("biocfd_ad_ghost_download", g, w, H.ptr(a), n_ts)
            out.append(a)
        return out

    def rand(rng):
        return [rng.standard_normal(n2) for _ in FIELDS] + [rng.standard_normal(n_ts) for _ in H.BGH_NAMES]

    return put, get, rand


def _dot(a, b):
    return sum(float(np.dot(x, y)) for x, y in zip(a, b))


def _recorded_step(run, n):
    """The state before step n (pre) and after it (post), step n's forcing
    calls recorded in slot 0."""
    from biocfd_amd import hip as H
    from biocfd_amd.adjoint import POST_SLOT, PRE_SLOT

    run._start(None, None)
    for _ in range(n - 1):
        run.sim.step()
    run._save(PRE_SLOT)
    H.call("biocfd_ad_record", run.g, 0)
    run.sim.step()
    H.call("biocfd_ad_record", run.g, -1)
    run._save(POST_SLOT)
    return PRE_SLOT, POST_SLOT


PIECES = [("forcing0", "pre"), ("forcing1", "pre"), ("forcing2", "post"), ("forcing3", "post"), ("forcing4", "post"),
          ("forcing5", "post"), ("step_start", "pre"), ("step_end", "post"), ("stress", "post")]


@pytest.mark.parametrize("piece,when", PIECES)
def test_geometric_piece_dot_product(run, piece, when):
    from biocfd_amd import hip as H

    print(f"geometric dot product: {piece}", flush=True)
    c, g = run.case, run.g
    pre, post = _recorded_step(run, 2)
    run._restore(pre if when == "pre" else post)
    put, get, rand = _sens_io(run)
    rng = np.random.default_rng(21)
    nn = run.st.nn
    kv = run._kin()
    x = rand(rng)
    kd = rng.standard_normal(H.BK_N)
    X = [rng.standard_normal(nn) for _ in range(3)]
    put(x)
    H.call("biocfd_ad_geometry_tangent", g, float(c.inor), *(H.ptr(a) for a in X))
    out = np.zeros(4)
    if piece.startswith("forcing"):
        H.call("biocfd_tlm_forcing", g, int(piece[-1]), H.ptr(kv), H.ptr(kd), 0)
    elif piece == "step_start":
        H.call("biocfd_tlm_step_start", g, run.n_iter, c.dt, c.re, c.uc, H.ptr(kv), H.ptr(kd), 0)
    elif piece == "step_end":
        H.call("biocfd_tlm_step_end", g, H.ptr(kv), H.ptr(kd), 0)
    else:
        H.call("biocfd_tlm_stress", g, c.mu_f, c.rho_f, H.ptr(kv), H.ptr(kd), 1, H.ptr(out))
    tx = get()
    y = rand(rng)
    sbar = rng.standard_normal(4)
    if piece == "stress":
        y = [np.zeros_like(a) for a in y]
    put(y)
    if piece.startswith("forcing"):
        H.call("biocfd_adj_forcing", g, int(piece[-1]), H.ptr(kv), 0)
    elif piece == "step_start":
        H.call("biocfd_adj_step_start", g, run.n_iter, c.dt, c.re, c.uc, H.ptr(kv), 0)
    elif piece == "step_end":
        H.call("biocfd_adj_step_end", g, H.ptr(kv), 0)
    else:
        H.call("biocfd_adj_stress", g, c.mu_f, c.rho_f, H.ptr(kv), H.ptr(sbar), 1)
    ay = get()
    kb = np.zeros(H.BK_N)
    H.call("biocfd_ad_kin", g, H.ptr(kb), 1)
    Xb = [np.zeros(nn) for _ in range(3)]
    H.call("biocfd_ad_geometry_adjoint", g, float(c.inor), *(H.ptr(a) for a in Xb), 1)
    lhs = _dot(tx, y) + (float(np.dot(out, sbar)) if piece == "stress" else 0.0)
    rhs = _dot(x, ay) + float(np.dot(kd, kb)) + _dot(X, Xb)
    assert abs(lhs - rhs) <= DOT_RTOL * max(abs(lhs), abs(rhs), 1.0), (piece, lhs, rhs)
    geo = _dot(X, Xb)
    assert geo != 0.0, piece   # the geometry reaches the piece


def test_gradient_of_the_run(run):
    """dJ/dtheta and dJ/d(body nodes) . D: adjoint with checkpoints every
    two steps, tangent-linear runs and central differences of the run."""
    from biocfd_amd.adjoint import N_THETA

    print("run gradient: 4 steps, checkpoints every 2", flush=True)
    J, (grad, gnodes) = run.gradient()
    tlm = np.zeros(N_THETA)
    fd = np.zeros(N_THETA)
    for q in range(N_THETA):
        e = np.zeros(N_THETA)
        e[q] = 1.0
        Jt, tlm[q] = run.tangent(e)
        assert Jt == J
        fd[q] = (run.forward(run.theta0 + FD_H * e) - run.forward(run.theta0 - FD_H * e)) / (2 * FD_H)
    scale = float(np.max(np.abs(tlm)))
    print("J", J, "adjoint", grad, "tangent", tlm, "fd", fd, flush=True)
    assert np.max(np.abs(grad - tlm)) <= 1e-10 * scale, (grad, tlm)
    assert np.max(np.abs(tlm - fd)) <= FD_RTOL * scale, (tlm, fd)
    rng = np.random.default_rng(31)
    D = [rng.standard_normal(run.st.nn) * 0.1 for _ in range(3)]
    adj = float(sum(np.dot(gnodes[:, q], D[q]) for q in range(3)))
    _, tl = run.tangent(np.zeros(N_THETA), dnodes=D)
    n0 = run.nodes0
    Jp = run.forward(nodes=tuple(a + FD_H * d for a, d in zip(n0, D)))
    Jm = run.forward(nodes=tuple(a - FD_H * d for a, d in zip(n0, D)))
    fdn = (Jp - Jm) / (2 * FD_H)
    print("nodes: adjoint", adj, "tangent", tl, "fd", fdn, flush=True)
    assert abs(adj - tl) <= 1e-10 * abs(tl), (adj, tl)
    assert abs(tl - fdn) <= FD_RTOL * abs(tl), (tl, fdn)


def test_autograd_wrapper_with_nodes(run):
    import torch
    from adjoint_harness import BodyForce

    theta = torch.tensor(run.theta0, dtype=torch.float64, requires_grad=True)
    nodes = torch.from_numpy(np.stack(run.nodes0, axis=1)).requires_grad_(True)
    J = BodyForce.apply(theta, nodes, run)
    J.backward()
    _, (g, gn) = run.gradient()
    # two runs of the backward pass: its scatters add with atomics, whose
    # order (and so the last bits) differ from run to run (measured: 1e-15
    # of the largest entry); the wrapper must return that gradient
    assert np.max(np.abs(theta.grad.numpy() - g)) <= 1e-12 * np.max(np.abs(g))
    assert np.max(np.abs(nodes.grad.numpy() - gn)) <= 1e-12 * np.max(np.abs(gn))
