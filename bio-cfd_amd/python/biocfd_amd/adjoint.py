"""Tangent-linear and adjoint time steps of a single block with a moving body
over the C ABI (include/biocfd_hip.h: biocfd_tlm_*, biocfd_adj_*,
biocfd_ad_*, biocfd_snapshot_*), with step checkpointing and recomputation.

The function differentiated is

    J = mean over the last M steps of F_x,

F_x the body's x-force after a step: the viscous and pressure x sums of
stress_calculation (StressCalc.f90:173-176), as a function of the kinematic
parameters theta = (heave amplitude, pitch amplitude [rad], phase [rad],
frequency of heave and pitch) and of the body's marker nodes (the body-frame
node coordinates the kinematics place every step, Search.F90:141-208).

Two models:

* frozen=False (the default): the run itself.  theta and the nodes move the
  body, so every step's element centroids, normals and lengths depend on
  them, and with them every reconstruction (normal distance, image point,
  cube weights, surface term) and the force sums; the body velocities and
  accelerations depend on theta.  The derivative is taken at the run's
  discrete structure (IB cells, nearest elements, image-point cubes, TS ghost
  choices), which a small perturbation leaves unchanged; it is the
  derivative central differences of the run see.
* frozen=True: the body's positions, and with them all geometry, stay those
  of the base parameters; theta enters through the velocities and
  accelerations alone.

The SOR count is fixed (epsi 0: pcItaMax + 1 iterations per solve), so
every step is differentiable.  Checkpoints: the forward pass keeps the
block's state every `every` steps (biocfd_snapshot_save); the backward pass
replays each segment from its checkpoint, keeping every step of it and what
its forcing calls read (biocfd_ad_record), and runs the adjoint steps
backwards through it.
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

from . import cases as CS
from . import hip as H
from .driver import S, Simulation

# kinematic vector slots (enum biocfd_kin)
KX, KYDOT, KYDDOT, KTD1, KTDD1, KTD2, KTDD2, KALPHA, KPIVX, KPIVY, KPIVZ = range(11)
K_PI = 3.141592653589793115997963468544185161590576171875   # the library's pi
N_THETA = 4
THETA_NAMES = ("heave amplitude", "pitch amplitude", "phase", "frequency")
SEG_SLOT = 2048   # first snapshot slot of the replayed segment
PRE_SLOT, POST_SLOT = 4000, 4001


def kin_velocities(theta, t: float, dt: float) -> np.ndarray:
    """(ydot, yddot, thetaDot1, thetaDDot1) of computeSurfaceVariables at time
    t (Search.F90:141-165), the library's expressions in its operand order
    (biocfd_compute_surface_variables), heave and pitch at one frequency."""
    yamp, a0, phase, freq = (float(x) for x in theta)
    ph = 2.0 * K_PI * freq * (t + dt) + phase
    wtheta = 2.0 * K_PI * freq
    td1 = wtheta * a0 * math.cos(ph)
    tdd1 = -wtheta * wtheta * a0 * math.sin(ph)
    angt = 2.0 * K_PI * freq
    ydot = angt * yamp * math.cos(angt * t)
    yddot = -angt * angt * yamp * math.sin(angt * t)
    return np.array([ydot, yddot, td1, tdd1])


def kin_velocities_jacobian(theta, t: float, dt: float) -> np.ndarray:
    """d(ydot, yddot, thetaDot1, thetaDDot1) / d theta, (4, 4)."""
    yamp, a0, phase, freq = (float(x) for x in theta)
    tau = t + dt
    w = 2.0 * K_PI * freq
    ph = w * tau + phase
    c, s = math.cos(ph), math.sin(ph)
    a = 2.0 * K_PI * freq
    ca, sa = math.cos(a * t), math.sin(a * t)
    J = np.zeros((4, N_THETA))
    # ydot = a y cos(a t), yddot = -a^2 y sin(a t), a = 2 pi f
    J[0, 0] = a * ca
    J[0, 3] = 2.0 * K_PI * yamp * ca - a * yamp * sa * 2.0 * K_PI * t
    J[1, 0] = -a * a * sa
    J[1, 3] = -2.0 * a * 2.0 * K_PI * yamp * sa - a * a * yamp * ca * 2.0 * K_PI * t
    # thetaDot1 = w a0 cos(ph), thetaDDot1 = -w^2 a0 sin(ph), ph = w (t + dt) + phase
    J[2, 1] = w * c
    J[2, 2] = -w * a0 * s
    J[2, 3] = 2.0 * K_PI * a0 * c - w * a0 * s * 2.0 * K_PI * tau
    J[3, 1] = -w * w * s
    J[3, 2] = -w * w * a0 * c
    J[3, 3] = -2.0 * w * 2.0 * K_PI * a0 * s - w * w * a0 * c * 2.0 * K_PI * tau
    return J


def heave_jacobian(theta, t: float) -> np.ndarray:
    """d yt / d theta, yt = yamp sin(2 pi f t) (the heave displacement)."""
    yamp, _, _, freq = (float(x) for x in theta)
    a = 2.0 * K_PI * freq
    return np.array([math.sin(a * t), 0.0, 0.0, yamp * math.cos(a * t) * 2.0 * K_PI * t])


def pitch_jacobian(theta, t: float, dt: float) -> Tuple[float, np.ndarray]:
    """aoa1 = a0 sin(2 pi f (t + dt) + phase) and its gradient in theta."""
    _, a0, phase, freq = (float(x) for x in theta)
    tau = t + dt
    ph = 2.0 * K_PI * freq * tau + phase
    return a0 * math.sin(ph), np.array([0.0, math.sin(ph), a0 * math.cos(ph), a0 * math.cos(ph) * 2.0 * K_PI * tau])


def _embed(v4: np.ndarray, kv: np.ndarray) -> None:
    kv[KYDOT], kv[KYDDOT], kv[KTD1], kv[KTDD1] = v4
    kv[KTD2], kv[KTDD2] = -v4[2], -v4[3]


class AdjointRun:
    """The adjoint workload on the GPU: forward, tangent-linear and adjoint
    passes of `case` (one block with a body, epsi 0)."""

    def __init__(self, case: CS.Case, every: int = 8, avg_steps: Optional[int] = None, frozen: bool = False,
                 tune: bool = False):
        """every: steps per checkpoint; avg_steps: the averaging window M
        (default half the steps); tune: pick the SOR kernel variant for the
        block by timing the candidates (biocfd_amd.tune; every variant gives
        the same bits)."""
        if len(case.blocks) != 1 or case.blocks[0].body is None:
            raise ValueError("the adjoint driver runs a single block with a body")
        if case.epsi != 0.0:
            raise ValueError("the adjoint driver needs a fixed SOR count (epsi = 0)")
        body = case.blocks[0].body
        if body.bfreq != case.freq:
            raise ValueError("heave and pitch frequencies must be equal (one frequency parameter)")
        if case.dxmin != 0.0:
            raise ValueError("the adjoint driver needs dxmin = 0 (no drift)")
        self.case = case
        self.frozen = bool(frozen)
        self.every = max(1, int(every))
        self.nsteps = case.nsteps
        self.m = int(avg_steps) if avg_steps else max(1, case.nsteps // 2)
        if not 1 <= self.m <= self.nsteps:
            raise ValueError(f"averaging window {self.m} outside 1..{self.nsteps}")
        self.n_iter = case.pc_ita_max + 1
        self.theta0 = np.array([body.yamp, body.a0_deg * math.pi / 180.0, case.phase_angle_deg * math.pi / 180.0,
                                case.freq])
        self.theta = self.theta0.copy()
        self.sim = Simulation(case)
        self.st = self.sim.blocks[0]
        self.g = self.st.g
        self.nodes0 = (self.st.xn.copy(), self.st.yn.copy(), self.st.zn.copy())
        self.nodes = self.nodes0
        self.sim.kin_hook = self._hook if self.frozen else None
        self.sor = None
        if tune:
            from . import tune as T

            v, planes, _ = T.autotune(block=self.g, dt=case.dt, planes=(4, 8, 16, 32))
            self.sor = {"variant": int(v), "planes": int(planes)}
        self._host = {}
        self._save(0)
        self._ckpt = {0: 0}

    # ------------------------------------------------------------ state
    def close(self):
        H.call("biocfd_snapshot_free", self.g, -1)
        self.sim.close()

    def _host_state(self):
        st, sim = self.st, self.sim
        return (sim.totime, sim.ita, st.kin.copy(), st.kin_t, st.xnode1.copy(), st.ynode1.copy(), st.znode1.copy(),
                st.blk_mv_tag, st.move_check)

    def _save(self, slot: int):
        H.call("biocfd_snapshot_save", self.g, slot)
        self._host[slot] = self._host_state()

    def _restore(self, slot: int):
        H.call("biocfd_snapshot_restore", self.g, slot)
        st, sim = self.st, self.sim
        (sim.totime, sim.ita, kin, st.kin_t, x, y, z, st.blk_mv_tag, st.move_check) = self._host[slot]
        st.kin[:] = kin
        st.xnode1[:], st.ynode1[:], st.znode1[:] = x, y, z

    def _start(self, theta, nodes):
        """The state after the set-up, with theta and the body nodes applied
        (full model: the kinematic state and, for other nodes, the set-up
        geometry's element data and normal distances recomputed)."""
        self.theta = self.theta0.copy() if theta is None else np.asarray(theta, dtype=np.float64).copy()
        new_nodes = nodes is not None
        self.nodes = self.nodes0 if nodes is None else tuple(np.ascontiguousarray(a, dtype=np.float64) for a in nodes)
        self._restore(0)
        if self.frozen:
            return
        st, sim = self.st, self.sim
        st.kin[S["a0"]] = self.theta[1]
        st.kin[S["yamp"]] = self.theta[0]
        st.kin[S["bfreq"]] = self.theta[3]
        sim.phase = float(self.theta[2])
        sim.freq = float(self.theta[3])
        st.xn, st.yn, st.zn = self.nodes
        if new_nodes:
            c = self.case
            kin = st.kin.copy()
            kin[:6] = [st.body.centre[0], st.body.centre[1], st.body.centre[2], st.body.a0_deg, self.theta[0],
                       self.theta[3]]
            H.call("biocfd_shift_surface_nodes_initial", st.nn, H.ptr(st.xn), H.ptr(st.yn), H.ptr(st.zn),
                   H.ptr(st.nid), c.aoa, c.piv_pt, 0, c.dt, c.dxmin, 0.0, H.ptr(kin), H.ptr(st.xnode1),
                   H.ptr(st.ynode1), H.ptr(st.znode1))
            H.call("biocfd_compute_surface_norm", self.g, H.ptr(st.xnode1), H.ptr(st.ynode1), H.ptr(st.znode1),
                   c.inor)
            H.call("biocfd_compute_norm_distance", self.g)

    # ------------------------------------------------------------ kinematics
    def _hook(self, st, half):
        kv = st.kin_vec().copy()
        if st.kin_t is not None:
            _embed(kin_velocities(self.theta, st.kin_t, self.case.dt), kv)
        return kv

    def _kin(self) -> np.ndarray:
        return self._hook(self.st, None) if self.frozen else self.st.kin_vec().copy()

    def _kin_jacobian(self) -> np.ndarray:
        """d kin / d theta at the current kinematic state, (BK_N, 4): the
        velocities and, for the full model, the y pivot (piv_y accumulates the
        heave increments, Search.F90:152-170)."""
        J = np.zeros((H.BK_N, N_THETA))
        t = self.st.kin_t
        if t is not None:
            jv = kin_velocities_jacobian(self.theta, t, self.case.dt)
            J[KYDOT], J[KYDDOT], J[KTD1], J[KTDD1] = jv
            J[KTD2], J[KTDD2] = -jv[2], -jv[3]
            if not self.frozen:
                J[KPIVY] = self._pivy_jac[self.sim.ita]
        return J

    def _pivy_table(self):
        """d piv_y / d theta after each step: the sum of d ychg / d theta,
        ychg = yt(t) - yt(t - dt)."""
        dt = self.case.dt
        tab = {0: np.zeros(N_THETA)}
        acc = np.zeros(N_THETA)
        t = dt
        for n in range(1, self.nsteps + 1):
            t = t + dt
            acc = acc + heave_jacobian(self.theta, t) - heave_jacobian(self.theta, t - dt)
            tab[n] = acc.copy()
        self._pivy_jac = tab

    def _node_maps(self):
        """The current node positions' Jacobians: with respect to theta (nn, 3,
        4) and to the body-frame nodes (nn, 3, 3), at the kinematic state's
        time (computeSurfaceVariables; the set-up's placement for None)."""
        st, c = self.st, self.case
        nn = st.nn
        xb, yb, zb = self.nodes
        rot = (st.nid == 51) | (st.nid == 52)
        sgn = np.where(st.nid == 52, -1.0, 1.0)
        if st.kin_t is None:
            a = np.full(nn, c.aoa * K_PI / 180.0)
            da = np.zeros(N_THETA)
            dy = np.zeros(N_THETA)
        else:
            a1, da = pitch_jacobian(self.theta, st.kin_t, c.dt)
            a = np.full(nn, a1)
            dy = heave_jacobian(self.theta, st.kin_t)
        a = np.where(rot, sgn * a, 0.0)
        ca, sa = np.cos(a), np.sin(a)
        piv = c.piv_pt
        dth = np.zeros((nn, 3, N_THETA))
        dnode = np.zeros((nn, 3, 3))
        dnode[:, 0, 0] = 1.0
        # zr = z cos a + y sin a + piv - piv cos a ; yr = -z sin a + y cos a + piv sin a
        dzr_da = -zb * sa + yb * ca + piv * sa
        dyr_da = -zb * ca - yb * sa + piv * ca
        dnode[:, 1, 1] = np.where(rot, ca, 1.0)
        dnode[:, 1, 2] = np.where(rot, -sa, 0.0)
        dnode[:, 2, 1] = np.where(rot, sa, 0.0)
        dnode[:, 2, 2] = np.where(rot, ca, 1.0)
        da_n = np.where(rot, sgn, 0.0)[:, None] * da[None, :]
        dth[:, 1, :] = dyr_da[:, None] * da_n + dy[None, :]
        dth[:, 2, :] = dzr_da[:, None] * da_n
        return dth, dnode

    def _node_tangent(self, dtheta, dnodes) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        dth, dnode = self._node_maps()
        d = dth @ np.asarray(dtheta, dtype=np.float64)
        if dnodes is not None:
            d = d + np.einsum("nij,jn->ni", dnode, np.stack([np.asarray(x, dtype=np.float64) for x in dnodes]))
        return (np.ascontiguousarray(d[:, 0]), np.ascontiguousarray(d[:, 1]), np.ascontiguousarray(d[:, 2]))

    def force_x(self, kv: Optional[np.ndarray] = None) -> float:
        out = np.zeros(4)
        kv = self._kin() if kv is None else kv
        H.call("biocfd_stress_sums", self.g, self.case.mu_f, self.case.rho_f, H.ptr(kv), H.ptr(out))
        return float(out[0] + out[1])

    def _in_window(self, n: int) -> bool:
        return n > self.nsteps - self.m

    def discrete_state(self) -> bytes:
        """The current step's discrete structure: cell tags, the IB cell list
        and every IB cell's seven nearest elements (the decisions the
        derivative holds fixed)."""
        import hashlib

        g, st = self.g, self.st
        cnt = self.sim.counts(g)
        h = hashlib.sha256(self.sim.tag(g, "cell").tobytes())
        n = cnt["ib"]
        h.update(self.sim.index_list(g, "interceptedIndexPtr", n).tobytes())
        for q in range(len(H.BN_NAMES)):
            a = np.zeros(n, dtype=np.int64)
            H.call("biocfd_nel_download", g, q, H.ptr(a), n)
            h.update(a.tobytes())
        return h.digest()

    # ------------------------------------------------------------ passes
    def forward(self, theta=None, nodes=None, checkpoints: bool = False, states: Optional[list] = None) -> float:
        """J(theta, nodes); with checkpoints, the block state every `every`
        steps is kept for backward(); states (a list) receives each step's
        discrete_state()."""
        self._start(theta, nodes)
        if checkpoints:
            self._save(1)   # the started state (theta and the nodes applied)
            self._ckpt = {0: 1}
        J = 0.0
        for n in range(1, self.nsteps + 1):
            self.sim.step()
            if states is not None:
                states.append(self.discrete_state())
            if self._in_window(n):
                J += self.force_x() / self.m
            if checkpoints and n % self.every == 0 and n < self.nsteps:
                slot = 2 + n // self.every
                self._save(slot)
                self._ckpt[n] = slot
        self._fwd = (self.theta.copy(), self.nodes)
        return J

    def tangent(self, dtheta, dnodes=None, theta=None) -> Tuple[float, float]:
        """(J, dJ) along (dtheta, dnodes) by tangent-linear time steps next to
        the forward."""
        c = self.case
        self._start(theta, None)
        if not self.frozen:
            self._pivy_table()
        H.call("biocfd_ad_reset", self.g)
        geo = not self.frozen
        rec = 0 if geo else -1
        dtheta = np.asarray(dtheta, dtype=np.float64)
        J = dJ = 0.0
        for n in range(1, self.nsteps + 1):
            kv_a, kd_a = self._kin(), self._kin_jacobian() @ dtheta
            if geo:
                nt_a = self._node_tangent(dtheta, dnodes)
                self._save(PRE_SLOT)
                H.call("biocfd_ad_record", self.g, 0)
                self.sim.step()
                H.call("biocfd_ad_record", self.g, -1)
                self._save(POST_SLOT)
                self._restore(PRE_SLOT)
                H.call("biocfd_ad_geometry_tangent", self.g, float(c.inor), *(H.ptr(x) for x in nt_a))
                H.call("biocfd_tlm_step_start", self.g, self.n_iter, c.dt, c.re, c.uc, H.ptr(kv_a), H.ptr(kd_a), rec)
                self._restore(POST_SLOT)
                nt_b = self._node_tangent(dtheta, dnodes)
                H.call("biocfd_ad_geometry_tangent", self.g, float(c.inor), *(H.ptr(x) for x in nt_b))
            else:
                H.call("biocfd_tlm_step_start", self.g, self.n_iter, c.dt, c.re, c.uc, H.ptr(kv_a), H.ptr(kd_a), rec)
                self.sim.step()
            kv, kd = self._kin(), self._kin_jacobian() @ dtheta
            H.call("biocfd_tlm_step_end", self.g, H.ptr(kv), H.ptr(kd), rec)
            if self._in_window(n):
                out = np.zeros(4)
                H.call("biocfd_tlm_stress", self.g, c.mu_f, c.rho_f, H.ptr(kv), H.ptr(kd), int(geo), H.ptr(out))
                J += self.force_x(kv) / self.m
                dJ += (out[0] + out[1]) / self.m
        return J, dJ

    def _geometry_adjoint(self, gtheta, gnodes):
        """The element adjoints of the current geometry into theta and the
        body-frame nodes (and cleared)."""
        nn = self.st.nn
        xb, yb, zb = np.zeros(nn), np.zeros(nn), np.zeros(nn)
        H.call("biocfd_ad_geometry_adjoint", self.g, float(self.case.inor), H.ptr(xb), H.ptr(yb), H.ptr(zb), 1)
        X = np.stack([xb, yb, zb], axis=1)
        dth, dnode = self._node_maps()
        gtheta += np.einsum("nij,ni->j", dth, X)
        gnodes += np.einsum("nij,ni->nj", dnode, X)

    def backward(self):
        """The gradient at the last forward(checkpoints=True): dJ/dtheta, and
        for the full model also dJ/d(body-frame nodes) (nn, 3), by the
        adjoint time steps in reverse, each segment replayed from its
        checkpoint."""
        c = self.case
        self.theta, self.nodes = self._fwd[0].copy(), self._fwd[1]
        if not self.frozen:
            st, sim = self.st, self.sim
            sim.phase, sim.freq = float(self.theta[2]), float(self.theta[3])
            st.xn, st.yn, st.zn = self.nodes
        geo = not self.frozen
        if geo:
            self._pivy_table()
        H.call("biocfd_ad_reset", self.g)
        grad = np.zeros(N_THETA)
        gnodes = np.zeros((self.st.nn, 3))
        kin_bar = np.zeros(H.BK_N)
        sbar = np.array([1.0 / self.m, 1.0 / self.m, 0.0, 0.0])
        starts = sorted(self._ckpt)
        for s0 in reversed(starts):
            s1 = min(s0 + self.every, self.nsteps)
            self._restore(self._ckpt[s0])
            self._save(SEG_SLOT)
            for n in range(s0 + 1, s1 + 1):
                if geo:
                    H.call("biocfd_ad_record", self.g, n - s0)
                self.sim.step()
                self._save(SEG_SLOT + n - s0)
            if geo:
                H.call("biocfd_ad_record", self.g, -1)
            for n in range(s1, s0, -1):
                rec = n - s0 if geo else -1
                self._restore(SEG_SLOT + n - s0)
                kv = self._kin()
                if self._in_window(n):
                    H.call("biocfd_adj_stress", self.g, c.mu_f, c.rho_f, H.ptr(kv), H.ptr(sbar), int(geo))
                H.call("biocfd_adj_step_end", self.g, H.ptr(kv), rec)
                H.call("biocfd_ad_kin", self.g, H.ptr(kin_bar), 1)
                grad += self._kin_jacobian().T @ kin_bar
                if geo:
                    self._geometry_adjoint(grad, gnodes)
                self._restore(SEG_SLOT + n - 1 - s0)
                kv = self._kin()
                H.call("biocfd_adj_step_start", self.g, self.n_iter, c.dt, c.re, c.uc, H.ptr(kv), rec)
                H.call("biocfd_ad_kin", self.g, H.ptr(kin_bar), 1)
                grad += self._kin_jacobian().T @ kin_bar
        if geo:
            # the set-up geometry, used by the first step's start
            self._geometry_adjoint(grad, gnodes)
            return grad, gnodes
        return grad

    def gradient(self, theta=None):
        J = self.forward(theta, checkpoints=True)
        return J, self.backward()
