"""Tangent-linear and adjoint time steps of a single block with a moving body
over the C ABI (include/biocfd_hip.h: biocfd_tlm_*, biocfd_adj_*,
biocfd_snapshot_*), with step checkpointing and recomputation.

The function differentiated is

    J(theta) = mean over the last M steps of F_x,

F_x the body's x-force after a step: the viscous and pressure x sums of
stress_calculation (StressCalc.f90:173-176).  The kinematic parameters
theta = (heave amplitude, pitch amplitude [rad], phase [rad], frequency of
heave and pitch) enter through the body velocities and accelerations that
computeSurfaceVariables derives (Search.F90:141-165) and the forcing routines
and the force sums read.  The body's positions, and with them every step's
IB geometry (tags, lists, nearest elements, distances, normals), are those of
the base parameters ("frozen geometry"): the derivative is that of the
flow's response to the body's motion through its boundary values, not of the
moving geometry itself.  The SOR count is fixed (epsi 0: pcItaMax + 1
iterations per solve), so every step is differentiable.

Checkpoints: the forward pass keeps the block's state every `every` steps
(biocfd_snapshot_save); the backward pass replays each segment from its
checkpoint, keeping every step of it, and runs the adjoint steps backwards
through it.
"""
from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

from . import cases as CS
from . import hip as H
from .driver import Simulation

# kinematic vector slots (enum biocfd_kin)
KX, KYDOT, KYDDOT, KTD1, KTDD1, KTD2, KTDD2 = 0, 1, 2, 3, 4, 5, 6
K_PI = 3.141592653589793115997963468544185161590576171875   # the library's pi
N_THETA = 4
THETA_NAMES = ("heave amplitude", "pitch amplitude", "phase", "frequency")
SEG_SLOT = 2048   # first snapshot slot of the replayed segment


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


def _embed(v4: np.ndarray, kv: np.ndarray) -> None:
    kv[KYDOT], kv[KYDDOT], kv[KTD1], kv[KTDD1] = v4
    kv[KTD2], kv[KTDD2] = -v4[2], -v4[3]


class AdjointRun:
    """The adjoint workload on the GPU: forward, tangent-linear and adjoint
    passes of `case` (one block with a body, epsi 0)."""

    def __init__(self, case: CS.Case, every: int = 8, avg_steps: Optional[int] = None):
        if len(case.blocks) != 1 or case.blocks[0].body is None:
            raise ValueError("the adjoint driver runs a single block with a body")
        if case.epsi != 0.0:
            raise ValueError("the adjoint driver needs a fixed SOR count (epsi = 0)")
        body = case.blocks[0].body
        if body.bfreq != case.freq:
            raise ValueError("heave and pitch frequencies must be equal (one frequency parameter)")
        self.case = case
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
        self.sim.kin_hook = self._hook
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

    # ------------------------------------------------------------ kinematics
    def _hook(self, st, half):
        kv = st.kin_vec().copy()
        if st.kin_t is not None:
            _embed(kin_velocities(self.theta, st.kin_t, self.case.dt), kv)
        return kv

    def _kin(self) -> np.ndarray:
        return self._hook(self.st, None)

    def _kin_dot(self, dtheta) -> np.ndarray:
        kd = np.zeros(H.BK_N)
        if self.st.kin_t is not None:
            _embed(kin_velocities_jacobian(self.theta, self.st.kin_t, self.case.dt) @ np.asarray(dtheta, float), kd)
        return kd

    def _theta_bar(self, kin_bar: np.ndarray) -> np.ndarray:
        if self.st.kin_t is None:
            return np.zeros(N_THETA)
        v = np.array([kin_bar[KYDOT], kin_bar[KYDDOT], kin_bar[KTD1] - kin_bar[KTD2], kin_bar[KTDD1] - kin_bar[KTDD2]])
        return kin_velocities_jacobian(self.theta, self.st.kin_t, self.case.dt).T @ v

    def force_x(self, kv: Optional[np.ndarray] = None) -> float:
        out = np.zeros(4)
        kv = self._kin() if kv is None else kv
        H.call("biocfd_stress_sums", self.g, self.case.mu_f, self.case.rho_f, H.ptr(kv), H.ptr(out))
        return float(out[0] + out[1])

    def _in_window(self, n: int) -> bool:
        return n > self.nsteps - self.m

    # ------------------------------------------------------------ passes
    def forward(self, theta=None, checkpoints: bool = False) -> float:
        """J(theta); with checkpoints, the block state every `every` steps is
        kept for backward()."""
        self.theta = self.theta0.copy() if theta is None else np.asarray(theta, dtype=np.float64).copy()
        self._restore(0)
        if checkpoints:
            self._ckpt = {0: 0}
        J = 0.0
        for n in range(1, self.nsteps + 1):
            self.sim.step()
            if self._in_window(n):
                J += self.force_x() / self.m
            if checkpoints and n % self.every == 0 and n < self.nsteps:
                slot = 1 + n // self.every
                self._save(slot)
                self._ckpt[n] = slot
        self._theta_forward = self.theta.copy()
        return J

    def tangent(self, dtheta, theta=None) -> Tuple[float, float]:
        """(J, dJ . dtheta) by the tangent-linear time steps alongside the forward."""
        self.theta = self.theta0.copy() if theta is None else np.asarray(theta, dtype=np.float64).copy()
        c = self.case
        self._restore(0)
        H.call("biocfd_ad_reset", self.g)
        J = dJ = 0.0
        for n in range(1, self.nsteps + 1):
            kv, kd = self._kin(), self._kin_dot(dtheta)
            H.call("biocfd_tlm_step_start", self.g, self.n_iter, c.dt, c.re, c.uc, H.ptr(kv), H.ptr(kd))
            self.sim.step()
            kv, kd = self._kin(), self._kin_dot(dtheta)
            H.call("biocfd_tlm_step_end", self.g, H.ptr(kv), H.ptr(kd))
            if self._in_window(n):
                out = np.zeros(4)
                H.call("biocfd_tlm_stress", self.g, c.mu_f, c.rho_f, H.ptr(kv), H.ptr(kd), H.ptr(out))
                J += self.force_x(kv) / self.m
                dJ += (out[0] + out[1]) / self.m
        return J, dJ

    def backward(self) -> np.ndarray:
        """dJ/dtheta at the last forward(checkpoints=True): the adjoint time
        steps in reverse, each segment replayed from its checkpoint."""
        c = self.case
        self.theta = self._theta_forward.copy()
        H.call("biocfd_ad_reset", self.g)
        grad = np.zeros(N_THETA)
        kin_bar = np.zeros(H.BK_N)
        sbar = np.array([1.0 / self.m, 1.0 / self.m, 0.0, 0.0])
        starts = sorted(self._ckpt)
        for s0 in reversed(starts):
            s1 = min(s0 + self.every, self.nsteps)
            self._restore(self._ckpt[s0])
            self._save(SEG_SLOT)
            for n in range(s0 + 1, s1 + 1):
                self.sim.step()
                self._save(SEG_SLOT + n - s0)
            for n in range(s1, s0, -1):
                self._restore(SEG_SLOT + n - s0)
                kv = self._kin()
                if self._in_window(n):
                    H.call("biocfd_adj_stress", self.g, c.mu_f, c.rho_f, H.ptr(kv), H.ptr(sbar))
                H.call("biocfd_adj_step_end", self.g, H.ptr(kv))
                H.call("biocfd_ad_kin", self.g, H.ptr(kin_bar), 1)
                grad += self._theta_bar(kin_bar)
                self._restore(SEG_SLOT + n - 1 - s0)
                kv = self._kin()
                H.call("biocfd_adj_step_start", self.g, self.n_iter, c.dt, c.re, c.uc, H.ptr(kv))
                H.call("biocfd_ad_kin", self.g, H.ptr(kin_bar), 1)
                grad += self._theta_bar(kin_bar)
        return grad

    def gradient(self, theta=None) -> Tuple[float, np.ndarray]:
        J = self.forward(theta, checkpoints=True)
        return J, self.backward()
