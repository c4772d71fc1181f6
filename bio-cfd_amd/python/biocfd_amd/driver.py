"""Time-step driver over the C ABI, in the call order of the reference driver
(app/main.F90:43-265): set-up, then per step momentum + boundary conditions +
IB velocity forcing, poissonSolver, pressure forcing, body kinematics and the
IB geometry/tag/list refresh.

Single-block-with-body cases run the IB routines on block 1 together with the
domain boundary conditions, as the oracle's harness does.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional

import numpy as np

from . import cases as CS
from . import hip as H

# kinematic state slots (include/biocfd_hip.h, biocfd_shift_surface_nodes_initial)
S = dict(xshift=0, yshift=1, zshift=2, a0=3, yamp=4, bfreq=5, xpth1=6, xpth2=7, ypth2=8, xchg=9,
         ychg=10, yt=11, ydot=12, yddot=13, xt=14, xdot=15, thetaDot=16, thetaDDot=17, thetaDot1=18,
         thetaDDot1=19, thetaDot2=20, thetaDDot2=21, alphaDot=22, piv_x=23, piv_y=24, piv_z=25,
         xmove=26, ymove=27, zmove=28)


class BlockState:
    def __init__(self, g: int, spec: CS.BlockSpec):
        self.g = g
        self.spec = spec
        self.n = tuple(int(v) for v in spec.n)
        self.grids = spec.grids()
        self.kin = np.zeros(H.BKS_N)
        self.body = spec.body
        self.xnode1 = self.ynode1 = self.znode1 = None
        # block move: marker nodes of findDistnode, move_check, blk_mv_tag
        self.markers = np.zeros(3, dtype=np.int64)
        self.move_check = 0
        self.blk_mv_tag = 0
        # the time the kinematic state was last computed at (None: the set-up's)
        self.kin_t = None

    @property
    def shape(self):
        nx, ny, nz = self.n
        return (nx + 2, ny + 2, nz + 2)

    def kin_vec(self) -> np.ndarray:
        k = self.kin
        return np.array([k[S["xdot"]], k[S["ydot"]], k[S["yddot"]], k[S["thetaDot1"]],
                         k[S["thetaDDot1"]], k[S["thetaDot2"]], k[S["thetaDDot2"]],
                         k[S["alphaDot"]], k[S["piv_x"]], k[S["piv_y"]], k[S["piv_z"]]])


def grid_arrays(st: BlockState) -> List[np.ndarray]:
    gx, gy, gz = st.grids
    return [gx.x1, gy.x1, gz.x1, gx.delta, gy.delta, gz.delta, gx.x1, gy.xc, gz.xc,
            gx.xc, gy.x1, gz.xc, gx.xc, gy.xc, gz.x1, gx.xc, gy.xc, gz.xc]


class Simulation:
    """One case on the GPU; the device holds all state between calls."""

    def __init__(self, case: CS.Case, device: int = 0, precision: str = "fp64", threads: int = 1):
        """precision: "fp64" (the reference's method) or "mixed" (FP32 SOR
        smoothing inside FP64 defect correction, include/biocfd_hip.h).
        threads > 1 runs the per-block loops of a time step on that many host
        threads, as the reference's OpenMP loops over blocks do
        (main.F90:133-148, 232-263); the library serialises the calls."""
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', not {precision!r}")
        self.case = case
        self.precision = precision
        self.threads = max(1, int(threads))
        self._pool = None
        if self.threads > 1:
            from concurrent.futures import ThreadPoolExecutor

            self._pool = ThreadPoolExecutor(max_workers=self.threads)
        H.call("biocfd_init", device)
        H.call("biocfd_set_precision", 0 if precision == "fp64" else 1)
        self.blocks = [BlockState(g + 1, spec) for g, spec in enumerate(case.blocks)]
        # kin_hook(st, half) -> kinematic vector of the forcing calls ("start":
        # momentum stage and pressureForcing1, "end": after the body moved);
        # None: st.kin_vec() (the adjoint driver substitutes velocities)
        self.kin_hook = None
        self.totime = 0.0
        self.ita = 0
        self.phase = case.phase_angle_deg * math.pi / 180.0
        self.freq = case.freq
        self._setup()

    # ---------------------------------------------------------------- set-up
    def _setup(self):
        c = self.case
        for st in self.blocks:
            s = st.spec
            gx, gy, gz = st.grids
            H.call("biocfd_block_create", st.g, st.n[0], st.n[1], st.n[2], s.omega, s.cintp,
                   float(gx.delta[1]), float(gy.delta[1]), float(gz.delta[1]))
            for w, a in enumerate(grid_arrays(st)):
                a = np.ascontiguousarray(a, dtype=np.float64)
                H.call("biocfd_grid_upload", st.g, w, H.ptr(a), a.size)
        for q, itf in enumerate(c.interfaces):
            ext = np.array([itf.lo[0], itf.hi[0], itf.lo[1], itf.hi[1], itf.lo[2], itf.hi[2]], dtype=np.float64)
            H.call("biocfd_interface_create", q + 1, itf.a_blk, itf.b_blk, itf.a_msh, itf.b_msh, H.ptr(ext))
        for st in self.blocks:
            if st.body is None:
                continue
            b = st.body.body
            self._body_arrays(st, b)
            box = np.array(CS.search_box(st.spec, st.grids), dtype=np.int64)
            H.call("biocfd_body_create", st.g, b.nodes.shape[0], b.elems.shape[0], H.ptr(st.e1), H.ptr(st.e2),
                   H.ptr(st.e3), H.ptr(st.sid), H.ptr(box))
            H.call("biocfd_find_dist_node", st.nn, H.ptr(st.xn), H.ptr(st.zn), H.ptr(st.nid), H.ptr(st.markers))
            st.kin[:6] = [st.body.centre[0], st.body.centre[1], st.body.centre[2], st.body.a0_deg,
                          st.body.yamp, st.body.bfreq]
            H.call("biocfd_shift_surface_nodes_initial", st.nn, H.ptr(st.xn), H.ptr(st.yn), H.ptr(st.zn),
                   H.ptr(st.nid), c.aoa, c.piv_pt, 0, c.dt, c.dxmin, 0.0, H.ptr(st.kin), H.ptr(st.xnode1),
                   H.ptr(st.ynode1), H.ptr(st.znode1))
            H.call("biocfd_compute_surface_norm", st.g, H.ptr(st.xnode1), H.ptr(st.ynode1), H.ptr(st.znode1),
                   c.inor)
        for st in self.blocks:
            if st.body is not None:
                H.call("biocfd_tagging", st.g)
                H.call("biocfd_cell_count_solid", st.g)
        if self.blocks[0].body is None:
            H.call("biocfd_fine_block_cell", 1)
            H.call("biocfd_cell_count_solid_coarse", 1)
        for st in self.blocks[1:]:
            if st.body is None:
                # a fine block without a body: every interior cell fluid, the
                # lists as for the coarse block
                H.call("biocfd_cell_count_solid_coarse", st.g)
        for st in self.blocks:
            self._initial_fields(st)
        for st in self.blocks:
            if st.body is not None:
                H.call("biocfd_compute_norm_distance", st.g)
                H.call("biocfd_find_ts_cells", st.g)
            H.call("biocfd_coefficient_matrix", st.g, 1 if st.g == 1 else 0)
        for st in self.blocks:
            H.call("biocfd_non_uni_coeff", st.g)
        self.totime = c.dt

    def _body_arrays(self, st: BlockState, b: CS.Body):
        st.nn = b.nodes.shape[0]
        st.xn = np.ascontiguousarray(b.nodes[:, 0])
        st.yn = np.ascontiguousarray(b.nodes[:, 1])
        st.zn = np.ascontiguousarray(b.nodes[:, 2])
        st.nid = np.ascontiguousarray(b.node_id, dtype=np.int64)
        st.e1 = np.ascontiguousarray(b.elems[:, 0], dtype=np.int64)
        st.e2 = np.ascontiguousarray(b.elems[:, 1], dtype=np.int64)
        st.e3 = np.ascontiguousarray(b.elems[:, 2], dtype=np.int64)
        st.sid = np.ascontiguousarray(b.surf_id, dtype=np.int64)
        st.xnode1 = np.zeros(st.nn)
        st.ynode1 = np.zeros(st.nn)
        st.znode1 = np.zeros(st.nn)

    def _initial_fields(self, st: BlockState):
        """initialConditions (InitialConditions.f90:10-36) plus the seeded noise
        on fluid cells, computed on the host and uploaded."""
        c = self.case
        shp = st.shape
        cell = self.tag(st.g, "cell")
        u = np.full(shp, c.uc, order="F")
        v = np.zeros(shp, order="F")
        w = np.zeros(shp, order="F")
        p = np.zeros(shp, order="F")
        if c.noise_amp != 0.0:
            fl = np.zeros(shp, dtype=bool, order="F")
            fl[1:-1, 1:-1, 1:-1] = cell[1:-1, 1:-1, 1:-1] == 0
            lin = np.arange(u.size, dtype=np.int64).reshape(shp, order="F")[fl]
            seed = c.seed + 1000 * st.g
            u[fl] = u[fl] + c.noise_amp * (2.0 * CS.noise_uniform(seed, 1, lin) - 1.0)
            v[fl] = v[fl] + c.noise_amp * (2.0 * CS.noise_uniform(seed, 2, lin) - 1.0)
            w[fl] = w[fl] + c.noise_amp * (2.0 * CS.noise_uniform(seed, 3, lin) - 1.0)
        for name, a in (("u", u), ("ut", u), ("v", v), ("vt", v), ("w", w), ("wt", w), ("p", p)):
            self.upload(st.g, name, a)
        zero_pr = np.zeros(tuple(s + 1 for s in shp), dtype=np.int64, order="F")
        H.call("biocfd_tag_upload", st.g, H.BT["cell_pr"], H.ptr(zero_pr))

    # ---------------------------------------------------------------- time step
    def _each(self, fn, bodies_only: bool = False):
        """fn(st) for every block (with a body), one after the other or on the
        thread pool; exceptions propagate."""
        blocks = [st for st in self.blocks if st.body is not None] if bodies_only else self.blocks
        if self._pool is None or len(blocks) < 2:
            for st in blocks:
                fn(st)
        else:
            list(self._pool.map(fn, blocks))

    def kin(self, st: BlockState, half: str) -> np.ndarray:
        return st.kin_vec() if self.kin_hook is None else self.kin_hook(st, half)

    def step(self):
        c = self.case
        self.ita += 1
        self.totime = self.totime + c.dt

        def momentum(st):
            # nsMomentum2order itself ran for every block above
            if st.g == 1:
                H.call("biocfd_velocity_bc", st.g, c.dt, c.uc)
            if st.body is not None:
                H.call("biocfd_solid_cell_bc", st.g)
                kv = self.kin(st, "start")
                H.call("biocfd_velocity_forcing1", st.g, H.ptr(kv))
            if st.g == 1:
                H.call("biocfd_velocity_bc", st.g, c.dt, c.uc)

        def pressure_forcing(st):
            kv = self.kin(st, "start")
            H.call("biocfd_pressure_forcing1", st.g, H.ptr(kv))

        def surface_variables(st):
            st.blk_mv_tag = 0
            H.call("biocfd_compute_surface_variables", st.nn, H.ptr(st.xn), H.ptr(st.yn), H.ptr(st.zn),
                   H.ptr(st.nid), self.phase, c.piv_pt, self.ita, c.dt, c.dxmin, self.freq, self.totime,
                   H.ptr(st.kin), H.ptr(st.xnode1), H.ptr(st.ynode1), H.ptr(st.znode1))
            st.kin_t = self.totime

        def surface_norm(st):
            H.call("biocfd_compute_surface_norm", st.g, H.ptr(st.xnode1), H.ptr(st.ynode1), H.ptr(st.znode1),
                   c.inor)

        def tagging_move(st):
            H.call("biocfd_tagging_move", st.g, st.move_check)

        def retag_and_force(st):
            g = st.g
            kv = self.kin(st, "end")
            H.call("biocfd_selective_retagging", g, st.blk_mv_tag)
            st.blk_mv_tag = 0
            H.call("biocfd_cell_count_solid", g)
            H.call("biocfd_solid_cell_bc_move", g, st.move_check)
            H.call("biocfd_update_velocity_newv", g, st.move_check)
            st.move_check = 0
            H.call("biocfd_compute_norm_distance", g)
            H.call("biocfd_find_ts_cells", g)
            H.call("biocfd_velocity_forcing_field", g, H.ptr(kv))
            H.call("biocfd_pressure_forcing_field", g, H.ptr(kv))
            H.call("biocfd_velocity_forcing_ghost", g, H.ptr(kv))
            H.call("biocfd_pressure_forcing_ghost", g, H.ptr(kv))

        # the momentum step of every block in one call (one dense launch for
        # all blocks; each block's step reads and writes only its own arrays,
        # so the reference's per-block order is kept by the blocks' BCs and
        # forcing that follow)
        H.call("biocfd_ns_momentum_set", 1, len(self.blocks), c.dt, c.re)
        self._each(momentum)
        H.call("biocfd_poisson_solver", 1, len(self.blocks), c.pc_ita_max, c.epsi, c.dt, c.uc)
        self._each(pressure_forcing, bodies_only=True)
        self._each(surface_variables, bodies_only=True)
        self._block_move()
        self._each(surface_norm, bodies_only=True)
        self._each(tagging_move, bodies_only=True)
        self._each(retag_and_force, bodies_only=True)

    def _block_move(self):
        """block_move_check and the moves it decides (main.F90:178-203).  The
        move check reads the marker nodes findDistnode sets only from ID-51
        nodes, so it runs only where every fine block has them."""
        c = self.case
        flcnt = np.zeros(1, dtype=np.int64)
        fine = [self.blocks[itf.b_blk - 1] for itf in c.interfaces]
        if fine and all(np.all(st.markers > 0) for st in fine):
            moved = np.zeros(1, dtype=np.int64)
            for q, st in enumerate(fine, start=1):
                H.call("biocfd_block_move_check", q, H.ptr(st.xnode1), H.ptr(st.ynode1), H.ptr(st.markers),
                       H.ptr(st.kin), H.ptr(moved), H.ptr(flcnt))
                if moved[0] == 1:
                    st.move_check = 1
                    st.blk_mv_tag = 1
        for st in self.blocks[1:]:
            H.call("biocfd_change_block_coords", st.g, st.move_check)
        for q, st in enumerate(fine, start=1):
            H.call("biocfd_change_block_coords_interfaces", q, st.move_check)
        for q, st in enumerate(fine, start=1):
            H.call("biocfd_change_block_interface", q, st.move_check)
        if self.blocks[0].body is None:
            # the coarse block's interface masks and lists (main.F90:195-200)
            H.call("biocfd_fine_block_cell", 1)
            H.call("biocfd_cell_count_solid_coarse_mv", 1, H.ptr(flcnt))

    # ---------------------------------------------------------------- data access
    def upload(self, g: int, name: str, a: np.ndarray):
        a = np.asfortranarray(a, dtype=np.float64)
        H.call("biocfd_field_upload", g, H.BF[name], H.ptr(a))

    def field(self, g: int, name: str) -> np.ndarray:
        st = self.blocks[g - 1]
        a = np.zeros(st.shape, order="F")
        H.call("biocfd_field_download", g, H.BF[name], H.ptr(a))
        return a

    def tag(self, g: int, name: str) -> np.ndarray:
        st = self.blocks[g - 1]
        shp = st.shape if name in ("cell", "cell_n") else tuple(s + 1 for s in st.shape)
        a = np.zeros(shp, dtype=np.int64, order="F")
        H.call("biocfd_tag_download", g, H.BT[name], H.ptr(a))
        return a

    def counts(self, g: int) -> Dict[str, int]:
        a = np.zeros(6, dtype=np.int64)
        H.call("biocfd_counts", g, H.ptr(a))
        return {k: int(a[v]) for k, v in H.BC.items()}

    def dump(self) -> List[Dict[str, np.ndarray]]:
        """Every array the oracle dumps, keyed like the oracle's records."""
        out = []
        for st in self.blocks:
            g = st.g
            d: Dict[str, np.ndarray] = {}
            for name in ("u", "v", "w", "p", "ut", "vt", "wt", "pc", "pco", "b"):
                d[name] = self.field(g, name)
            for name in ("cell", "cell2", "nodeIdTag", "cell_n", "cell_pr"):
                d[name] = self.tag(g, name)
            for name, w, n in (("x1", 0, st.n[0] + 3), ("y1", 1, st.n[1] + 3), ("z1", 2, st.n[2] + 3)):
                a = np.zeros(n)
                H.call("biocfd_grid_download", g, w, H.ptr(a), n)
                d[name] = a
            it = np.zeros(1, dtype=np.int64)
            e2 = np.zeros(1)
            es = np.zeros(1)
            H.call("biocfd_sor_state", g, H.ptr(it), H.ptr(e2), H.ptr(es))
            d["nIterPcor"] = np.int64(it[0])
            d["derr2"] = np.float64(e2[0])
            d["derrStdSt"] = np.float64(es[0])
            cnt = self.counts(g)
            d["fluidCellCount"] = np.int64(cnt["fluid"])
            d["solidCellCount"] = np.int64(cnt["solid"])
            d["ibCellCount"] = np.int64(cnt["ib"])
            d["TSCellCount"] = np.int64(cnt["ts"])
            nred = cnt["red"]
            nblack = cnt["black"]
            d["redCellCount"] = np.int64(nred)
            d["blackCellCount"] = np.int64(nblack)
            for name, n in (("fluidIndexPtr", cnt["fluid"]), ("redCellIndexPtr", nred),
                            ("blackCellIndexPtr", nblack)):
                d[name] = self.index_list(g, name, n)
            nx, ny, nz = st.n
            for ax, (nm, n) in enumerate((("Acx", nx), ("Acy", ny), ("Acz", nz))):
                a = np.zeros((n, 3), order="F")
                H.call("biocfd_coef_download", g, ax, H.ptr(a), n)
                d[nm] = a
            ncols = [nx, ny, nz, nx, nx, ny, ny, nz, nz]
            for t, nm in enumerate(H.BTB_NAMES):
                for is_ck, pre in ((0, "ca_"), (1, "ck_")):
                    a = np.zeros((6, ncols[t] + 2), order="F")
                    H.call("biocfd_table_download", g, t, is_ck, H.ptr(a), ncols[t] + 2)
                    d[pre + nm] = a
            if st.body is not None:
                d["interceptedIndexPtr"] = self.index_list(g, "interceptedIndexPtr", cnt["ib"])
                d["solidIndexPtr"] = self.index_list(g, "solidIndexPtr", cnt["solid"])
                d["TSIndexPtr"] = self.index_list(g, "TSIndexPtr", cnt["ts"])
                a = np.zeros(cnt["ts"], dtype=np.int64)
                H.call("biocfd_index_ts_download", g, H.ptr(a), cnt["ts"])
                d["index_ts"] = a
                for q, nm in enumerate(H.BN_NAMES):
                    a = np.zeros(cnt["ib"], dtype=np.int64)
                    H.call("biocfd_nel_download", g, q, H.ptr(a), cnt["ib"])
                    d[nm] = a
                for q, nm in enumerate(H.BND_NAMES):
                    a = np.zeros(cnt["ib"])
                    H.call("biocfd_normdis_download", g, q, H.ptr(a), cnt["ib"])
                    d[nm] = a
                ne = st.body.body.elems.shape[0]
                for q, nm in enumerate(H.BE_NAMES):
                    a = np.zeros(ne)
                    H.call("biocfd_elem_download", g, q, H.ptr(a), ne)
                    d[nm] = a
                for q, nm in enumerate(H.BGH_NAMES):
                    a = np.zeros(cnt["ts"])
                    H.call("biocfd_ghost_download", g, q, H.ptr(a), cnt["ts"])
                    d[nm] = a
                d["xnode1"], d["ynode1"], d["znode1"] = st.xnode1.copy(), st.ynode1.copy(), st.znode1.copy()
                k = st.kin
                for nm in ("xdot", "ydot", "yddot", "thetaDot1", "thetaDDot1", "piv_x", "piv_y", "piv_z",
                           "xpth2", "ymove"):
                    d[nm] = np.float64(k[S[nm]])
                res = np.zeros(4)
                kv = st.kin_vec()
                H.call("biocfd_stress_calculation", g, self.case.mu_f, self.case.rho_f, H.ptr(kv), H.ptr(res))
                d["viscous_drag"], d["pressure_drag"], d["viscous_lift"], d["pressure_lift"] = (
                    np.float64(x) for x in res)
            out.append(d)
        return out

    def index_list(self, g: int, name: str, n: int) -> np.ndarray:
        a = np.zeros((n, 3), dtype=np.int64, order="F")
        H.call("biocfd_list_download", g, H.BL[name], H.ptr(a), n)
        return a

    def residual_norm(self, g: int = 1) -> float:
        """max |b/dt - A pc| over the fluid cells of block g (FP64)."""
        out = np.zeros(1)
        H.call("biocfd_residual_norm", g, self.case.dt, H.ptr(out))
        return float(out[0])

    def close(self):
        if self._pool is not None:
            self._pool.shutdown()
            self._pool = None
        H.call("biocfd_set_precision", 0)
        H.call("biocfd_finalize")
