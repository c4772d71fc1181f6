"""A case as the reference's own input deck (the files readInput,
readBlockInterface and readSurfaceMeshGmsh read, ReadInput.f90:25-456).

The deck carries what the reference's main program can express: block 1 is
the coarse block without a body (main.F90 reads a surface mesh for blocks
g /= 1 only, :63-68), every body block carries the one mesh
`geometries/butterflyMedium.msh` (ReadInput.f90:370), lengths are in
millimetres (scaled by 0.001 on reading), the viscosity in units of 1e-5, the
time step follows from the flapping frequency, dt = 1/(4 freq dt_order)
(:157), and the heave amplitude is counted in dxmin (:178).  The values are
written in shortest round-trip decimal, so a reader parses the exact doubles
written; what the reference derives from them (u0, re, deltat, the grids) it
computes itself, and the driver that reads this deck repeats that arithmetic.

    write_deck(case, directory, dxmin_mm=..., freq=...)

produces input_data.nml, body_search.dat, grid_shift.dat, shift.dat,
flap_amp.dat, butter_move.dat, block_details.dat, interface_details.dat,
the grid files {x,y,z}grid_bk<g:03d>_<n+1>.txt and
geometries/butterflyMedium.msh (Gmsh 2.2 ASCII).
"""
from __future__ import annotations

import os
from typing import List, Optional

import numpy as np

from . import cases as CS


def _f(v: float) -> str:
    """Shortest decimal that reads back to the same double."""
    return repr(float(v))


class DeckError(ValueError):
    """The case cannot be written as the reference's input deck."""


def write_deck(case: CS.Case, path: str, dxmin_mm: Optional[float] = None, l_c_mm: float = 1000.0) -> List[str]:
    """Write `case` as a reference input deck into directory `path`; returns
    the file names written.

    dxmin_mm: the reference's dxmin in millimetres (default 1000 case.dxmin).
    The reference drifts every body by 0.05 dxmin per step and counts the
    heave amplitude in dxmin (Search.F90:76-96, ReadInput.f90:178), so a
    heaving body needs dxmin > 0; a small value keeps the drift small.
    The flapping frequency is case.freq (1 when it is 0, which only changes
    the sign of zero angles of a body that does not flap): it also fixes the
    time step, dt_order = 1 / (4 freq dt)."""
    blocks = case.blocks
    if len(blocks) < 2 or blocks[0].body is not None:
        raise DeckError("the deck needs a body-free block 1 and at least one body block (main.F90:63-68)")
    bodies = [b.body for b in blocks[1:]]
    if any(p is None for p in bodies):
        raise DeckError("main.F90 gives every block g > 1 the body")
    mesh = bodies[0].body
    for p in bodies[1:]:
        same = all(np.array_equal(getattr(p.body, a), getattr(mesh, a)) for a in ("nodes", "elems", "surf_id"))
        if not same:
            raise DeckError("the deck has one surface mesh for every body block (ReadInput.f90:370)")
    if not np.any(mesh.surf_id == 51):
        raise DeckError("main.F90's block-move check reads the marker nodes that only ID-51 nodes define "
                        "(Search.F90:29-44, :1018-1021): the body needs ID-51 elements (see cases.flap_marked)")
    if any(int(b.cintp) != 3 for b in blocks):
        raise DeckError("readInput sets cintp = 3 for every block (ReadInput.f90:227)")
    if len({float(b.omega) for b in blocks[3:]}) > 1:
        raise DeckError("blocks 4 and up share omega4 (ReadInput.f90:352-357)")
    if dxmin_mm is None:
        dxmin_mm = 1000.0 * case.dxmin
    dxmin = 0.001 * dxmin_mm
    for p in bodies:
        if p.yamp != 0.0 and dxmin == 0.0:
            raise DeckError("a heave amplitude is counted in dxmin: give dxmin_mm > 0")
    freq = case.freq if case.freq > 0.0 else 1.0
    rho_f = case.rho_f
    os.makedirs(os.path.join(path, "geometries"), exist_ok=True)
    written = []

    def put(name: str, lines: List[str]):
        with open(os.path.join(path, name), "w") as f:
            f.write("\n".join(lines) + "\n")
        written.append(name)

    # input_data.nml: readInput derives mu_f = 1e-5 mu, l_c = 1e-3 l_c,
    # u0 = re mu_f / (rho_f l_c), re = rho_f / mu_f, freq = freq u0 / l_c,
    # deltat = 1 / (4 freq dt_order), dxmin = 1e-3 dxmin (ReadInput.f90:128-157)
    mu_f = rho_f / case.re                  # the code's re is rho_f / mu_f
    l_c = 0.001 * l_c_mm
    re_nml = case.uc * rho_f * l_c / mu_f   # so that u0 = uc
    freq_nml = freq * l_c / case.uc         # so that freq = freq * u0 / l_c
    dt_order = 1.0 / (4.0 * freq * case.dt)
    omegas = [b.omega for b in blocks[:4]]
    omegas += [omegas[-1]] * (4 - len(omegas))
    nml = {
        "nblocks": str(len(blocks)), "intflines": str(len(case.interfaces)), "itamax": str(case.nsteps),
        "epsi": _f(case.epsi), "pcItaMax": str(case.pc_ita_max),
        "omega1": _f(omegas[0]), "omega2": _f(omegas[1]), "omega3": _f(omegas[2]), "omega4": _f(omegas[3]),
        "re": _f(re_nml), "rho_f": _f(rho_f), "mu_f": _f(mu_f * 1e5), "l_c": _f(l_c_mm), "u0": _f(case.uc),
        "surGeoPoints": "0", "phase_angle": _f(case.phase_angle_deg), "freq": _f(freq_nml),
        "aoa": _f(case.aoa), "piv_pt": _f(case.piv_pt), "alpha_m": _f(0.0), "theta_m": _f(1.0),
        "istart": "0", "dt_order": _f(dt_order), "inor": str(case.inor), "dxmin": _f(dxmin_mm),
    }
    put("input_data.nml", ["&input_data"] + [f"  {k} = {v}" for k, v in nml.items()] + ["/"])

    grids = [b.grids() for b in blocks]
    boxes = []
    for g, b in enumerate(blocks):
        full = [2, int(b.n[0]) + 1, 2, int(b.n[1]) + 1, 2, int(b.n[2]) + 1]
        boxes.append(CS.search_box(b, grids[g]) if b.body is not None else full)
    put("body_search.dat", [" ".join(str(int(v)) for v in box) for box in boxes])
    put("grid_shift.dat", ["0.0 0.0 0.0" for _ in blocks])
    put("shift.dat", ["0.0 0.0 0.0"] + [" ".join(_f(1000.0 * c) for c in p.centre) for p in bodies])
    put("flap_amp.dat", [_f(p.a0_deg) for p in bodies])
    put("butter_move.dat", [f"{_f(p.yamp / dxmin if dxmin else 0.0)} {_f(p.bfreq)}" for p in bodies])
    # block_details.dat: extents (mm, printed only), cells, and 1/(1000 h)
    # (the reference stores dx = (1/dx) * 0.001, ReadInput.f90:226-231)
    rows = []
    for b, gr in zip(blocks, grids):
        ext = []
        for d in range(3):
            ext += [_f(1000.0 * b.lo[d]), _f(1000.0 * b.hi[d])]
        inv = [_f(0.001 / float(gr[d].delta[1])) for d in range(3)]
        rows.append(" ".join(ext + [str(int(n)) for n in b.n] + inv))
    put("block_details.dat", rows)
    for g, (b, gr) in enumerate(zip(blocks, grids), start=1):
        for d, axis in enumerate("xyz"):
            n = int(b.n[d])
            faces = gr[d].x1[1:n + 2]      # Fortran faces 2 .. n+2
            put(f"{axis}grid_bk{g:03d}_{n + 1}.txt", [_f(1000.0 * v) for v in faces])
    lines = []
    for it in case.interfaces:
        ext = []
        for d in range(3):
            ext += [_f(1000.0 * it.lo[d]), _f(1000.0 * it.hi[d])]
        lines.append(" ".join([str(it.a_blk), str(it.a_msh), "1", str(it.b_blk), str(it.b_msh), "1"] + ext))
    put("interface_details.dat", lines)

    # Gmsh 2.2: 4 header lines, the node count, nodes (mm), 2 lines, the
    # element count, elements "id type ntags physical elementary n1 n2 n3"
    nn, ne = mesh.nodes.shape[0], mesh.elems.shape[0]
    msh = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(nn)]
    for q in range(nn):
        msh.append(f"{q + 1} " + " ".join(_f(1000.0 * v) for v in mesh.nodes[q]))
    msh += ["$EndNodes", "$Elements", str(ne)]
    for q in range(ne):
        e = mesh.elems[q]
        msh.append(f"{q + 1} 2 2 {int(mesh.surf_id[q])} 1 {int(e[0])} {int(e[1])} {int(e[2])}")
    msh.append("$EndElements")
    put(os.path.join("geometries", "butterflyMedium.msh"), msh)
    return written


def config_deck_wing(nsteps: int = 2, pc_ita_max: int = 10, markers: int = 512) -> CS.Case:
    """Coarse block (32, 44, 20) on [0,4]x[0,5.5]x[0,2.5] (h = 1/8) and a
    refinement-ratio-2 block (44, 72, 24) on [0.5,3.25]x[0.5,5]x[0.5,2]
    carrying a flapping, heaving ellipsoid wing (semi-axes 0.2, 0.06, 0.18,
    ID 51, a0 = 20 deg, heave 0.05 at frequency 0.5) at (2.75, 2.75, 1.25).
    The wing keeps more than the block-move margins (13.5, 15 and 4.56
    coarse cells in x, y and z; Search.F90:1013-1033) from the fine block's
    faces, so main.F90 never moves the block."""
    coarse = CS.BlockSpec((0.0, 0.0, 0.0), (4.0, 5.5, 2.5), (32, 44, 20), omega=1.7)
    lo, hi = (0.5, 0.5, 0.5), (3.25, 5.0, 2.0)
    body = CS.BodyPlacement(CS.ellipsoid_body(markers, (0.2, 0.06, 0.18), surf_id=51), (2.75, 2.75, 1.25),
                            a0_deg=20.0, yamp=0.05, bfreq=0.5)
    fine = CS.BlockSpec(lo, hi, (44, 72, 24), omega=1.7, body=body, moving=True)
    itf = CS.InterfaceSpec(1, 2, 1, 2, lo, hi)
    return CS.Case("deck_wing", [coarse, fine], [itf], dt=0.002, re=300.0, nsteps=nsteps, pc_ita_max=pc_ita_max,
                   epsi=0.0, freq=0.5, phase_angle_deg=90.0, dxmin=1e-6, noise_amp=0.0)


flap_marked = CS.flap_marked

__all__ = ["write_deck", "DeckError", "flap_marked", "config_deck_wing"]
