"""Synthetic inputs for the Bio-CFD time-step hot path.

Builds the per-block grid arrays, the triangulated immersed bodies, the seeded
initial noise and the case file that the reference driver harness, the CPU
oracle and the HIP runtime all read, so the three start from bit-identical
inputs.

Grid arrays follow the reference's grid construction
(reference src/ReadInput.f90:254-351): nodes x1(2..n+2) are given, cell widths
deltax(i) = x1(i+1)-x1(i) with the end widths copied into the ghost entries,
ghost nodes extrapolated by one width, xu = x1, and xv = xw = xp the cell
centres 0.5*(x1(i)+x1(i+1)).  Arrays here are 0-based, so Fortran index i is
Python index i-1.

Bodies are triangle meshes whose element centroids are the "markers" (reference
src/Search.F90:221-253): a sphere of M markers is the convex hull of
V = (M+4)/2 Fibonacci points, which has exactly M triangles.
"""
from __future__ import annotations

import dataclasses
import struct
from typing import List, Optional, Sequence

import numpy as np

CASE_MAGIC = 20261203
CASE_VERSION = 1
P_HASH = 2147483647  # Mersenne prime 2^31-1


# --------------------------------------------------------------------------
# grids
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Grid1D:
    x1: np.ndarray      # (n+3) faces incl. ghosts
    delta: np.ndarray   # (n+2)
    xc: np.ndarray      # (n+2) cell centres (xp = xv = xw for x)


def grid_1d(x0: float, x1: float, n: int) -> Grid1D:
    """Uniform 1-D grid of n interior cells on [x0, x1] (ReadInput.f90:254-287)."""
    h = (x1 - x0) / n
    nodes = x0 + np.arange(n + 1, dtype=np.float64) * h
    f = np.empty(n + 3)
    f[1:n + 2] = nodes
    d = np.empty(n + 2)
    d[1:n + 1] = f[2:n + 2] - f[1:n + 1]
    d[0] = d[1]
    d[n + 1] = d[n]
    f[0] = f[1] - d[0]
    f[n + 2] = f[n + 1] + d[n + 1]
    xc = 0.5 * (f[0:n + 2] + f[1:n + 3])
    return Grid1D(f, d, xc)


# --------------------------------------------------------------------------
# bodies
# --------------------------------------------------------------------------
@dataclasses.dataclass
class Body:
    nodes: np.ndarray     # (N,3) body-frame node coordinates
    node_id: np.ndarray   # (N,) int64 surface ID of each node (50 rigid, 51/52 flapping)
    elems: np.ndarray     # (E,3) int64 1-based node indices, outward orientation
    surf_id: np.ndarray   # (E,) int64


def _fibonacci_points(v: int) -> np.ndarray:
    i = np.arange(v, dtype=np.float64)
    golden = (1.0 + 5.0 ** 0.5) / 2.0
    z = 1.0 - (2.0 * i + 1.0) / v
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = 2.0 * np.pi * i / golden
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1)


def _canonical_triangles(pts: np.ndarray, tris: np.ndarray, centre: np.ndarray) -> np.ndarray:
    """Orient every triangle outward and fix a deterministic order."""
    out = []
    for t in tris:
        a, b, c = (int(x) for x in t)
        n = np.cross(pts[b] - pts[a], pts[c] - pts[a])
        cen = (pts[a] + pts[b] + pts[c]) / 3.0
        if np.dot(n, cen - centre) < 0:
            b, c = c, b
        # rotate so the smallest index comes first (keeps orientation)
        tri = [a, b, c]
        k = tri.index(min(tri))
        tri = tri[k:] + tri[:k]
        out.append(tri)
    out.sort()
    return np.asarray(out, dtype=np.int64)


def ellipsoid_body(markers: int, semi_axes: Sequence[float], surf_id: int = 50) -> Body:
    """Closed ellipsoid surface with `markers` triangles (Fibonacci hull)."""
    from scipy.spatial import ConvexHull

    if markers % 2 != 0 or markers < 4:
        raise ValueError("marker count must be even and >= 4")
    v = (markers + 4) // 2
    unit = _fibonacci_points(v)
    hull = ConvexHull(unit)
    tris = _canonical_triangles(unit, hull.simplices, np.zeros(3))
    if tris.shape[0] != markers:
        raise RuntimeError(f"hull has {tris.shape[0]} triangles, expected {markers}")
    pts = unit * np.asarray(semi_axes, dtype=np.float64)[None, :]
    return Body(pts, np.full(v, surf_id, np.int64), tris + 1, np.full(markers, surf_id, np.int64))


def sphere_body(markers: int, diameter: float, surf_id: int = 50) -> Body:
    r = 0.5 * diameter
    return ellipsoid_body(markers, (r, r, r), surf_id)


def plate_body(nc: int, ns: int, chord: float, span: float, thickness: float,
               surf_id: int = 51) -> Body:
    """Thin box (chord along x, span along z, thickness along y); regular grid
    of nc x ns quads on the two large faces plus the four edge strips."""
    from scipy.spatial import ConvexHull  # noqa: F401  (kept for symmetry)

    xs = np.linspace(-0.5 * chord, 0.5 * chord, nc + 1)
    zs = np.linspace(-0.5 * span, 0.5 * span, ns + 1)
    h = 0.5 * thickness
    nodes = []
    index = {}

    def node(x, y, z):
        key = (round(x, 12), round(y, 12), round(z, 12))
        if key not in index:
            index[key] = len(nodes)
            nodes.append((x, y, z))
        return index[key]

    tris = []
    for sgn in (1.0, -1.0):
        for a in range(nc):
            for b in range(ns):
                p00 = node(xs[a], sgn * h, zs[b])
                p10 = node(xs[a + 1], sgn * h, zs[b])
                p01 = node(xs[a], sgn * h, zs[b + 1])
                p11 = node(xs[a + 1], sgn * h, zs[b + 1])
                tris += [(p00, p10, p11), (p00, p11, p01)]
    # edge strips (x = +-c/2 and z = +-s/2)
    for sx in (0, nc):
        for b in range(ns):
            q0 = node(xs[sx], h, zs[b]); q1 = node(xs[sx], h, zs[b + 1])
            r0 = node(xs[sx], -h, zs[b]); r1 = node(xs[sx], -h, zs[b + 1])
            tris += [(q0, q1, r1), (q0, r1, r0)]
    for sz in (0, ns):
        for a in range(nc):
            q0 = node(xs[a], h, zs[sz]); q1 = node(xs[a + 1], h, zs[sz])
            r0 = node(xs[a], -h, zs[sz]); r1 = node(xs[a + 1], -h, zs[sz])
            tris += [(q0, q1, r1), (q0, r1, r0)]
    pts = np.asarray(nodes, dtype=np.float64)
    t = _canonical_triangles(pts, np.asarray(tris), np.zeros(3))
    return Body(pts, np.full(len(pts), surf_id, np.int64), t + 1,
                np.full(t.shape[0], surf_id, np.int64))


def flap_marked(body: Body) -> Body:
    """The body with its rigid (ID 50) elements and nodes marked 51.  With
    flap amplitude a0 = 0 and pivot 0 an ID-51 node is rotated by a zero
    angle, z cos 0 + y sin 0 + 0 - 0 cos 0 = z and -z sin 0 + y cos 0 = y
    (Search.F90:186-195), so the body stays where an ID-50 body would be, and
    main.F90's block-move check has the marker nodes it reads."""
    sid = np.where(body.surf_id == 50, 51, body.surf_id)
    nid = np.where(body.node_id == 50, 51, body.node_id)
    return Body(body.nodes.copy(), nid, body.elems.copy(), sid)


# --------------------------------------------------------------------------
# blocks and cases
# --------------------------------------------------------------------------
@dataclasses.dataclass
class BodyPlacement:
    body: Body
    centre: Sequence[float]          # xshift, yshift, zshift
    a0_deg: float = 0.0              # flap amplitude (IDs 51/52 rotate about x)
    yamp: float = 0.0                # heave amplitude
    bfreq: float = 0.0               # heave frequency
    margin_cells: int = 3


@dataclasses.dataclass
class BlockSpec:
    lo: Sequence[float]
    hi: Sequence[float]
    n: Sequence[int]
    omega: float = 1.7
    cintp: int = 3
    body: Optional[BodyPlacement] = None
    moving: bool = False

    def grids(self):
        return [grid_1d(self.lo[d], self.hi[d], int(self.n[d])) for d in range(3)]


@dataclasses.dataclass
class InterfaceSpec:
    a_blk: int
    b_blk: int
    a_msh: int
    b_msh: int
    lo: Sequence[float]
    hi: Sequence[float]


@dataclasses.dataclass
class Case:
    name: str
    blocks: List[BlockSpec]
    interfaces: List[InterfaceSpec] = dataclasses.field(default_factory=list)
    dt: float = 0.002
    re: float = 500.0
    uc: float = 1.0
    epsi: float = 1e-8
    pc_ita_max: int = 500
    nsteps: int = 20
    inor: int = 1
    seed: int = 20261203
    noise_amp: float = 1e-3
    aoa: float = 0.0
    phase_angle_deg: float = 0.0
    piv_pt: float = 0.0
    dxmin: float = 0.0
    freq: float = 0.0
    mu_f: float = 1.0
    rho_f: float = 1.0
    dump_steps: Sequence[int] = (0,)

    @property
    def ib_block1(self) -> bool:
        return self.blocks[0].body is not None

    def cells(self) -> int:
        return int(sum(int(np.prod(b.n)) for b in self.blocks))


def search_box(spec: BlockSpec, grids) -> List[int]:
    """Fortran-indexed search box around the body's bounding sphere."""
    pl = spec.body
    assert pl is not None
    rad = float(np.max(np.linalg.norm(pl.body.nodes, axis=1)))
    ext = rad + abs(pl.yamp)
    box = []
    for d in range(3):
        g = grids[d]
        n = int(spec.n[d])
        c = float(pl.centre[d])
        cen = g.xc  # 0-based; Fortran index = python + 1
        lo_i = int(np.searchsorted(cen, c - ext)) + 1 - pl.margin_cells
        hi_i = int(np.searchsorted(cen, c + ext)) + 1 + pl.margin_cells
        box += [max(2, lo_i), min(n + 1, hi_i)]
    return box


def noise_uniform(seed: int, field_id: int, lin: np.ndarray) -> np.ndarray:
    """Integer-hash uniform numbers in [0,1]; the same hash is restated in the
    harness (oracle/ref/ref_harness.F90) and the C oracle."""
    x = (np.int64(seed) + np.int64(2654435) * np.int64(field_id) + lin.astype(np.int64)) % P_HASH
    for _ in range(3):
        x = (np.int64(48271) * x + np.int64(11)) % P_HASH
    return x.astype(np.float64) / 2147483647.0


def _i8(*vals) -> bytes:
    return struct.pack(f"<{len(vals)}q", *[int(v) for v in vals])


def _f8(*vals) -> bytes:
    return struct.pack(f"<{len(vals)}d", *[float(v) for v in vals])


def _arr(a: np.ndarray, dtype) -> bytes:
    return np.ascontiguousarray(a, dtype=dtype).tobytes()


def write_case(case: Case, path: str) -> None:
    """Serialise a case in the layout ref_harness.F90 reads."""
    out = bytearray()
    out += _i8(CASE_MAGIC, CASE_VERSION, len(case.blocks), len(case.interfaces),
               1 if case.ib_block1 else 0, case.pc_ita_max, case.nsteps, case.inor, case.seed)
    out += _f8(case.dt, case.re, case.uc, case.epsi, case.noise_amp, case.aoa,
               case.phase_angle_deg, case.piv_pt, case.dxmin, case.freq, case.mu_f, case.rho_f)
    for spec in case.blocks:
        gx, gy, gz = spec.grids()
        nx, ny, nz = (int(v) for v in spec.n)
        has_ib = spec.body is not None
        out += _i8(nx, ny, nz, spec.cintp, 1 if has_ib else 0, 1 if spec.moving else 0)
        out += _f8(gx.delta[1], gy.delta[1], gz.delta[1], spec.omega)
        out += _arr(gx.x1, "<f8") + _arr(gy.x1, "<f8") + _arr(gz.x1, "<f8")
        out += _arr(gx.delta, "<f8") + _arr(gy.delta, "<f8") + _arr(gz.delta, "<f8")
        # xu, yu, zu ; xv, yv, zv ; xw, yw, zw ; xp, yp, zp
        out += _arr(gx.x1, "<f8") + _arr(gy.xc, "<f8") + _arr(gz.xc, "<f8")
        out += _arr(gx.xc, "<f8") + _arr(gy.x1, "<f8") + _arr(gz.xc, "<f8")
        out += _arr(gx.xc, "<f8") + _arr(gy.xc, "<f8") + _arr(gz.x1, "<f8")
        out += _arr(gx.xc, "<f8") + _arr(gy.xc, "<f8") + _arr(gz.xc, "<f8")
        if has_ib:
            pl = spec.body
            b = pl.body
            out += _i8(b.nodes.shape[0], b.elems.shape[0])
            out += _i8(*search_box(spec, (gx, gy, gz)))
            out += _f8(pl.centre[0], pl.centre[1], pl.centre[2], pl.a0_deg, pl.yamp, pl.bfreq)
            out += _arr(b.nodes[:, 0], "<f8") + _arr(b.nodes[:, 1], "<f8") + _arr(b.nodes[:, 2], "<f8")
            out += _arr(b.node_id, "<i8")
            out += _arr(b.elems[:, 0], "<i8") + _arr(b.elems[:, 1], "<i8") + _arr(b.elems[:, 2], "<i8")
            out += _arr(b.surf_id, "<i8")
    for itf in case.interfaces:
        out += _i8(itf.a_blk, itf.b_blk, itf.a_msh, itf.b_msh)
        out += _f8(itf.lo[0], itf.hi[0], itf.lo[1], itf.hi[1], itf.lo[2], itf.hi[2])
    out += _i8(len(case.dump_steps))
    if case.dump_steps:
        out += _i8(*case.dump_steps)
    with open(path, "wb") as f:
        f.write(out)


# --------------------------------------------------------------------------
# named configurations (full-size workloads and small parity variants)
# --------------------------------------------------------------------------
def re_code(reynolds: float, u: float, length: float) -> float:
    """The reference's `re` is an inverse kinematic viscosity
    (NavierStokes.f90:672), so re_code = Re / (U L)."""
    return reynolds / (u * length)


def config_sphere_64(n: int = 64, markers: int = 1024, nsteps: int = 20, **kw) -> Case:
    """Single block on the unit cube, rigid sphere d=0.2 at (0.35,0.5,0.5)."""
    body = BodyPlacement(sphere_body(markers, 0.2), (0.35, 0.5, 0.5))
    blk = BlockSpec((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), (n, n, n), omega=1.7, body=body)
    args = dict(dt=0.002, re=re_code(100.0, 1.0, 0.2), nsteps=nsteps, pc_ita_max=500)
    args.update(kw)
    return Case(f"sphere{n}", [blk], **args)


def config_sphere_256(n=(256, 256, 256), markers: int = 8192, nsteps: int = 100, **kw) -> Case:
    """Single block on [0,8]x[0,4]x[0,4], sphere d=1 at (2,2,2), Re=300."""
    body = BodyPlacement(sphere_body(markers, 1.0), (2.0, 2.0, 2.0))
    blk = BlockSpec((0.0, 0.0, 0.0), (8.0, 4.0, 4.0), n, omega=1.8, body=body)
    args = dict(dt=0.0005, re=re_code(300.0, 1.0, 1.0), nsteps=nsteps, pc_ita_max=500)
    args.update(kw)
    return Case("sphere256", [blk], **args)


def config_nested(nc: int = 128, nf: int = 128, markers: int = 4096, nsteps: int = 50,
                  moving: bool = True, **kw) -> Case:
    """Coarse block on [0,4]^3 plus a refinement-ratio-2 fine block on
    [1,3]^3 carrying a sphere d=0.5 at (2,2,2) that heaves in y."""
    coarse = BlockSpec((0.0, 0.0, 0.0), (4.0, 4.0, 4.0), (nc, nc, nc), omega=1.7)
    lo, hi = (1.0, 1.0, 1.0), (3.0, 3.0, 3.0)
    body = BodyPlacement(sphere_body(markers, 0.5), (2.0, 2.0, 2.0),
                         yamp=0.125 if moving else 0.0, bfreq=0.2 if moving else 0.0)
    fine = BlockSpec(lo, hi, (nf, nf, nf), omega=1.7, body=body, moving=moving)
    itf = InterfaceSpec(1, 2, 1, 2, lo, hi)
    args = dict(dt=0.001, re=re_code(150.0, 1.0, 0.5), nsteps=nsteps, pc_ita_max=500)
    args.update(kw)
    return Case(f"nested{nc}", [coarse, fine], [itf], **args)


def config_channel(n=(16, 12, 10), nsteps: int = 3, **kw) -> Case:
    """Single block without a body (inlet/outlet/slip walls only)."""
    blk = BlockSpec((0.0, 0.0, 0.0), (1.0, 0.75, 0.625), n, omega=1.7)
    args = dict(dt=0.002, re=500.0, nsteps=nsteps, pc_ita_max=40, epsi=0.0)
    args.update(kw)
    return Case("channel", [blk], **args)


def config_nested_move(nc: int = 24, nf: int = 24, markers: int = 256, nsteps: int = 2, **kw) -> Case:
    """Coarse block on [0,6]^3 plus a refinement-ratio-2 block on [1.5,4.5]^3
    carrying a heaving sphere d=0.5 at (3,3,3) whose nodes are ID 51 (the
    marker nodes of the reference's block-move check) with flap amplitude 0.
    At these sizes the sphere is within the move margins of the fine block's
    faces (4.5, 5 and 1.52 times cintp coarse cells), so the fine block moves
    every step: one coarse cell in -x and in +y.  The fine block is wide
    enough that two moves keep the stress probes (one cell diagonal off the
    surface) on its grid, and the search box is widened to keep the sphere
    inside it."""
    coarse = BlockSpec((0.0, 0.0, 0.0), (6.0, 6.0, 6.0), (nc, nc, nc), omega=1.7)
    lo, hi = (1.5, 1.5, 1.5), (4.5, 4.5, 4.5)
    body = BodyPlacement(flap_marked(sphere_body(markers, 0.5)), (3.0, 3.0, 3.0), yamp=0.125, bfreq=0.2,
                         margin_cells=6)
    fine = BlockSpec(lo, hi, (nf, nf, nf), omega=1.7, body=body, moving=True)
    itf = InterfaceSpec(1, 2, 1, 2, lo, hi)
    args = dict(dt=0.001, re=re_code(150.0, 1.0, 0.5), nsteps=nsteps, pc_ita_max=10, epsi=0.0, dxmin=1e-6)
    args.update(kw)
    return Case(f"move{nc}", [coarse, fine], [itf], **args)


def config_batched(nc=(56, 32, 32), wake=(16, 20, 24), chord_cells: int = 8, span_cells: int = 16, nsteps: int = 3,
                   **kw) -> Case:
    """Uneven multi-block case: a coarse block on [0,14]x[0,8]x[0,8] and
    refinement-ratio-2 fine blocks, the first ([2,4]x[2.5,5.5]x[2.5,5.5])
    carrying a flat plate (chord 1 along x, span 2 along z, thickness 0.04;
    ID 51) that heaves (0.5 at frequency 0.25) and pitches about x (30 deg,
    phase 90 deg), then cubes of the given sides (fine cells) one after the
    other along the wake, without bodies.  The reduced form of the batched
    multi-block workload; the fine blocks' SOR calls run in one batched
    launch per pass."""
    nx, ny, nz = nc
    coarse = BlockSpec((0.0, 0.0, 0.0), (14.0, 8.0, 8.0), (nx, ny, nz), omega=1.7)
    hf = 14.0 / nx / 2          # fine spacing
    plate = plate_body(chord_cells, span_cells, 1.0, 2.0, 0.04, surf_id=51)
    body = BodyPlacement(plate, (3.0, 4.0, 4.0), a0_deg=30.0, yamp=0.5, bfreq=0.25, margin_cells=2)
    lo, hi = (2.0, 2.5, 2.5), (4.0, 5.5, 5.5)
    nb = [int(round((hi[d] - lo[d]) / hf)) for d in range(3)]
    blocks = [coarse, BlockSpec(lo, hi, nb, omega=1.7, body=body, moving=True)]
    itfs = [InterfaceSpec(1, 2, 1, 2, lo, hi)]
    x = 4.5
    for side in wake:
        ext = side * hf
        blo = (x, 4.0 - ext / 2, 4.0 - ext / 2)
        bhi = (x + ext, 4.0 + ext / 2, 4.0 + ext / 2)
        blocks.append(BlockSpec(blo, bhi, (side, side, side), omega=1.7))
        itfs.append(InterfaceSpec(1, len(blocks), 1, 2, blo, bhi))
        x += ext + 0.5
    args = dict(dt=0.0005, re=200.0, nsteps=nsteps, pc_ita_max=20, epsi=0.0, freq=0.25, phase_angle_deg=90.0)
    args.update(kw)
    return Case("batched", blocks, itfs, **args)


def _pack_boxes(extents, lo, hi, gap):
    """Non-overlapping positions (lower corners) for boxes of the given
    extents inside [lo, hi] with at least `gap` between any two, or None.
    Largest volume first, each at the first free candidate corner in (x, z,
    y) order, candidates being the domain corner and the faces of the boxes
    already placed (bottom-left fill)."""
    n = len(extents)
    order = sorted(range(n), key=lambda q: -float(np.prod(extents[q])))
    placed = []
    out = [None] * n

    def fits(p, e):
        for d in range(3):
            if p[d] < lo[d] - 1e-12 or p[d] + e[d] > hi[d] + 1e-12:
                return False
        for a, b in placed:
            if all(p[d] < b[d] + gap - 1e-12 and a[d] < p[d] + e[d] + gap - 1e-12 for d in range(3)):
                return False
        return True

    for q in order:
        e = extents[q]
        cands = {tuple(lo)}
        for a, b in placed:
            for d in range(3):
                c = list(a)
                c[d] = b[d] + gap
                cands.add(tuple(c))
                for d2 in range(3):
                    if d2 != d:
                        c2 = list(c)
                        c2[d2] = lo[d2]
                        cands.add(tuple(c2))
        pos = next((c for c in sorted(cands, key=lambda c: (c[0], c[2], c[1])) if fits(c, e)), None)
        if pos is None:
            return None
        placed.append((pos, tuple(pos[d] + e[d] for d in range(3))))
        out[q] = pos
    return out


def config_uneven(n_wake: int = 15, seed: int = 1, sides=tuple(range(32, 129, 16)), markers_chord: int = 32,
                  markers_span: int = 32, nsteps: int = 5, **kw) -> Case:
    """The batched multi-block workload: a coarse block 192x128x128 on
    [0,12]x[0,8]x[0,8] (h = 1/16) and 16 refinement-ratio-2 fine blocks:
    one of 80^3 fine cells on [1,3.5]x[2.75,5.25]^2 carrying a flat plate
    (chord 1, span 2, thickness 0.04, ID 51, 2 x chord x span quads per
    face plus edge strips) that heaves (0.5 at 0.25) and pitches about x
    (30 deg, phase 90 deg), and n_wake body-free blocks behind it, each side
    (fine cells) drawn with the seed from `sides` ({32, 48, ..., 128}),
    placed without overlap (gaps >= 1/4) in the wake [3.75,11.75]x[0.25,7.75]^2
    inside the coarse block.  A draw that does not fit is replaced by the
    draw of the next seed (seed 1 fits: 6.06 M wake cells)."""
    coarse = BlockSpec((0.0, 0.0, 0.0), (12.0, 8.0, 8.0), (192, 128, 128), omega=1.7)
    hf = 1.0 / 32
    plate = plate_body(markers_chord, markers_span, 1.0, 2.0, 0.04, surf_id=51)
    body = BodyPlacement(plate, (2.25, 4.0, 4.0), a0_deg=30.0, yamp=0.5, bfreq=0.25, margin_cells=3)
    lo, hi = (1.0, 2.75, 2.75), (3.5, 5.25, 5.25)
    blocks = [coarse, BlockSpec(lo, hi, (80, 80, 80), omega=1.7, body=body, moving=True)]
    itfs = [InterfaceSpec(1, 2, 1, 2, lo, hi)]
    wake_lo, wake_hi, gap = (3.75, 0.25, 0.25), (11.75, 7.75, 7.75), 0.25
    for s in range(seed, seed + 1000):
        rng = np.random.default_rng(s)
        dims = [tuple(int(rng.choice(sides)) for _ in range(3)) for _ in range(n_wake)]
        pos = _pack_boxes([tuple(v * hf for v in d) for d in dims], wake_lo, wake_hi, gap)
        if pos is not None:
            break
    else:
        raise ValueError("no draw of the wake blocks fits the coarse block")
    for (bx, by, bz), d in zip(pos, dims):
        blo = (bx, by, bz)
        bhi = (bx + d[0] * hf, by + d[1] * hf, bz + d[2] * hf)
        blocks.append(BlockSpec(blo, bhi, d, omega=1.7))
        itfs.append(InterfaceSpec(1, len(blocks), 1, 2, blo, bhi))
    args = dict(dt=0.0005, re=200.0, nsteps=nsteps, pc_ita_max=500, epsi=1e-8, freq=0.25, phase_angle_deg=90.0)
    args.update(kw)
    return Case("uneven17", blocks, itfs, **args)


def config_adjoint(n: int = 96, markers: int = 2048, nsteps: int = 40, **kw) -> Case:
    """The adjoint workload: single block on [0,3]^3, an ellipsoid (semi-axes
    0.4, 0.2, 0.1; ID 51, so it pitches about x with the reference's flap
    kinematics, Search.F90:186-195) at (1, 1.5, 1.5) heaving 0.2 in y,
    pitch amplitude 20 deg, phase 90 deg, frequency 0.5, Re = 100 on the
    major axis (re_code = 100 / 0.8), dt = 0.001."""
    body = BodyPlacement(ellipsoid_body(markers, (0.4, 0.2, 0.1), surf_id=51), (1.0, 1.5, 1.5), a0_deg=20.0,
                         yamp=0.2, bfreq=0.5)
    blk = BlockSpec((0.0, 0.0, 0.0), (3.0, 3.0, 3.0), (n, n, n), omega=1.7, body=body, moving=True)
    args = dict(dt=0.001, re=re_code(100.0, 1.0, 0.8), nsteps=nsteps, pc_ita_max=500, freq=0.5,
                phase_angle_deg=90.0)
    args.update(kw)
    return Case(f"adjoint{n}", [blk], **args)
