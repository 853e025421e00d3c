"""The benchmark problems as concrete inputs (host arrays).

The reference publishes no meshes for this path; these are its documented
benchmark configurations built on structured C3D8 boxes (node and element
order of ufc_mesh_box).  Equations are 0-based eq = 3*node + d.  Faces are
0-based indices into PH_ELEM_C3D8_FACE_NODES (PH_Elem_C3D8.f90:120-121):
0 zeta=-1, 1 zeta=+1, 2 eta=-1, 3 eta=+1, 4 xi=-1, 5 xi=+1.
"""
from __future__ import annotations

from dataclasses import dataclass, field as dfield
import math

import numpy as np

from . import static as S

SEED = 20251203
PENALTY = 1.0e12  # RT_Asm_Solv.f90:326


@dataclass
class Problem:
    name: str
    nx: int
    ny: int
    nz: int
    coords: np.ndarray
    conn: np.ndarray
    young: np.ndarray
    nu: float
    bc_eq: np.ndarray
    bc_val: np.ndarray
    load_elem: np.ndarray
    load_face: np.ndarray
    load_kind: str            # "traction" or "pressure"
    load_value: tuple         # traction vector or (p,)
    extra: dict = dfield(default_factory=dict)

    @property
    def n_dof(self):
        return 3 * self.coords.shape[0]

    @property
    def n_elems(self):
        return self.conn.shape[0]


def _elem_grid(nx, ny, nz):
    return np.arange(nx * ny * nz, dtype=np.int32).reshape(nz, ny, nx)


def _nodes_where(coords, axis, value, tol=1e-12):
    return np.nonzero(np.abs(coords[:, axis] - value) < tol)[0].astype(np.int32)


def _clamp_eqs(nodes):
    return (3 * nodes[:, None] + np.arange(3, dtype=np.int32)[None, :]).ravel().astype(np.int32)


def cantilever(nx=20, ny=4, nz=4, lx=10.0, ly=2.0, lz=2.0, young=210e9, nu=0.3,
               total=-1.0e3) -> Problem:
    """Cantilever on a box, x=0 clamped, uniform traction on x=lx summing to
    `total` in z (lumped area/4 per face node)."""
    coords, conn = S.mesh_box(nx, ny, nz, lx, ly, lz)
    clamp = _clamp_eqs(_nodes_where(coords, 0, 0.0))
    elem = _elem_grid(nx, ny, nz)[:, :, nx - 1].ravel()
    t = total / (ly * lz)
    return Problem("cantilever", nx, ny, nz, coords, conn, np.full(conn.shape[0], young), nu,
                   clamp, np.zeros(clamp.size), elem, np.full(elem.size, 5, np.int32),
                   "traction", (0.0, 0.0, t))


def cube_pressure(n, young=None, nu=0.3, p=1.0e6, name="cube", jitter=0.0, seed=SEED):
    """Unit cube n^3, z=0 clamped, pressure p on z=1 (outward normal, +z)."""
    coords, conn = S.mesh_box(n, n, n, 1.0, 1.0, 1.0, jitter, seed + 2)
    clamp = _clamp_eqs(_nodes_where(coords, 2, 0.0))
    elem = _elem_grid(n, n, n)[n - 1, :, :].ravel()
    if young is None:
        young = S.field("lognormal", conn.shape[0], seed, math.log(1.0e9), 0.5)
    return Problem(name, n, n, n, coords, conn, np.asarray(young, np.float64), nu, clamp,
                   np.zeros(clamp.size), elem, np.full(elem.size, 1, np.int32), "pressure", (p,))


def lognormal_cube(n=128, seed=SEED) -> Problem:
    """128^3 unit cube, E_e ~ lognormal(ln 1e9, 0.5), nu 0.3, z=0 clamped, p=1e6 on z=1."""
    return cube_pressure(n, name=f"cube{n}_lognormal", seed=seed)


def slab_share(n=128, parts=8, seed=SEED) -> Problem:
    """One GPU's share of the n^3 lognormal cube cut into `parts` z-slabs: the
    bottom slab, n x n x n/parts hexes of size 1/n (the first elements of the
    cube's numbering, so the same seeded E field), z=0 clamped, p=1e6 on its
    top face (where the next slab's halo would act)."""
    nz = n // parts
    coords, conn = S.mesh_box(n, n, nz, 1.0, 1.0, nz / n, 0.0, seed + 2)
    clamp = _clamp_eqs(_nodes_where(coords, 2, 0.0))
    elem = _elem_grid(n, n, nz)[nz - 1, :, :].ravel()
    young = S.field("lognormal", conn.shape[0], seed, math.log(1.0e9), 0.5)
    return Problem(f"slab{n}x{n}x{nz}", n, n, nz, coords, conn, young, 0.3, clamp,
                   np.zeros(clamp.size), elem, np.full(elem.size, 1, np.int32), "pressure",
                   (1.0e6,))


def simp_cube(n=96, seed=SEED):
    """96^3 SIMP cube: rho_e ~ U[0.3, 1], E_e = 1e-3 + rho_e^3 1e9."""
    rho = S.field("uniform", n ** 3, seed + 1, 0.3, 1.0)
    young = 1.0e-3 + rho ** 3 * 1.0e9
    pr = cube_pressure(n, young=young, name=f"cube{n}_simp", seed=seed)
    pr.extra["rho"] = rho
    return pr


def j2_cube(n=64, seed=SEED, strain=0.005, increments=10):
    """64^3 jittered cube for J2: z=0 clamped, u_z prescribed on z=1 up to `strain`."""
    coords, conn = S.mesh_box(n, n, n, 1.0, 1.0, 1.0, 0.1, seed + 2)
    bottom = _nodes_where(coords, 2, 0.0)
    top = _nodes_where(coords, 2, 1.0)
    eq = np.concatenate([_clamp_eqs(bottom), (3 * top + 2).astype(np.int32)])
    pr = Problem(f"cube{n}_j2", n, n, n, coords, conn, np.full(conn.shape[0], 200e9), 0.3, eq,
                 np.zeros(eq.size), np.zeros(0, np.int32), np.zeros(0, np.int32), "pressure",
                 (0.0,))
    pr.extra.update(props=(200e9, 0.3, 250e6, 2e9), top_eq=(3 * top + 2).astype(np.int32),
                    n_clamp=3 * bottom.size, strain=strain, increments=increments)
    return pr


def polar_prism(sectors=5, rings=3, nz=4, radius=1.0, height=1.0, nu=0.3, p=1.0e6, seed=SEED,
                name=None) -> Problem:
    """An unstructured C3D8 mesh: a polar disk of quadrilaterals around a
    centre node, extruded in z.  `sectors` kite quads meet at the centre
    (c, p_i, q_i, p_{i+1}: p at radius dr, q between them at 1.15 dr), and
    rings - 1 rings of 2 sectors quads surround them, so the axis nodes of
    the interior layers lie in 2 sectors elements (10 for 5 sectors) with
    3 (2 sectors + 1) neighbours (33), beyond an 8-element / 32-neighbour
    node map.  z=0 clamped, pressure p on z=height, E_e lognormal(ln 1e9, 0.5)."""
    ns, R = sectors, rings
    dr = radius / R
    pts = [(0.0, 0.0)]
    ring = []  # ring[j][k]: 2 ns points at angle pi k / ns
    for j in range(1, R + 1):
        ids = []
        for k in range(2 * ns):
            r = j * dr * (1.15 if (j == 1 and k % 2) else 1.0)
            t = math.pi * k / ns
            ids.append(len(pts))
            pts.append((r * math.cos(t), r * math.sin(t)))
        ring.append(ids)
    quads = []
    r1 = ring[0]
    for i in range(ns):  # counter-clockwise kites around the centre
        quads.append((0, r1[2 * i], r1[2 * i + 1], r1[(2 * i + 2) % (2 * ns)]))
    for j in range(R - 1):
        a, b = ring[j], ring[j + 1]
        for k in range(2 * ns):
            k1 = (k + 1) % (2 * ns)
            quads.append((a[k], b[k], b[k1], a[k1]))
    n2 = len(pts)
    xy = np.array(pts)
    coords = np.concatenate([np.column_stack([xy, np.full(n2, height * k / nz)])
                             for k in range(nz + 1)])
    q = np.array(quads, np.int32)
    conn = np.concatenate([np.concatenate([q + k * n2, q + (k + 1) * n2], axis=1)
                           for k in range(nz)]).astype(np.int32)
    clamp = _clamp_eqs(np.arange(n2, dtype=np.int32))
    top = np.arange((nz - 1) * len(quads), nz * len(quads), dtype=np.int32)
    young = S.field("lognormal", conn.shape[0], seed, math.log(1.0e9), 0.5)
    return Problem(name or f"polar{ns}x{R}x{nz}", ns, R, nz, coords, conn, young, nu, clamp,
                   np.zeros(clamp.size), top, np.full(top.size, 1, np.int32), "pressure", (p,))


def setup_context(ctx: "S.Context", pr: Problem, precision, penalty=PENALTY):
    """Pattern, loads and BCs of `pr` on `ctx` (assembly is separate)."""
    ctx.build_pattern(pr.coords, pr.conn)
    if pr.load_elem.size:
        if pr.load_kind == "traction":
            ctx.load_traction(pr.load_elem, pr.load_face, pr.load_value)
        else:
            ctx.load_pressure(pr.load_elem, pr.load_face, pr.load_value[0])
    ctx.apply_bc(pr.bc_eq, pr.bc_val, penalty)
