"""Input decks for the static C3D8 step.

read_inp() calls ufc_ap_read_inp (the keyword reading of
AP_Input_Domain_ParseKeyword, L6_AP/Input/Parser/AP_Inp_Mgr.f90:99-125, for
the keywords of this path) and copies the model into numpy arrays;
setup_context() puts its loads and boundary conditions on a Context the way
problems.setup_context does for the built-in benchmarks; write_inp() writes a
problems.Problem as a deck that read_inp() reads back.

Equations are 0-based eq = 3*node + d; faces are 0-based columns of
PH_ELEM_C3D8_FACE_NODES (PH_Elem_C3D8.f90:120-121), label Sk = column k as in
the reference's *SURFACE mapping (MD_KW_Mapper.f90:2137-2155).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
import os

import numpy as np

from . import _abi as A

PENALTY = 1.0e12  # RT_Asm_Solv.f90:326


@dataclass
class Deck:
    node_label: np.ndarray   # (n_nodes,) labels of the deck
    elem_label: np.ndarray   # (n_elems,)
    coords: np.ndarray       # (n_nodes, 3)
    conn: np.ndarray         # (n_elems, 8) 0-based node indices
    young: np.ndarray        # (n_elems,)
    nu: float
    bc_eq: np.ndarray        # prescribed equations
    bc_val: np.ndarray
    cload_eq: np.ndarray     # concentrated loads, repeated equations add up
    cload_val: np.ndarray
    face_elem: np.ndarray    # pressure-loaded faces (0-based element index)
    face_id: np.ndarray      # 0..5
    face_p: np.ndarray       # pressure along the outward normal
    variant: int = 1         # UFC_VARIANT_* of the element type (C3D8R: 2, ...)

    @property
    def n_dof(self):
        return 3 * self.coords.shape[0]

    @property
    def n_elems(self):
        return self.conn.shape[0]


def _copy(ptr, n, ctype, dtype, shape=None):
    if n == 0:
        return np.zeros((0,) if shape is None else (0,) + shape[1:], dtype)
    arr = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), (n,)).copy().astype(dtype)
    return arr if shape is None else arr.reshape(shape)


def read_inp(path: str) -> Deck:
    """Reads a deck; raises UfcError (code STATUS_UNSUPPORTED for what this
    path does not model, STATUS_INVALID for a malformed deck, with the deck
    line in the message)."""
    L = A.lib()
    m = A.ApModel()
    st = A.Status()
    A.check(L.ufc_ap_read_inp(os.fsencode(path), C.byref(m), C.byref(st)), st)
    try:
        nn, ne = m.n_nodes, m.n_elems
        i32, f64 = C.c_int32, C.c_double
        return Deck(
            node_label=_copy(m.node_label, nn, i32, np.int32),
            elem_label=_copy(m.elem_label, ne, i32, np.int32),
            coords=_copy(m.coords, 3 * nn, f64, np.float64, (nn, 3)),
            conn=_copy(m.conn, 8 * ne, i32, np.int32, (ne, 8)),
            young=_copy(m.young, ne, f64, np.float64),
            nu=float(m.nu),
            bc_eq=_copy(m.bc_eq, m.n_bc, i32, np.int32),
            bc_val=_copy(m.bc_val, m.n_bc, f64, np.float64),
            cload_eq=_copy(m.cload_eq, m.n_cload, i32, np.int32),
            cload_val=_copy(m.cload_val, m.n_cload, f64, np.float64),
            face_elem=_copy(m.face_elem, m.n_face, i32, np.int32),
            face_id=_copy(m.face_id, m.n_face, i32, np.int32),
            face_p=_copy(m.face_p, m.n_face, f64, np.float64),
            variant=int(m.variant),
        )
    finally:
        L.ufc_ap_free(C.byref(m), C.byref(st))


def setup_context(ctx, deck: Deck, penalty=PENALTY):
    """Pattern, F (consistent face pressures, one library call per distinct
    pressure, then the concentrated loads) and the penalty boundary
    conditions of `deck` on `ctx`; assembly is separate
    (ctx.assemble_k(deck.young, deck.nu, precision, variant=deck.variant))."""
    ctx.build_pattern(deck.coords, deck.conn)
    F = np.zeros(deck.n_dof)
    first = True
    for p in np.unique(deck.face_p):
        sel = deck.face_p == p
        ctx.load_pressure(deck.face_elem[sel], deck.face_id[sel], float(p), accumulate=not first)
        first = False
    if not first:
        F = ctx.get_rhs()
    np.add.at(F, deck.cload_eq, deck.cload_val)
    ctx.set_rhs(F)
    ctx.apply_bc(deck.bc_eq, deck.bc_val, penalty)


def write_inp(path: str, pr, part="PART-1", instance="PART-1-1", element="C3D8") -> None:
    """Writes problems.Problem `pr` as a deck: one part with one instance,
    nodes and elements labelled from 1, one elastic material per distinct
    Young's modulus, *BOUNDARY on the prescribed equations (with their
    values), and a pressure face load as *DSLOAD on a *SURFACE.  Traction
    loads (TRVEC) are not part of the deck reader and raise ValueError."""
    if pr.load_elem.size and pr.load_kind != "pressure":
        raise ValueError("write_inp writes pressure loads only")
    lines = [f"** {pr.name}: {pr.coords.shape[0]} nodes, {pr.conn.shape[0]} C3D8",
             f"*PART, NAME={part}", "*NODE"]
    lines += [f"{i + 1}, {x!r}, {y!r}, {z!r}" for i, (x, y, z) in enumerate(pr.coords.tolist())]
    lines.append(f"*ELEMENT, TYPE={element}, ELSET=ALL")
    lines += [f"{e + 1}, " + ", ".join(str(n + 1) for n in row)
              for e, row in enumerate(pr.conn.tolist())]
    young = np.asarray(pr.young, np.float64)
    mats, inv = np.unique(young, return_inverse=True)
    for k in range(mats.size):
        elems = np.nonzero(inv == k)[0] + 1
        lines.append(f"*ELSET, ELSET=E{k + 1}")
        lines += [", ".join(str(v) for v in elems[i:i + 16]) for i in range(0, elems.size, 16)]
        lines.append(f"*SOLID SECTION, ELSET=E{k + 1}, MATERIAL=M{k + 1}")
    if pr.load_elem.size:
        lines.append("*SURFACE, NAME=LOADED, TYPE=ELEMENT")
        lines += [f"{e + 1}, S{f + 1}" for e, f in zip(pr.load_elem.tolist(),
                                                      pr.load_face.tolist())]
    lines.append("*END PART")
    for k, E in enumerate(mats.tolist()):
        lines += [f"*MATERIAL, NAME=M{k + 1}", "*ELASTIC", f"{E!r}, {pr.nu!r}"]
    lines += ["*ASSEMBLY, NAME=ASSEMBLY", f"*INSTANCE, NAME={instance}, PART={part}",
              "*END INSTANCE", "*END ASSEMBLY", "*STEP, NAME=LOAD", "*STATIC",
              "1.0, 1.0", "*BOUNDARY"]
    for eq, v in zip(np.asarray(pr.bc_eq).tolist(), np.asarray(pr.bc_val).tolist()):
        lines.append(f"{instance}.{eq // 3 + 1}, {eq % 3 + 1}, {eq % 3 + 1}, {v!r}")
    if pr.load_elem.size:
        lines += ["*DSLOAD", f"{instance}.LOADED, P, {pr.load_value[0]!r}"]
    lines += ["*OUTPUT, FIELD", "*NODE OUTPUT", "U", "*END STEP"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")
