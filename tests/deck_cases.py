"""Input decks written for the deck-reader tests, with the arrays the reader
must return for them, derived here from the grid that generates the deck
(not from the reader).

The block deck uses every keyword form the reader supports: a part with one
instance, lower/mixed case keywords, CRLF line ends, comments, element lines
continued after a trailing comma, node and element sets by list and by
GENERATE, instance-qualified labels, two materials with one Poisson ratio,
*SOLID SECTION per element set, a *SURFACE on an element set, *BOUNDARY with
ENCASTRE and a dof range with a value, *CLOAD on a set and on a node, *DLOAD
Pn on an element set and *DSLOAD on the surface.
"""
from __future__ import annotations

import numpy as np

NX, NY, NZ = 3, 2, 2
SX, SY = NX + 1, NY + 1


def node_label(i, j, k):
    return 101 + i + SX * (j + SY * k)


def elem_label(i, j, k):
    return 1001 + i + NX * (j + NY * k)


def block_deck(path, extra_step="", replace=None):
    """Writes the block deck to `path`; returns the expected reader output
    (dict of arrays, 0-based).  `replace` maps deck text to substitutes (for
    the error cases); `extra_step` is added before *END STEP."""
    nodes = [(i, j, k) for k in range(NZ + 1) for j in range(NY + 1) for i in range(NX + 1)]
    rng = np.random.default_rng(3)
    xyz = {n: (n[0] * 1.0 + rng.uniform(-0.1, 0.1) * (0 < n[0] < NX),
               n[1] * 0.5, n[2] * 0.75) for n in nodes}
    elems = [(i, j, k) for k in range(NZ) for j in range(NY) for i in range(NX)]

    def corners(i, j, k):
        return [(i, j, k), (i + 1, j, k), (i + 1, j + 1, k), (i, j + 1, k),
                (i, j, k + 1), (i + 1, j, k + 1), (i + 1, j + 1, k + 1), (i, j + 1, k + 1)]

    L = ["** block deck for the reader tests", "*Heading", " block 3x2x2",
         "*Part, name=Block"]
    # nodes in two *NODE blocks, the second in reverse order
    half = len(nodes) // 2
    order = nodes[:half] + nodes[half:][::-1]
    L.append("*Node")
    for n in order[:half]:
        L.append(f"{node_label(*n)}, {xyz[n][0]!r}, {xyz[n][1]!r}, {xyz[n][2]!r}")
    L.append("*NODE, NSET=second")
    for n in order[half:]:
        L.append(f"  {node_label(*n)},{xyz[n][0]!r},{xyz[n][1]!r},{xyz[n][2]!r}")
    L.append("*Element, type=C3D8, elset=Everything")
    for e in elems:
        c = [str(node_label(*n)) for n in corners(*e)]
        if e[0] == 1:  # continued on the next line
            L.append(f"{elem_label(*e)}, " + ", ".join(c[:5]) + ",")
            L.append("    " + ", ".join(c[5:]))
        else:
            L.append(f"{elem_label(*e)}, " + ", ".join(c))
    bottom = [node_label(i, j, 0) for j in range(SY) for i in range(SX)]
    L += ["*Nset, nset=Bottom, generate", f"{bottom[0]}, {bottom[-1]}, 1"]
    tip = [node_label(NX, j, NZ) for j in range(SY)]
    L += ["*Nset, nset=Tip", ", ".join(str(t) for t in tip)]
    top = [elem_label(i, j, NZ - 1) for j in range(NY) for i in range(NX)]
    L += ["*Elset, elset=Top, generate", f"{top[0]}, {top[-1]}"]
    soft = [elem_label(0, 0, 0), elem_label(1, 0, 0)]
    hard = [elem_label(*e) for e in elems if elem_label(*e) not in soft]
    L += ["*Elset, elset=Soft", ", ".join(map(str, soft)), "*ELSET, ELSET=Hard"]
    L += [", ".join(map(str, hard[i:i + 4])) for i in range(0, len(hard), 4)]
    L += ["*Surface, type=ELEMENT, name=TopS", "Top, S2"]
    L += ["*Solid Section, elset=Soft, material=Rubberish", ",",
          "*Solid Section, elset=Hard, material=Steelish", ","]
    L += ["*End Part", "**", "*Material, name=Rubberish", "*Elastic", " 1.0e9, 0.3",
          "*Material, name=Steelish", "*Elastic", " 3.0e9, 0.3"]
    L += ["*Assembly, name=Assembly", "*Instance, name=Blk-1, part=Block", "*End Instance",
          "*End Assembly"]
    L += ["*Step, name=Load, nlgeom=NO", "*Static", "1., 1., 1e-05, 1.",
          "*Boundary", "Blk-1.Bottom, ENCASTRE", "Blk-1.Tip, 1, 1, 0.001",
          "*Cload", "Blk-1.Tip, 3, -500.", f"Blk-1.{node_label(1, 1, 1)}, 2, 250.",
          "*Dload", "Blk-1.Top, P2, 2.0e5",
          "*Dsload", "Blk-1.TopS, P, 1.0e5"]
    if extra_step:
        L.append(extra_step)
    L += ["*Output, field", "*Node Output", "U, RF", "*End Step"]
    text = "\r\n".join(L) + "\r\n"
    for a, b in (replace or {}).items():
        assert a in text, a
        text = text.replace(a, b)
    with open(path, "w", newline="") as fh:
        fh.write(text)

    index = {node_label(*n): q for q, n in enumerate(order)}
    coords = np.array([xyz[n] for n in order])
    conn = np.array([[index[node_label(*n)] for n in corners(*e)] for e in elems], np.int32)
    young = np.array([1.0e9 if elem_label(*e) in soft else 3.0e9 for e in elems])
    bc = {}
    for lab in bottom:
        for d in range(3):
            bc[3 * index[lab] + d] = 0.0
    for lab in tip:
        bc[3 * index[lab]] = 0.001
    bc_eq = np.array(sorted(bc), np.int32)
    cl = [(3 * index[lab] + 2, -500.0) for lab in tip] + [(3 * index[node_label(1, 1, 1)] + 1,
                                                           250.0)]
    elem_index = {elem_label(*e): q for q, e in enumerate(elems)}
    faces = [(elem_index[t], 1, 2.0e5) for t in top] + [(elem_index[t], 1, 1.0e5) for t in top]
    return dict(
        node_label=np.array([node_label(*n) for n in order], np.int32),
        elem_label=np.array([elem_label(*e) for e in elems], np.int32),
        coords=coords, conn=conn, young=young, nu=0.3,
        bc_eq=bc_eq, bc_val=np.array([bc[q] for q in bc_eq]),
        cload_eq=np.array([c[0] for c in cl], np.int32),
        cload_val=np.array([c[1] for c in cl]),
        face_elem=np.array([f[0] for f in faces], np.int32),
        face_id=np.array([f[1] for f in faces], np.int32),
        face_p=np.array([f[2] for f in faces]))
