"""ufc_mesh_box_numbering (host only): the lexicographic numbering of a
structured box of hexes under any node numbering, element order and element
node rotation, recovered from the element edges -- what lets a renumbered
box use the symmetric block-diagonal FP32 view.  Meshes that are not boxes
are refused."""
import numpy as np


def _offsets(conn):
    c = np.asarray(conn, np.int64)
    d = (c[:, None, :] - c[:, :, None]).ravel()
    return np.unique(d[d > 0])


def _check_box(conn, perm, dims):
    nn = perm.size
    assert np.array_equal(np.sort(perm), np.arange(nn))
    new = perm[conn]
    offs = _offsets(new)
    assert offs.size <= 13
    # every element a unit cell of some axis order of the lattice dims
    import itertools
    for NX, NY, NZ in itertools.permutations([d + 1 for d in dims]):
        if NX * NY * NZ != nn:
            continue
        i, j, k = new % NX, (new // NX) % NY, new // (NX * NY)
        ok = all(np.all(np.ptp(a, axis=1) == 1) for a in (i, j, k))
        if ok:
            cells = set()
            for e in range(new.shape[0]):
                cells.add((i[e].min(), j[e].min(), k[e].min()))
            assert len(cells) == new.shape[0]
            return
    raise AssertionError("no lattice order fits")


def _renumber(coords, conn, seed, rotate=True):
    rng = np.random.default_rng(seed)
    nn = coords.shape[0]
    p = rng.permutation(nn)
    conn2 = p[conn][rng.permutation(conn.shape[0])]
    if rotate:
        # the same hex with its bottom/top faces' node order rotated
        r = rng.integers(0, 4, conn2.shape[0])
        rot = np.array([[0, 1, 2, 3], [1, 2, 3, 0], [2, 3, 0, 1], [3, 0, 1, 2]])
        conn2 = np.concatenate([np.take_along_axis(conn2[:, :4], rot[r], 1),
                                np.take_along_axis(conn2[:, 4:], rot[r], 1)], axis=1)
    return conn2.astype(np.int32)


def test_box_numbering_recovers_renumbered_boxes():
    import ufc_hip
    for dims, jit, seed in (((4, 3, 5), 0.0, 1), ((6, 6, 6), 0.1, 2), ((12, 1, 1), 0.0, 3),
                            ((1, 1, 1), 0.0, 4), ((20, 4, 4), 0.0, 5)):
        coords, conn = ufc_hip.mesh_box(*dims, jitter=jit, seed=seed)
        perm = ufc_hip.static.box_numbering(conn, coords.shape[0])
        assert perm is not None, dims
        _check_box(conn, perm, dims)
        conn2 = _renumber(coords, conn, seed)
        assert _offsets(conn2).size > 13 or dims[1] == 1
        perm2 = ufc_hip.static.box_numbering(conn2, coords.shape[0])
        assert perm2 is not None, dims
        _check_box(conn2, perm2, dims)
        # 1-based connectivity gives the same numbering
        perm3 = ufc_hip.static.box_numbering(conn2 + 1, coords.shape[0], index_base=1)
        assert np.array_equal(perm2, perm3)


def test_box_numbering_refuses_other_meshes():
    import ufc_hip
    coords, conn = ufc_hip.mesh_box(6, 5, 4)
    e = np.arange(6 * 5 * 4).reshape(4, 5, 6)
    keep = np.ones_like(e, bool)
    keep[2:, 3:, 4:] = False  # a notch
    notched = conn[e[keep].ravel()]
    used = np.unique(notched)
    remap = np.full(coords.shape[0], -1)
    remap[used] = np.arange(used.size)
    assert ufc_hip.static.box_numbering(remap[notched], used.size) is None
    # two boxes side by side without a shared face
    c2, k2 = ufc_hip.mesh_box(2, 2, 2)
    two = np.concatenate([k2, k2 + c2.shape[0]])
    assert ufc_hip.static.box_numbering(two, 2 * c2.shape[0]) is None
    # an element whose nodes are not a hex of the lattice (two nodes swapped)
    bad = conn.copy()
    bad[7, [0, 6]] = bad[7, [6, 0]]
    assert ufc_hip.static.box_numbering(bad, coords.shape[0]) is None
