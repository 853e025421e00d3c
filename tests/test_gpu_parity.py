"""HIP library vs the FP64 oracle, through the C-ABI (libufc_hip.so).

Tolerances (stated per test): CSR row_ptr/col_idx bit-exact; FP64 matrix
values |dK| <= 1e-12 max|K|; FP32 values within FP32 rounding of the FP64
ones; displacements 1e-8 relative where both sides are driven to a tight
residual.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

UNIT = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                 [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _problem_small(kind):
    from ufc_hip import problems as P
    if kind == "cantilever":
        return P.cantilever()
    if kind == "cube8":
        return P.cube_pressure(8, name="cube8")
    if kind == "jitter6":
        return P.cube_pressure(6, name="jitter6", jitter=0.1)
    if kind == "irregular":
        return _irregular()
    if kind == "notched":
        return _irregular(shuffle=False)
    raise KeyError(kind)


def _n_diagonals(conn):
    """Distinct offsets m - n > 0 over the node pairs that share an element:
    at most 13 allow the symmetric block-diagonal FP32 view."""
    c = np.asarray(conn, np.int64)
    d = (c[:, None, :] - c[:, :, None]).ravel()
    return np.unique(d[d > 0]).size


def _irregular(shuffle=True):
    """Notched, jittered 8x6x5 block with shuffled node and element numbering:
    node degrees from 1 to 8 elements, ragged rows, no structure in the ids.
    shuffle=False keeps the box's node order (compacted) and element order."""
    from ufc_hip import problems as P
    base = P.cube_pressure(1, name="irr")  # for the Problem shape only
    import ufc_hip
    nx, ny, nz = 8, 6, 5
    coords, conn = ufc_hip.mesh_box(nx, ny, nz, 2.0, 1.5, 1.0, 0.1, 77)
    e = np.arange(nx * ny * nz).reshape(nz, ny, nx)
    keep = np.ones_like(e, bool)
    keep[2:, 3:, 4:] = False      # notch through the top corner
    keep[0, 0, nx - 1] = False    # a missing corner element
    conn = conn[e[keep].ravel()]
    used = np.unique(conn)
    rng = np.random.default_rng(5)
    perm = rng.permutation(used.size) if shuffle else np.arange(used.size)
    new_of = np.full(coords.shape[0], -1, np.int64)   # old id used[i] -> new id perm[i]
    new_of[used] = perm
    coords_n = np.empty((used.size, 3))
    coords_n[perm] = coords[used]
    eorder = rng.permutation(conn.shape[0]) if shuffle else np.arange(conn.shape[0])
    conn_n = new_of[conn].astype(np.int32)[eorder]
    z0 = np.nonzero(coords_n[:, 2] == 0.0)[0]
    clamp = (3 * z0[:, None] + np.arange(3)).ravel().astype(np.int32)
    top = np.nonzero(np.all(coords_n[conn_n[:, 4:8], 2] == 1.0, axis=1))[0].astype(np.int32)
    young = 1e9 * (1.0 + rng.random(conn_n.shape[0]))
    base.name = "irregular" if shuffle else "notched"
    base.coords, base.conn, base.young = coords_n, conn_n, young
    base.bc_eq, base.bc_val = clamp, np.zeros(clamp.size)
    base.load_elem, base.load_face = top, np.full(top.size, 1, np.int32)
    return base


def _oracle_system(orc, pr, penalty=1e12, allow_degenerate=False):
    rp, ci = orc.csr_pattern(pr.coords.shape[0], pr.conn)
    val = orc.assemble(pr.conn, pr.coords, pr.young, pr.nu, rp, ci,
                       allow_degenerate=allow_degenerate)
    F = np.zeros(pr.n_dof)
    if pr.load_kind == "traction":
        orc.face_loads_traction(pr.load_elem, pr.load_face, pr.conn, pr.coords, pr.load_value, F)
    else:
        orc.face_loads_pressure(pr.load_elem, pr.load_face, pr.conn, pr.coords,
                                pr.load_value[0], F)
    f0 = F.copy()
    orc.apply_penalty(pr.bc_eq, pr.bc_val, penalty, rp, ci, val, F)
    return rp, ci, val, F, f0


def _difference_form(orc, pr, rp, ci, penalty=1e12):
    val = orc.assemble(pr.conn, pr.coords, pr.young, pr.nu, rp, ci)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    val[(ci // 3) == (rows // 3)] = 0.0
    orc.apply_penalty(pr.bc_eq, pr.bc_val, penalty, rp, ci, val, None)
    return val


def test_ke_batch_fp64_fp32(ctx, orc):
    rng = np.random.default_rng(0)
    n = 200
    xe = UNIT[None] * rng.uniform(0.5, 2.0, (n, 1, 3)) + 0.15 * rng.uniform(-1, 1, (n, 8, 3))
    E = rng.uniform(1e9, 2e11, n)
    say("eval_ke n", n)
    K = ctx.eval_ke(xe, E, 0.3)
    Ko = np.stack([orc.ke(xe[i], E[i], 0.3) for i in range(n)])
    err = np.abs(K - Ko).max(axis=(1, 2)) / np.abs(Ko).max(axis=(1, 2))
    assert err.max() < 1e-12
    K32 = ctx.eval_ke(xe, E, 0.3, precision=1)
    err32 = np.abs(K32 - Ko).max(axis=(1, 2)) / np.abs(Ko).max(axis=(1, 2))
    assert err32.max() < 2e-5


@pytest.mark.parametrize("mesh", ["1x1x1", "2x1x1", "3x2x2", "20x4x4"])
def test_pattern_bitexact_vs_reference(ctx, golden_dir, mesh):
    import ufc_hip
    g = np.load(os.path.join(golden_dir, f"ref_csr_{mesh}.npz"))
    nx, ny, nz = (int(v) for v in g["mesh"])
    coords, conn = ufc_hip.mesh_box(nx, ny, nz, index_base=1)
    say("pattern", mesh)
    info = ctx.build_pattern(coords, conn, index_base=1)
    assert info.nnz == g["ja"].size
    rp, ci = ctx.get_pattern(index_base=1)
    assert np.array_equal(rp, g["ia"]) and np.array_equal(ci, g["ja"])


@pytest.mark.parametrize("kind", ["cantilever", "cube8", "jitter6", "irregular"])
def test_assembly_fp64_fp32_vs_oracle(ctx, orc, kind):
    from ufc_hip import problems as P
    pr = _problem_small(kind)
    rp_o, ci_o, val_o, F_o, _ = _oracle_system(orc, pr)
    say("assembly", kind, "dof", pr.n_dof)
    P.setup_context(ctx, pr, 0)
    rp, ci = ctx.get_pattern()
    assert np.array_equal(rp, rp_o) and np.array_equal(ci, ci_o)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    v64 = ctx.get_values(0)
    scale = np.abs(val_o).max()
    assert np.abs(v64 - val_o).max() <= 1e-12 * scale
    assert np.array_equal(ctx.get_rhs(), F_o) or np.abs(ctx.get_rhs() - F_o).max() < 1e-12 * np.abs(F_o).max()
    # bitwise run-to-run determinism of the atomics-free scatter
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    assert np.array_equal(ctx.get_values(0), v64)
    ctx.assemble_k(pr.young, pr.nu, precision=1)
    v32 = ctx.get_values(1).astype(np.float64)
    # FP32 values are kept in difference form: self 3x3 blocks hold only the
    # penalty diagonal (every other block as assembled, rounded to FP32)
    expect = _difference_form(orc, pr, rp_o, ci_o)
    assert np.abs(v32 - expect).max() <= 1e-6 * scale
    # and the block rows of the FP64 K sum to zero (what makes that exact)
    node_of_row = np.repeat(np.arange(pr.coords.shape[0]), 3)
    rows = np.repeat(np.arange(rp_o.size - 1), np.diff(rp_o))
    v_nopen = orc.assemble(pr.conn, pr.coords, pr.young, pr.nu, rp_o, ci_o)
    rowsum = np.zeros((rp_o.size - 1, 3))
    np.add.at(rowsum, (rows, ci_o % 3), v_nopen)
    assert np.abs(rowsum).max() <= 1e-12 * scale


def test_spmv_fp64_fp32(ctx, orc):
    from ufc_hip import problems as P
    pr = _problem_small("cube8")
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    x = np.random.default_rng(1).normal(size=pr.n_dof)
    y = ctx.spmv(x, precision=0)
    yo = orc.spmv(rp, ci, val, x)
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()
    ctx.assemble_k(pr.young, pr.nu, precision=1)
    y32 = ctx.spmv(x, precision=1)
    assert np.abs(y32 - yo).max() <= 1e-5 * np.abs(yo).max()


@pytest.mark.parametrize("kind", ["cantilever", "jitter6", "irregular"])
def test_pcg_fp64_vs_oracle(ctx, orc, kind):
    from ufc_hip import problems as P
    pr = _problem_small(kind)
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    uo, st, it_o, rn_o = orc.pcg(rp, ci, val, F, rtol=1e-10)
    assert st == 0
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    say("pcg fp64", kind)
    u, info = ctx.pcg(precision=0, rtol=1e-10)
    assert info.status == 0 and info.converged and info.rnorm < 1e-10
    # the same algorithm: the stopping iteration matches up to rounding drift
    assert abs(info.niter - it_o) <= max(2, it_o // 100)
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8
    # both driven to a tight residual: displacements agree to 1e-8 (here ~1e-12)
    uo2, _, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-14, max_iter=200000)
    u2, info2 = ctx.pcg(precision=0, rtol=1e-14, max_iter=200000)
    assert np.linalg.norm(u2 - uo2) / np.linalg.norm(uo2) < 1e-10


@pytest.mark.parametrize("kind,replace_rtol,view", [("cube12", 1.0, None), ("cube12", 0.1, None),
                                                    ("cube12", 0.1, "pattern"),
                                                    ("irregular", 0.1, None),
                                                    ("notched", 0.1, None),
                                                    ("cantilever", 0.1, None),
                                                    ("cantilever", 0.1, "pattern")])
def test_pcg_mixed_refinement_vs_oracle(ctx, orc, kind, replace_rtol, view):
    """FP32 inner PCG in FP64 refinement, both forms: restarted steps to
    inner_rtol (replace_rtol >= 1), and residual replacement every 0.1 with
    the direction kept.  The structured cube and cantilever run the symmetric
    block-diagonal view by default and the stencil-pattern view when asked;
    the notched block in box order runs the stencil-pattern view with node
    degrees 8 to 27; the cantilever has its clamped (penalty) nodes spread
    over the numbering (x = 0 face)."""
    from ufc_hip import _abi as A
    from ufc_hip import problems as P
    pr = P.cube_pressure(12, name="cube12") if kind == "cube12" else _problem_small(kind)
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    uo, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-14, max_iter=200000)
    P.setup_context(ctx, pr, 2)
    if kind in ("cube12", "cantilever"):
        assert _n_diagonals(pr.conn) == 13
        assert ctx.fp32_view == A.FP32_VIEW_SYMDIAG and ctx.nb_index_bytes == 0, kind
    if kind == "notched":
        assert _n_diagonals(pr.conn) > 13
        assert ctx.fp32_view == A.FP32_VIEW_PATTERN and ctx.nb_index_bytes == 1, kind
    if view == "pattern":
        assert ctx.set_fp32_view(A.FP32_VIEW_PATTERN) == A.FP32_VIEW_PATTERN
        assert ctx.nb_index_bytes == 1
    ctx.assemble_k(pr.young, pr.nu, precision=2)
    say("pcg mixed", kind, "replace_rtol", replace_rtol, "view", ctx.fp32_view)
    u, info = ctx.pcg(precision=2, rtol=1e-10, inner_rtol=1e-5, replace_rtol=replace_rtol)
    assert info.status == 0 and info.rnorm < 1e-10 and info.n_outer >= 2
    # true FP64 residual of the oracle's operator
    K_u = orc.spmv(rp, ci, val, u)
    assert np.linalg.norm(F - K_u) / np.linalg.norm(F) < 2e-10
    err = np.linalg.norm(u - uo) / np.linalg.norm(uo)
    say("mixed err vs oracle", err, "outer", info.n_outer, "inner", info.niter)
    assert err < 1e-8


def test_pcg_status_codes(ctx):
    from ufc_hip import problems as P
    pr = _problem_small("cantilever")
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    u, info = ctx.pcg(precision=0, rtol=1e-10, max_iter=5)
    assert info.status == 1 and not info.converged and info.niter == 5
    # negative modulus: p.Ap <= 0 breakdown is reported, not a NaN result
    ctx.assemble_k(-pr.young, pr.nu, precision=0)
    ctx.apply_bc(pr.bc_eq, pr.bc_val, -1e12)
    u, info = ctx.pcg(precision=0, rtol=1e-10, max_iter=50)
    assert info.status == 2 and not info.converged


def test_adjoint_and_sensitivity_vs_oracle_and_fd(ctx, orc):
    from ufc_hip import problems as P
    pr = P.simp_cube(4)
    rho = pr.extra["rho"]
    rp, ci, val, F, f0 = _oracle_system(orc, pr)
    uo, _, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-14, max_iter=100000)
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    u, info = ctx.pcg(precision=0, rtol=1e-13)
    lam, ainfo = ctx.adjoint_solve(f0, precision=0, rtol=1e-13)
    # reference adjoint algorithm (NM_CG_Solve_Transpose) on the same K
    lo, st, _, _ = orc.cg_transpose(rp, ci, val, f0, tol=1e-13, max_iter=100000)
    assert st == 0 and np.linalg.norm(lam - lo) / np.linalg.norm(lo) < 1e-9
    dedrho = 3 * rho ** 2 * 1e9
    g = ctx.sensitivity(u, lam, dedrho, pr.nu)
    go = orc.sensitivity(pr.conn, pr.coords, pr.nu, dedrho, uo, lo)
    assert np.abs(g - go).max() < 1e-9 * np.abs(go).max()


def test_j2_update_and_newton_vs_oracle(ctx, orc):
    import ufc_hip
    from ufc_hip import problems as P
    coords, conn = ufc_hip.mesh_box(3, 3, 3, jitter=0.1, seed=9)
    bottom = np.nonzero(coords[:, 2] == 0)[0]
    top = np.nonzero(coords[:, 2] == 1)[0]
    eq = np.concatenate([(3 * bottom[:, None] + np.arange(3)).ravel(), 3 * top + 2]).astype(np.int32)
    props = (200e9, 0.3, 250e6, 2e9)
    ctx.build_pattern(coords, conn)
    ctx.j2_init(*props)
    n = 3 * coords.shape[0]
    u_n = np.zeros(n)
    state = np.zeros(conn.shape[0] * 8 * 13)
    uo = np.zeros(n)
    for inc in (1, 2, 3):
        val = np.concatenate([np.zeros(3 * bottom.size), np.full(top.size, 1e-3 * inc)])
        st, uo, state, nn, rn = orc.j2_newton_step(coords.shape[0], conn, coords, props, eq,
                                                   val, 1e12, state, uo)
        assert st == 0
        ctx.apply_bc(eq, val, 1e12)
        u = u_n.copy()
        for it in range(25):
            ctx.j2_update(u, u_n)
            r, norms = ctx.residual(u)
            if it > 0 and norms[0] <= 1e-8 * norms[1]:
                break
            ctx.assemble_k(None, 0.3, precision=0, material=1)
            du, info = ctx.pcg(precision=0, rtol=1e-12, b=r)
            u = u + du
        ctx.j2_commit()
        u_n = u.copy()
        say("j2 inc", inc, "newton", it, "oracle", nn)
        assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-7
    sg = ctx.j2_get_state().ravel()
    so = state
    assert np.abs(sg - so).max() < 1e-6 * np.abs(so).max()


def test_pcg_zero_rhs_and_converged_start(ctx, orc):
    """b = 0 gives x = 0 with no iteration (NM_Solver_PCG would form 0/0);
    a start at the solution returns it unchanged with no iteration."""
    from ufc_hip import problems as P
    pr = _problem_small("cube8")
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    zero = np.zeros(pr.n_dof)
    u, info = ctx.pcg(precision=0, rtol=1e-10, b=zero)
    assert info.status == 0 and info.converged and info.niter == 0 and not np.any(u)
    u_ref, info_ref = ctx.pcg(precision=0, rtol=1e-13)
    u2, info2 = ctx.pcg(precision=0, rtol=1e-10, x0=u_ref)
    assert info2.converged and info2.niter == 0 and np.array_equal(u2, u_ref)
    ctx.assemble_k(pr.young, pr.nu, precision=2)
    u3, info3 = ctx.pcg(precision=2, rtol=1e-10, b=zero)
    assert info3.converged and info3.niter == 0 and not np.any(u3)


def test_pattern_rejects_invalid_meshes(ctx):
    """Host-side validation of the mesh: status codes, no device work."""
    import ufc_hip
    from ufc_hip import _abi as A
    coords, conn = ufc_hip.mesh_box(2, 2, 2)
    bad = conn.copy()
    bad[3, 5] = coords.shape[0]            # node id out of range
    with pytest.raises(A.UfcError) as e:
        ctx.build_pattern(coords, bad)
    assert e.value.code == A.STATUS_INVALID
    bad = conn.copy()
    bad[1, 2] = bad[1, 0]                  # element repeats a node
    with pytest.raises(A.UfcError) as e:
        ctx.build_pattern(coords, bad)
    assert e.value.code == A.STATUS_INVALID
    # node 0 in 40 more elements of 7 fresh nodes each: 1 + 27 + 280 neighbours,
    # beyond the general nodes' 256
    nn = coords.shape[0]
    fresh = nn + np.arange(40 * 7).reshape(40, 7)
    many = np.concatenate([conn, np.column_stack([np.zeros(40, np.int32), fresh])]).astype(np.int32)
    far = np.random.default_rng(0).uniform(3, 4, (fresh.size, 3))
    with pytest.raises(A.UfcError) as e:
        ctx.build_pattern(np.concatenate([coords, far]), many)
    assert e.value.code == A.STATUS_UNSUPPORTED and "256 neighbours" in e.value.message
    iso = np.concatenate([coords, [[9.0, 9.0, 9.0]]])   # a node in no element
    with pytest.raises(A.UfcError) as e:
        ctx.build_pattern(iso, conn)
    assert e.value.code == A.STATUS_INVALID
    # the context stays usable after the rejections
    info = ctx.build_pattern(coords, conn)
    assert info.nnz > 0


def test_fp32_view_wide_numbering_vs_oracle(ctx, orc):
    """A 32^3 cube with randomly permuted node ids (bandwidth > 2^15): by
    default the symmetric view in the box's own lexicographic numbering;
    forced to the views of the given numbering it takes 4-byte neighbour ids.
    Both SpMVs and mixed solves match the oracle."""
    from ufc_hip import problems as P
    pr = P.cube_pressure(32, name="cube32perm")
    perm = np.random.default_rng(11).permutation(pr.coords.shape[0])
    coords = np.empty_like(pr.coords)
    coords[perm] = pr.coords
    pr.conn = perm[pr.conn].astype(np.int32)
    pr.coords = coords
    pr.bc_eq = (3 * perm[pr.bc_eq // 3] + pr.bc_eq % 3).astype(np.int32)
    from ufc_hip import _abi as A
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    uo, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-13, max_iter=200000)
    x = np.random.default_rng(2).normal(size=pr.n_dof)
    yo = orc.spmv(rp, ci, val, x)
    P.setup_context(ctx, pr, 2)
    assert ctx.fp32_view == A.FP32_VIEW_SYMDIAG and ctx.nb_index_bytes == 0
    for view in (A.FP32_VIEW_SYMDIAG, A.FP32_VIEW_ID32):
        assert ctx.set_fp32_view(view) == view
        ctx.assemble_k(pr.young, pr.nu, precision=1)
        y32 = ctx.spmv(x, precision=1)
        assert np.abs(y32 - yo).max() <= 1e-5 * np.abs(yo).max(), view
        ctx.assemble_k(pr.young, pr.nu, precision=2)
        say("wide numbering mixed solve, view", view)
        u, info = ctx.pcg(precision=2, rtol=1e-10)
        assert info.converged
        assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8, view
    assert ctx.nb_index_bytes == 4
    # the structured numbering takes the symmetric block-diagonal view (13
    # diagonals); the stencil-pattern view (27 patterns) is there on request
    P.setup_context(ctx, P.cube_pressure(8, name="cube8"), 2)
    assert ctx.fp32_view == A.FP32_VIEW_SYMDIAG and ctx.nb_index_bytes == 0
    assert ctx.set_fp32_view(A.FP32_VIEW_PATTERN) == A.FP32_VIEW_PATTERN
    assert ctx.nb_index_bytes == 1


def _renumbered(pr, perm):
    coords = np.empty_like(pr.coords)
    coords[perm] = pr.coords
    pr.conn = perm[pr.conn].astype(np.int32)
    pr.coords = coords
    pr.bc_eq = (3 * perm[pr.bc_eq // 3] + pr.bc_eq % 3).astype(np.int32)
    return pr


def test_fp32_views_vs_oracle(ctx, orc):
    """Neighbour indexes of the FP32 view on one 16^3 cube: the structured
    numbering (13 block diagonals, each off-diagonal block stored once; and,
    on request, one stencil-pattern byte per node, 27 patterns), a numbering
    shuffled inside windows of 64 ids and a random numbering (both more than
    256 patterns, offsets below 2^15: 16-bit offsets on request; by default
    the symmetric view in the box's lexicographic numbering; the 4-byte ids
    are covered on a 32^3 cube above).  Each SpMV, and the mixed solve,
    matches the oracle; the shuffled numberings give the structured solution
    back after un-permuting."""
    from ufc_hip import _abi as A
    from ufc_hip import problems as P
    n = 16
    base = P.cube_pressure(n, name="cube16")
    nn = base.coords.shape[0]
    rng = np.random.default_rng(5)
    window = np.concatenate([w + rng.permutation(min(64, nn - w)) for w in range(0, nn, 64)])
    rand = rng.permutation(nn)
    cases = [("symdiag", np.arange(nn), 0, None),
             ("pattern", np.arange(nn), 1, A.FP32_VIEW_PATTERN),
             ("window", window, 2, A.FP32_VIEW_OFF16), ("random", rand, 2, A.FP32_VIEW_OFF16),
             ("window-box", window, 0, None), ("random-box", rand, 0, None)]
    ref = None
    for name, perm, width, force in cases:
        pr = _renumbered(P.cube_pressure(n, name="cube16"), perm)
        rp, ci, val, F, _ = _oracle_system(orc, pr)
        say("fp32 view", name)
        P.setup_context(ctx, pr, 2)
        if force is not None:
            ctx.set_fp32_view(force)
        assert ctx.nb_index_bytes == width, name
        ctx.assemble_k(pr.young, pr.nu, precision=1)
        x = np.random.default_rng(2).normal(size=pr.n_dof)
        yo = orc.spmv(rp, ci, val, x)
        y32 = ctx.spmv(x, precision=1)
        assert np.abs(y32 - yo).max() <= 1e-5 * np.abs(yo).max(), name
        ctx.assemble_k(pr.young, pr.nu, precision=2)
        u, info = ctx.pcg(precision=2, rtol=1e-10)
        assert info.converged, name
        uo, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-13, max_iter=200000)
        assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8, name
        u_nat = u.reshape(-1, 3)[perm].ravel()
        if ref is None:
            ref = u_nat
        else:
            assert np.linalg.norm(u_nat - ref) / np.linalg.norm(ref) < 1e-8, name


def test_fp32_symdiag_view(ctx, orc):
    """The symmetric block-diagonal FP32 view.  A beam one element thick in y
    and z has 10 distinct neighbour offsets (fewer than the 13 the kernel
    unrolls: the missing diagonals must contribute nothing); its SpMV and
    mixed solve match the oracle.  On a cube the view is switched after the
    FP32 assembly: the stencil-pattern view's K x agrees to FP32 rounding,
    and switching back gives the symmetric view's K x bit for bit.  A mesh
    that is no box (notched, shuffled) rejects the view
    (UFC_STATUS_UNSUPPORTED), an unknown view id is invalid, and the
    rejections leave the view unchanged."""
    from ufc_hip import _abi as A
    from ufc_hip import problems as P
    pr = P.cantilever(nx=12, ny=1, nz=1)
    assert _n_diagonals(pr.conn) == 10
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    P.setup_context(ctx, pr, 2)
    assert ctx.fp32_view == A.FP32_VIEW_SYMDIAG
    ctx.assemble_k(pr.young, pr.nu, precision=1)
    x = np.random.default_rng(3).normal(size=pr.n_dof)
    yo = orc.spmv(rp, ci, val, x)
    y32 = ctx.spmv(x, precision=1)
    say("symdiag beam spmv", np.abs(y32 - yo).max() / np.abs(yo).max())
    assert np.abs(y32 - yo).max() <= 1e-5 * np.abs(yo).max()
    ctx.assemble_k(pr.young, pr.nu, precision=2)
    u, info = ctx.pcg(precision=2, rtol=1e-10)
    assert info.converged
    uo, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-13, max_iter=200000)
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8

    pr = P.cube_pressure(10, name="cube10")
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    P.setup_context(ctx, pr, 2)
    ctx.assemble_k(pr.young, pr.nu, precision=1)
    x = np.random.default_rng(4).normal(size=pr.n_dof)
    yo = orc.spmv(rp, ci, val, x)
    y_sym = ctx.spmv(x, precision=1)
    assert ctx.set_fp32_view(A.FP32_VIEW_PATTERN) == A.FP32_VIEW_PATTERN
    y_pat = ctx.spmv(x, precision=1)
    assert ctx.set_fp32_view(-1) == A.FP32_VIEW_SYMDIAG
    y_sym2 = ctx.spmv(x, precision=1)
    scale = np.abs(yo).max()
    assert np.abs(y_sym - yo).max() <= 1e-5 * scale
    assert np.abs(y_pat - yo).max() <= 1e-5 * scale
    assert np.array_equal(y_sym, y_sym2)
    with pytest.raises(A.UfcError) as e:
        ctx.set_fp32_view(7)
    assert e.value.code == A.STATUS_INVALID
    assert ctx.info().fp32_view == A.FP32_VIEW_SYMDIAG

    pr = _irregular()  # notched, shuffled: no box numbering exists
    P.setup_context(ctx, pr, 2)
    before = ctx.fp32_view
    assert before in (A.FP32_VIEW_OFF16, A.FP32_VIEW_PATTERN)
    with pytest.raises(A.UfcError) as e:
        ctx.set_fp32_view(A.FP32_VIEW_SYMDIAG)
    assert e.value.code == A.STATUS_UNSUPPORTED
    assert ctx.info().fp32_view == before


@pytest.mark.parametrize("shape", ["cube24", "slab48x48x6", "slab40x40x2"])
def test_fp32_symdiag_view_tile_blocks(ctx, orc, shape):
    """The symmetric view's tile order over planes of F tiles in blocks of 16
    planes, the last block holding the planes that remain: a 24^3 cube (62
    tiles, F = 3: 21 planes, one full block and a block of 5), a 48 x 48 x 6
    slab (one GPU's share of an 8-way z-split of a 48^3 cube: 66 tiles, F =
    10: a block of 7 planes only) and a 40 x 40 x 2 slab (20 tiles, F = 7: a
    block of 3 planes); SpMV and the mixed solve match the oracle."""
    from ufc_hip import _abi as A
    from ufc_hip import problems as P
    if shape == "cube24":
        pr = P.cube_pressure(24, name="cube24")
    elif shape == "slab48x48x6":
        pr = P.slab_share(48, 8)
    else:
        pr = P.slab_share(40, 20)
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    P.setup_context(ctx, pr, 2)
    assert ctx.fp32_view == A.FP32_VIEW_SYMDIAG
    ctx.assemble_k(pr.young, pr.nu, precision=1)
    x = np.random.default_rng(8).normal(size=pr.n_dof)
    yo = orc.spmv(rp, ci, val, x)
    y32 = ctx.spmv(x, precision=1)
    say("symdiag", shape, "spmv", np.abs(y32 - yo).max() / np.abs(yo).max())
    assert np.abs(y32 - yo).max() <= 1e-5 * np.abs(yo).max()
    ctx.assemble_k(pr.young, pr.nu, precision=2)
    u, info = ctx.pcg(precision=2, rtol=1e-10)
    assert info.converged
    uo, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-13, max_iter=200000)
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8


@pytest.mark.parametrize("precision", [0, 2])
def test_pcg_progress_record_matches_polling(ctx, precision):
    """The default host loop (the GPU reports each iteration into pinned host
    memory, three iterations queued) and fixed polling every 7 iterations run
    the same kernels: bit-identical solutions and iteration counts.  Event
    timing of every 5th SpMV samples about a fifth of the launches."""
    from ufc_hip import problems as P
    pr = P.cube_pressure(10, name="cube10")
    P.setup_context(ctx, pr, precision)
    ctx.assemble_k(pr.young, pr.nu, precision=precision)
    say("progress record vs polling, precision", precision)
    ctx.enable_timing(5)
    u0, i0 = ctx.pcg(precision=precision, rtol=1e-10, replace_rtol=0.1 if precision else 0.0)
    ctx.enable_timing(0)
    u1, i1 = ctx.pcg(precision=precision, rtol=1e-10, check_every=7,
                     replace_rtol=0.1 if precision else 0.0)
    assert i0.converged and i1.converged
    assert i0.niter == i1.niter and i0.n_outer == i1.n_outer
    assert np.array_equal(u0, u1)
    assert 0.15 * i0.niter <= i0.spmv_calls <= 0.25 * i0.niter
    if precision == 0:
        # max_iter stops at exactly max_iter iterations in both loops
        for ce in (0, 3):
            _, im = ctx.pcg(precision=0, rtol=1e-14, max_iter=11, check_every=ce)
            assert im.status == 1 and im.niter == 11


def test_fp32_symdiag_box_numbering_vs_oracle(ctx, orc):
    """A 12^3 cube with random node ids, shuffled element order and rotated
    element node order: the library finds the box's lexicographic numbering
    and runs the symmetric view in it.  FP32 SpMV, FP32 PCG and both mixed
    forms match the oracle; a boundary-condition change after assembly
    (penalties re-applied in the box numbering) solves again to the oracle;
    switching to the 16-bit-offset view of the given numbering and back
    gives K x bit for bit."""
    from ufc_hip import _abi as A
    from ufc_hip import problems as P
    pr = P.cube_pressure(12, name="cube12box")
    nn = pr.coords.shape[0]
    rng = np.random.default_rng(21)
    perm = rng.permutation(nn)
    pr = _renumbered(pr, perm)
    eorder = rng.permutation(pr.n_elems)
    rot = np.array([[0, 1, 2, 3], [1, 2, 3, 0], [2, 3, 0, 1], [3, 0, 1, 2]])
    r = rng.integers(0, 4, pr.n_elems)
    conn = pr.conn[eorder]
    pr.conn = np.concatenate([np.take_along_axis(conn[:, :4], rot[r], 1),
                              np.take_along_axis(conn[:, 4:], rot[r], 1)], 1).astype(np.int32)
    pr.young = pr.young[eorder]
    # the loaded faces follow the elements (face 1, zeta = +1, of the
    # rotated element is the same face)
    inv = np.empty_like(eorder)
    inv[eorder] = np.arange(eorder.size)
    pr.load_elem = inv[pr.load_elem].astype(np.int32)
    assert _n_diagonals(pr.conn) > 13
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    uo, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-13, max_iter=200000)
    P.setup_context(ctx, pr, 2)
    assert ctx.fp32_view == A.FP32_VIEW_SYMDIAG and ctx.nb_index_bytes == 0
    ctx.assemble_k(pr.young, pr.nu, precision=1)
    x = np.random.default_rng(4).normal(size=pr.n_dof)
    yo = orc.spmv(rp, ci, val, x)
    y32 = ctx.spmv(x, precision=1)
    say("box-numbered symdiag spmv", np.abs(y32 - yo).max() / np.abs(yo).max())
    assert np.abs(y32 - yo).max() <= 1e-5 * np.abs(yo).max()
    u1, i1 = ctx.pcg(precision=1, rtol=1e-5)
    assert i1.converged
    assert np.linalg.norm(u1 - uo) / np.linalg.norm(uo) < 1e-3
    ctx.assemble_k(pr.young, pr.nu, precision=2)
    for repl in (1.0, 0.1):
        u, info = ctx.pcg(precision=2, rtol=1e-10, replace_rtol=repl)
        err = np.linalg.norm(u - uo) / np.linalg.norm(uo)
        say("box-numbered mixed", repl, info.niter, err)
        assert info.converged and err < 1e-8
    # penalties re-applied after assembly: clamp z = 0 and also x = 0
    x0 = np.nonzero(pr.coords[:, 0] == 0.0)[0]
    eq2 = np.unique(np.concatenate([pr.bc_eq, (3 * x0[:, None] + np.arange(3)).ravel()]))
    ctx.apply_bc(eq2.astype(np.int32), None, 1e12)
    rp2, ci2 = rp, ci
    val2 = orc.assemble(pr.conn, pr.coords, pr.young, pr.nu, rp2, ci2)
    F2 = np.zeros(pr.n_dof)
    orc.face_loads_pressure(pr.load_elem, pr.load_face, pr.conn, pr.coords, pr.load_value[0], F2)
    orc.apply_penalty(eq2.astype(np.int32), np.zeros(eq2.size), 1e12, rp2, ci2, val2, F2)
    uo2, _, _, _ = orc.pcg(rp2, ci2, val2, F2, rtol=1e-13, max_iter=200000)
    u2, info2 = ctx.pcg(precision=2, rtol=1e-10, replace_rtol=0.1)
    assert info2.converged and np.linalg.norm(u2 - uo2) / np.linalg.norm(uo2) < 1e-8
    y_box = ctx.spmv(x, precision=1)
    assert ctx.set_fp32_view(A.FP32_VIEW_OFF16) == A.FP32_VIEW_OFF16
    y16 = ctx.spmv(x, precision=1)
    yo2 = orc.spmv(rp2, ci2, val2, x)
    assert np.abs(y16 - yo2).max() <= 1e-5 * np.abs(yo2).max()
    assert ctx.set_fp32_view(-1) == A.FP32_VIEW_SYMDIAG
    assert np.array_equal(ctx.spmv(x, precision=1), y_box)


@pytest.mark.parametrize("view", ["symdiag", "pattern", "id32"])
def test_fp32_bc_change_after_assembly(ctx, orc, view):
    """A new penalty set after the FP32 assembly: the FP32 self diagonals of
    equations no longer constrained return to exactly zero (the difference
    form holds only the penalty there), and the FP32 product and the mixed
    solve follow the new set (oracle with the new set: K x 1e-5, u 1e-8)."""
    from ufc_hip import _abi as A
    from ufc_hip import problems as P
    pr = P.cube_pressure(8, name="cube8")
    P.setup_context(ctx, pr, 2)
    views = {"symdiag": A.FP32_VIEW_SYMDIAG, "pattern": A.FP32_VIEW_PATTERN,
             "id32": A.FP32_VIEW_ID32}
    assert ctx.set_fp32_view(views[view]) == views[view]
    ctx.assemble_k(pr.young, pr.nu, precision=2)
    # z = 0 clamp -> x = 0 clamp plus one z-equation of the old set
    x0 = np.nonzero(pr.coords[:, 0] == 0.0)[0]
    new_eq = np.unique(np.concatenate([(3 * x0[:, None] + np.arange(3)).ravel(),
                                       pr.bc_eq[:1]])).astype(np.int32)
    ctx.apply_bc(new_eq, np.zeros(new_eq.size), 1e12)
    say("fp32 bc change, view", view)
    rp, ci = ctx.get_pattern()
    v32 = ctx.get_values(1)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    diag = v32[rows == ci]
    expect = np.zeros(pr.n_dof, np.float32)
    expect[new_eq] = np.float32(1e12)
    assert np.array_equal(diag, expect)
    pr2 = P.cube_pressure(8, name="cube8")
    pr2.bc_eq, pr2.bc_val = new_eq, np.zeros(new_eq.size)
    rp_o, ci_o, val_o, F_o, _ = _oracle_system(orc, pr2)
    x = np.random.default_rng(3).normal(size=pr.n_dof)
    yo = orc.spmv(rp_o, ci_o, val_o, x)
    assert np.abs(ctx.spmv(x, precision=1) - yo).max() <= 1e-5 * np.abs(yo).max()
    uo, st, _, _ = orc.pcg(rp_o, ci_o, val_o, F_o, rtol=1e-14, max_iter=200000)
    u, info = ctx.pcg(precision=2, rtol=1e-12, max_iter=200000, replace_rtol=0.1)
    assert info.status == 0 and np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8


@pytest.mark.parametrize("kind", ["cube8", "polar"])
def test_fp64_block_view_bitwise_equals_csr(ctx, orc, kind):
    """The FP64 node-block layout (one neighbour id per 3x3 block) gives the
    CSR SpMV's bits (same products, same order), so the FP64 PCG run on it is
    bit-identical too; the polar mesh (33 neighbours) runs its long-list form."""
    from ufc_hip import _abi as A
    from ufc_hip import problems as P
    pr = P.cube_pressure(8, name="cube8") if kind == "cube8" else P.polar_prism(5, 3, 4)
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    say("fp64 block view", kind)
    x = np.random.default_rng(5).normal(size=pr.n_dof)
    y_csr = ctx.spmv(x, precision=0)
    u_csr, i_csr = ctx.pcg(precision=0, rtol=1e-12, max_iter=100000)
    ctx.set_fp64_view(A.FP64_VIEW_BLOCK)
    y_blk = ctx.spmv(x, precision=0)
    u_blk, i_blk = ctx.pcg(precision=0, rtol=1e-12, max_iter=100000)
    assert np.array_equal(y_blk, y_csr)
    assert i_blk.niter == i_csr.niter and np.array_equal(u_blk, u_csr)
    yo = orc.spmv(rp, ci, val, x)
    assert np.abs(y_blk - yo).max() <= 1e-12 * np.abs(yo).max()
    with pytest.raises(A.UfcError) as e:
        ctx.set_fp64_view(2)
    assert e.value.code == A.STATUS_INVALID


@pytest.mark.parametrize("nx", [140, 150])
def test_fp32_symdiag_x_window_limit(ctx, orc, nx):
    """The symmetric SpMV stages its tile's x window in LDS when the 27
    offsets fit in four runs of at most 1664 nodes: an nx x 3 x 3 beam (rows of
    nx + 1 nodes) needs three runs of 2 (nx + 2) + 256 nodes, 1620 at nx = 140
    (window) and 1680 at nx = 150 (global gathers).  Both forms match the
    oracle's K x and solve to its displacements.  The wall is clamped along
    its base (z = 0) and pressed on its top, so no node is more than three
    elements from the clamp and the solves take a few hundred iterations."""
    import ufc_hip
    from ufc_hip import _abi as A
    from ufc_hip import problems as P
    pr = P.cube_pressure(1, name=f"wall{nx}")
    coords, conn = ufc_hip.mesh_box(nx, 3, 3, nx / 4.0, 0.75, 0.75)
    z0 = np.nonzero(coords[:, 2] == 0.0)[0]
    top = np.nonzero(np.all(coords[conn[:, 4:8], 2] == 0.75, axis=1))[0].astype(np.int32)
    pr.coords, pr.conn = coords, conn
    pr.young = np.full(conn.shape[0], 1e9)
    pr.bc_eq = (3 * z0[:, None] + np.arange(3)).ravel().astype(np.int32)
    pr.bc_val = np.zeros(pr.bc_eq.size)
    pr.load_elem, pr.load_face = top, np.full(top.size, 1, np.int32)
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    P.setup_context(ctx, pr, 2)
    assert ctx.fp32_view == A.FP32_VIEW_SYMDIAG
    ctx.assemble_k(pr.young, pr.nu, precision=1)
    x = np.random.default_rng(8).normal(size=pr.n_dof)
    yo = orc.spmv(rp, ci, val, x)
    say("symdiag x window, nx", nx)
    assert np.abs(ctx.spmv(x, precision=1) - yo).max() <= 1e-5 * np.abs(yo).max()
    ctx.assemble_k(pr.young, pr.nu, precision=2)
    u, info = ctx.pcg(precision=2, rtol=1e-12, max_iter=20000, replace_rtol=0.1)
    uo, st, it, _ = orc.pcg(rp, ci, val, F, rtol=1e-14, max_iter=20000)
    say("wall iterations: oracle", it, "gpu inner", info.niter)
    assert st == 0 and info.converged
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8


def test_degenerate_and_inverted_elements_vs_oracle(ctx, orc):
    """Gauss points with |det J| <= 1e-12 are skipped (PH_Elem_C3D8.f90:775)
    and a negative det J weighs its point negatively, as the reference's
    integration does.  A 3x3x2 box whose middle node plane is moved onto the
    bottom (the lower layer has zero volume: every point skipped, K_e = 0),
    and one upper element listed with its faces swapped (det J < 0
    everywhere): FP64 and FP32 values and both SpMVs against the oracle;
    with both node planes at z = 0 clamped, the upper layer carries the
    load and the FP64 solve converges to the oracle's displacements."""
    import ufc_hip
    from ufc_hip import problems as P
    coords, conn = ufc_hip.mesh_box(3, 3, 2, 1.0, 1.0, 1.0)
    coords = coords.copy()
    coords[np.abs(coords[:, 2] - 0.5) < 1e-12, 2] = 0.0
    pr = P.cube_pressure(1, name="degenerate")
    pr.coords, pr.conn = coords, conn
    pr.young = np.full(conn.shape[0], 1e9)
    z0 = np.nonzero(coords[:, 2] == 0.0)[0]   # both node planes at z = 0
    assert z0.size == 32
    pr.bc_eq = (3 * z0[:, None] + np.arange(3)).ravel().astype(np.int32)
    pr.bc_val = np.zeros(pr.bc_eq.size)
    top = np.arange(9, 18, dtype=np.int32)
    pr.load_elem, pr.load_face = top, np.full(top.size, 1, np.int32)
    rp, ci, val, F, _ = _oracle_system(orc, pr, allow_degenerate=True)
    lower = orc.ke(coords[conn[0]], 1e9, 0.3)
    assert np.abs(lower).max() == 0.0               # every point skipped
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    say("degenerate layer")
    scale = np.abs(val).max()
    assert np.abs(ctx.get_values(0) - val).max() <= 1e-12 * scale
    x = np.random.default_rng(9).normal(size=pr.n_dof)
    yo = orc.spmv(rp, ci, val, x)
    assert np.abs(ctx.spmv(x, precision=0) - yo).max() <= 1e-12 * np.abs(yo).max()
    ctx.assemble_k(pr.young, pr.nu, precision=1)
    assert np.abs(ctx.spmv(x, precision=1) - yo).max() <= 1e-5 * np.abs(yo).max()
    uo, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-13, max_iter=100000)
    assert st == 0
    ctx.assemble_k(pr.young, pr.nu, precision=0)   # the FP32 assembly replaced the FP64 values
    u, info = ctx.pcg(precision=0, rtol=1e-13, max_iter=100000)
    assert info.converged and np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-9
    # one upper element with its bottom and top faces swapped: det J < 0
    inv = conn.copy()
    inv[13] = conn[13][[4, 5, 6, 7, 0, 1, 2, 3]]
    pr.conn = inv
    rp, ci, val, F, _ = _oracle_system(orc, pr, allow_degenerate=True)
    assert np.trace(orc.ke(coords[inv[13]], 1e9, 0.3)) < 0.0
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    assert np.abs(ctx.get_values(0) - val).max() <= 1e-12 * np.abs(val).max()
    ctx.assemble_k(pr.young, pr.nu, precision=1)
    yo = orc.spmv(rp, ci, val, x)
    assert np.abs(ctx.spmv(x, precision=1) - yo).max() <= 1e-5 * np.abs(yo).max()
