"""Meshes beyond the 8-element / 32-neighbour node map, through the C-ABI
against the oracle (the reference assembles any hex mesh,
RT_Asm_Solv.f90:1840-1900, RT_Solv_Sparse.f90:336-447).

Two meshes:
  - polar: a disk of kite and ring quadrilaterals around a centre node,
    extruded in 4 layers (problems.polar_prism): the interior axis nodes lie
    in 10 elements with 33 neighbours, so their blocks run the general
    assembly and gather kernels and every SpMV the long-list path;
  - duplicated: the 3^3 box with its corner element listed 9 times: nodes in
    up to 16 elements but at most 27 neighbours (general assembly blocks, the
    32-lane SpMV views).
Tolerances as the structured-mesh tests: pattern bit-exact, FP64 values
1e-12 max|K|, FP32 values 1e-6 max|K| (difference form), SpMV 1e-12 / 1e-5,
displacements 1e-8 (both sides driven to a tight residual), J2 u 1e-7 and
plastic strains 1e-6, sensitivities 1e-9.
"""
import numpy as np
import pytest

from test_gpu_parity import _difference_form, _oracle_system

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _mesh(kind):
    from ufc_hip import problems as P
    if kind == "polar":
        return P.polar_prism(5, 3, 4)
    pr = P.cube_pressure(3, name="duplicated")
    pr.conn = np.concatenate([pr.conn] + [pr.conn[:1]] * 8)
    pr.young = np.concatenate([pr.young, np.full(8, 2e9)])
    return pr


def _node_stats(pr):
    cnt = np.bincount(pr.conn.ravel(), minlength=pr.coords.shape[0])
    return cnt


@pytest.mark.parametrize("kind", ["polar", "duplicated"])
def test_general_mesh_pattern_values_spmv(ctx, orc, kind):
    from ufc_hip import _abi as A
    from ufc_hip import problems as P
    pr = _mesh(kind)
    cnt = _node_stats(pr)
    assert cnt.max() > 8
    rp_o, ci_o, val_o, F_o, _ = _oracle_system(orc, pr)
    deg = np.diff(rp_o)[::3] // 3
    assert (deg.max() > 32) == (kind == "polar")
    say("general mesh", kind, "max elements per node", cnt.max(), "max neighbours", deg.max())
    P.setup_context(ctx, pr, 0)
    rp, ci = ctx.get_pattern()
    assert np.array_equal(rp, rp_o) and np.array_equal(ci, ci_o)
    if kind == "polar":
        assert ctx.fp32_view == A.FP32_VIEW_ID32
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    v64 = ctx.get_values(0)
    scale = np.abs(val_o).max()
    assert np.abs(v64 - val_o).max() <= 1e-12 * scale
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    assert np.array_equal(ctx.get_values(0), v64)  # deterministic
    x = np.random.default_rng(1).normal(size=pr.n_dof)
    yo = orc.spmv(rp_o, ci_o, val_o, x)
    assert np.abs(ctx.spmv(x, precision=0) - yo).max() <= 1e-12 * np.abs(yo).max()
    ctx.assemble_k(pr.young, pr.nu, precision=1)
    v32 = ctx.get_values(1).astype(np.float64)
    assert np.abs(v32 - _difference_form(orc, pr, rp_o, ci_o)).max() <= 1e-6 * scale
    assert np.abs(ctx.spmv(x, precision=1) - yo).max() <= 1e-5 * np.abs(yo).max()
    # every FP32 view the mesh allows gives the same product
    for view in (A.FP32_VIEW_OFF16, A.FP32_VIEW_PATTERN):
        try:
            got = ctx.set_fp32_view(view)
        except A.UfcError as e:
            assert e.code == A.STATUS_UNSUPPORTED and kind == "polar"
            continue
        assert got == view
        assert np.abs(ctx.spmv(x, precision=1) - yo).max() <= 1e-5 * np.abs(yo).max()


@pytest.mark.parametrize("kind", ["polar", "duplicated"])
@pytest.mark.parametrize("precision", [0, 2])
def test_general_mesh_solves_vs_oracle(ctx, orc, kind, precision):
    """FP64 PCG and the mixed refinement (FP64 element-by-element residual
    through the general gather) to the oracle's displacements."""
    from ufc_hip import problems as P
    pr = _mesh(kind)
    rp, ci, val, F, _ = _oracle_system(orc, pr)
    uo, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-14, max_iter=200000)
    assert st == 0
    P.setup_context(ctx, pr, precision)
    ctx.assemble_k(pr.young, pr.nu, precision=precision)
    say("general mesh solve", kind, "precision", precision)
    u, info = ctx.pcg(precision=precision, rtol=1e-13, max_iter=200000, replace_rtol=0.1)
    assert info.status == 0 and info.converged
    err = np.linalg.norm(u - uo) / np.linalg.norm(uo)
    say("err vs oracle", err)
    assert err < 1e-8


@pytest.mark.parametrize("precision", [0, 2])
def test_valence10_deck_solves_vs_oracle(ctx, orc, tmp_path, precision):
    """The polar mesh as an input deck (write_inp -> ufc_ap_read_inp ->
    setup): a deck with valence-10 nodes solves to the oracle's
    displacements."""
    from ufc_hip import inp
    from test_gpu_inp import _oracle_solution
    pr = _mesh("polar")
    path = str(tmp_path / "polar.inp")
    inp.write_inp(path, pr)
    d = inp.read_inp(path)
    assert np.bincount(d.conn.ravel()).max() == 10
    uo = _oracle_solution(orc, d)
    say("valence-10 deck precision", precision)
    inp.setup_context(ctx, d)
    ctx.assemble_k(d.young, d.nu, precision=precision)
    u, info = ctx.pcg(precision=precision, rtol=1e-12, max_iter=200000, replace_rtol=0.1)
    assert info.status == 0 and info.converged
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8


@pytest.mark.parametrize("variant", [2, 3, 4, 6])
def test_general_mesh_element_variants(ctx, orc, variant):
    """Reduced integration (2), selective (3), incompatible modes (4) and the
    27-point rule (6) on the polar mesh: FP64 values against the oracle's
    variant assembly (1e-12), FP32 values against its difference form
    (1e-6), and the mixed solve against the oracle's displacements (1e-8)."""
    from ufc_hip import problems as P
    pr = _mesh("polar")
    rp, ci = orc.csr_pattern(pr.coords.shape[0], pr.conn)
    vo = orc.assemble_variant(pr.conn, pr.coords, pr.young, pr.nu, variant, rp, ci)
    vd = vo.copy()
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    vd[(ci // 3) == (rows // 3)] = 0.0
    orc.apply_penalty(pr.bc_eq, pr.bc_val, 1e12, rp, ci, vd, None)
    F = np.zeros(pr.n_dof)
    orc.face_loads_pressure(pr.load_elem, pr.load_face, pr.conn, pr.coords, pr.load_value[0], F)
    orc.apply_penalty(pr.bc_eq, pr.bc_val, 1e12, rp, ci, vo, F)
    P.setup_context(ctx, pr, 0)
    say("general mesh variant", variant)
    ctx.assemble_k(pr.young, pr.nu, precision=0, variant=variant)
    v = ctx.get_values(0)
    scale = np.abs(vo).max()
    assert np.abs(v - vo).max() <= 1e-12 * scale
    ctx.assemble_k(pr.young, pr.nu, precision=1, variant=variant)
    assert np.abs(ctx.get_values(1).astype(np.float64) - vd).max() <= 1e-6 * scale
    uo, st, _, _ = orc.pcg(rp, ci, vo, F, rtol=1e-14, max_iter=200000)
    ctx.assemble_k(pr.young, pr.nu, precision=2, variant=variant)
    u, info = ctx.pcg(precision=2, rtol=1e-12, max_iter=200000, replace_rtol=0.1)
    assert st == 0 and info.converged
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8


def test_general_mesh_transpose_and_sensitivity(ctx, orc):
    """K^T x (long neighbour lists in the transposed read) and the compliance
    sensitivity through the adjoint on the polar mesh."""
    from ufc_hip import problems as P
    pr = _mesh("polar")
    rho = np.random.default_rng(4).uniform(0.3, 1.0, pr.n_elems)
    pr.young = 1e-3 + rho ** 3 * 1e9
    rp, ci, val, F, f0 = _oracle_system(orc, pr)
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    x = np.random.default_rng(2).normal(size=pr.n_dof)
    yo = orc.spmv_t(rp, ci, val, x)
    assert np.abs(ctx.spmv_t(x) - yo).max() <= 1e-12 * np.abs(yo).max()
    say("general mesh adjoint")
    uo, _, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-14, max_iter=200000)
    u, _ = ctx.pcg(precision=0, rtol=1e-13, max_iter=200000)
    lam, _ = ctx.adjoint_solve(f0, precision=0, rtol=1e-13)
    lo, st, _, _ = orc.cg_transpose(rp, ci, val, f0, tol=1e-13, max_iter=200000)
    assert st == 0 and np.linalg.norm(lam - lo) / np.linalg.norm(lo) < 1e-9
    dedrho = 3 * rho ** 2 * 1e9
    g = ctx.sensitivity(u, lam, dedrho, pr.nu)
    go = orc.sensitivity(pr.conn, pr.coords, pr.nu, dedrho, uo, lo)
    assert np.abs(g - go).max() < 1e-9 * np.abs(go).max()


@pytest.mark.parametrize("precision", [0, 2])
def test_general_mesh_j2_newton_vs_oracle(ctx, orc, precision):
    """J2 (linear hardening) on the polar mesh, u_z prescribed on top to 0.5%
    strain in 5 increments: the tangent assembly and the internal-force
    gather of the general nodes, against the oracle's Newton loop."""
    from ufc_hip.nonlinear import j2_static
    pr = _mesh("polar")
    bottom = np.nonzero(pr.coords[:, 2] == 0.0)[0]
    top = np.nonzero(pr.coords[:, 2] == 1.0)[0]
    eq = np.concatenate([(3 * bottom[:, None] + np.arange(3)).ravel(), 3 * top + 2]).astype(np.int32)
    target = np.concatenate([np.zeros(3 * bottom.size), np.full(top.size, 0.005)])
    props = (200e9, 0.3, 250e6, 2e9)
    ctx.build_pattern(pr.coords, pr.conn)
    say("general mesh j2 precision", precision)
    res = j2_static(ctx, eq, target, props, n_increments=5, precision=precision, lin_rtol=1e-12)
    assert res.converged and len(res.increments) == 5
    assert res.increments[-1].n_yield > 0
    sg = ctx.j2_get_state().reshape(-1, 13)
    state = np.zeros(pr.n_elems * 8 * 13)
    uo = np.zeros(pr.n_dof)
    for k in range(1, 6):
        st, uo, state, nn, rn = orc.j2_newton_step(pr.coords.shape[0], pr.conn, pr.coords, props,
                                                   eq, target * (k / 5.0), 1e12, state, uo)
        assert st == 0
    so = state.reshape(-1, 13)
    err = np.linalg.norm(res.u - uo) / np.linalg.norm(uo)
    say("j2 err vs oracle", err)
    assert err < 1e-7
    assert np.abs(sg[:, 6] - so[:, 6]).max() < 1e-6 * so[:, 6].max()
