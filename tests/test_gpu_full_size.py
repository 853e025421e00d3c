"""Parity at the benchmark sizes, through the C-ABI.

128^3 lognormal cube (2,097,152 C3D8, 6,440,067 DOF, nnz 513,599,625; the
plain-FP64 and mixed-precision workloads), against the FP64 oracle on the
same seeded mesh and field:
  - row_ptr / col_idx bit-exact with orc_csr_pattern (RT_CSR_FromTripletMerged
    order, RT_Solv_Sparse.f90:336-447);
  - FP64 values within 1e-12 max|K| of orc_assemble;
  - the FP32 SpMV of the symmetric block-diagonal view (512-node tiles at
    this size, 33 zero tiles in front, int32 address arithmetic over 1e9
    entries) within
    1e-5 max|y| of the oracle's FP64 K x;
  - the FP64 and the mixed-precision GPU solutions: the oracle's true FP64
    residual ||F - K u|| / ||F|| below 2e-10 for both (NM_Solver_PCG stops
    at 1e-10 on its recurrence residual, NM_Solv_Core.f90:389-399), and the
    two solutions within 1e-8 of each other.

16^3 jittered J2 cube (the 64^3 nonlinear workload's mesh generator, seed and
material at a size the oracle's Newton loop runs in seconds): the GPU
Newton loop with exact linear solves against orc_j2_newton_step over all 10
increments, displacements within 1e-7, stresses and plastic strains within
1e-6.

96^3 SIMP adjoint workload (config 4 at full size): the oracle's residuals of
the GPU's forward and adjoint solutions, lambda = u for compliance, and the
oracle's element sensitivities from the GPU's u and lambda against the
torch.autograd gradient.

64^3 J2 job (config 5 at full size): the oracle replays the J2 state from the
GPU's converged displacement of every increment and checks the residual of
each increment and the final state of every Gauss point.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def cube128(orc):
    """The 128^3 system on both sides: oracle CSR, values and F; a GPU
    context with the same pattern, loads and BCs."""
    import ufc_hip
    from ufc_hip import problems as P
    pr = P.lognormal_cube(128)
    say("cube128 oracle pattern")
    rp, ci = orc.csr_pattern(pr.coords.shape[0], pr.conn)
    say("cube128 oracle assembly, threads", _threads())
    val = orc.assemble(pr.conn, pr.coords, pr.young, pr.nu, rp, ci, nthreads=_threads())
    F = np.zeros(pr.n_dof)
    orc.face_loads_pressure(pr.load_elem, pr.load_face, pr.conn, pr.coords,
                            pr.load_value[0], F)
    orc.apply_penalty(pr.bc_eq, pr.bc_val, P.PENALTY, rp, ci, val, F)
    say("cube128 gpu pattern")
    ctx = ufc_hip.Context(0)
    P.setup_context(ctx, pr, ufc_hip.PREC_MIXED)
    yield dict(pr=pr, rp=rp, ci=ci, val=val, F=F, ctx=ctx, sol={})
    ctx.close()


def test_cube128_pattern_bitexact(cube128):
    ctx = cube128["ctx"]
    assert ctx.nnz == 513599625 and ctx.n_dof == 6440067
    rp, ci = ctx.get_pattern()
    say("cube128 pattern compare")
    assert np.array_equal(rp, cube128["rp"])
    assert np.array_equal(ci, cube128["ci"])


def test_cube128_fp64_values(cube128):
    import ufc_hip
    ctx, pr = cube128["ctx"], cube128["pr"]
    ctx.assemble_k(pr.young, pr.nu, precision=ufc_hip.PREC_FP64)
    v = ctx.get_values(ufc_hip.PREC_FP64)
    val = cube128["val"]
    err = np.abs(v - val).max() / np.abs(val).max()
    say("cube128 FP64 values max rel err", err)
    assert err <= 1e-12
    assert np.array_equal(ctx.get_rhs(), cube128["F"]) or \
        np.abs(ctx.get_rhs() - cube128["F"]).max() <= 1e-12 * np.abs(cube128["F"]).max()


def test_cube128_fp32_symdiag_spmv(cube128, orc):
    import ufc_hip
    ctx, pr = cube128["ctx"], cube128["pr"]
    assert ctx.fp32_view == ufc_hip._abi.FP32_VIEW_SYMDIAG
    ctx.assemble_k(pr.young, pr.nu, precision=ufc_hip.PREC_FP32)
    x = np.random.default_rng(11).normal(size=pr.n_dof)
    y32 = ctx.spmv(x, precision=ufc_hip.PREC_FP32)
    yo = orc.spmv(cube128["rp"], cube128["ci"], cube128["val"], x, nthreads=_threads())
    err = np.abs(y32 - yo).max() / np.abs(yo).max()
    say("cube128 FP32 symdiag SpMV max rel err", err)
    assert err <= 1e-5


@pytest.mark.parametrize("precision", [0, 2])
def test_cube128_solution_true_residual(cube128, orc, precision):
    """PCG to 1e-10 (FP64, and FP32 inner in FP64 refinement with residual
    replacement): the residual of the returned u under the oracle's FP64
    operator is below 2e-10."""
    ctx, pr = cube128["ctx"], cube128["pr"]
    ctx.assemble_k(pr.young, pr.nu, precision=precision)
    say("cube128 solve precision", precision)
    u, info = ctx.pcg(precision=precision, rtol=1e-10, inner_rtol=1e-5,
                      replace_rtol=0.1 if precision == 2 else 0.0)
    assert info.status == 0 and info.converged and info.rnorm < 1e-10
    F = cube128["F"]
    r = F - orc.spmv(cube128["rp"], cube128["ci"], cube128["val"], u, nthreads=_threads())
    rel = np.linalg.norm(r) / np.linalg.norm(F)
    say("cube128 precision", precision, "iterations", info.niter, "outer", info.n_outer,
        "oracle true residual", rel)
    assert rel < 2e-10
    cube128["sol"][precision] = u
    if len(cube128["sol"]) == 2:
        a, b = cube128["sol"][0], cube128["sol"][2]
        d = np.linalg.norm(a - b) / np.linalg.norm(a)
        say("cube128 mixed vs FP64 solution", d)
        assert d < 1e-8


def test_j2_cube16_newton_vs_oracle(ctx, orc):
    """Config-5 mesh generator and material at 16^3: 10 increments of u_z on
    z = 1 to 0.5% strain, exact Newton (linear solves to 1e-12) on the GPU
    against orc_j2_newton_step increment by increment."""
    from ufc_hip import problems as P
    from ufc_hip.nonlinear import j2_static
    pr = P.j2_cube(16, increments=10)
    props = pr.extra["props"]
    target = np.zeros(pr.bc_eq.size)
    target[pr.extra["n_clamp"]:] = pr.extra["strain"]
    ctx.build_pattern(pr.coords, pr.conn)
    say("j2 16^3 gpu")
    res = j2_static(ctx, pr.bc_eq, target, props, n_increments=10, precision=0,
                    lin_rtol=1e-12)
    assert res.converged and len(res.increments) == 10
    sg = ctx.j2_get_state().reshape(-1, 13)
    n_gp = pr.n_elems * 8
    state = np.zeros(n_gp * 13)
    uo = np.zeros(pr.n_dof)
    newton_o = []
    for k in range(1, 11):
        st, uo, state, nn, rn = orc.j2_newton_step(pr.coords.shape[0], pr.conn, pr.coords,
                                                   props, pr.bc_eq, target * (k / 10.0), 1e12,
                                                   state, uo)
        assert st == 0
        newton_o.append(nn)
    say("j2 16^3 newton gpu", [i.newton_iters for i in res.increments], "oracle", newton_o)
    so = state.reshape(-1, 13)
    err = np.linalg.norm(res.u - uo) / np.linalg.norm(uo)
    say("j2 16^3 u err", err, "yielded", res.increments[-1].n_yield)
    assert res.increments[-1].n_yield > 0
    assert err < 1e-7
    assert np.abs(sg[:, :6] - so[:, :6]).max() < 1e-6 * np.abs(so[:, :6]).max()
    assert np.abs(sg[:, 6] - so[:, 6]).max() < 1e-6 * so[:, 6].max()
    assert np.abs(sg[:, 7:] - so[:, 7:]).max() < 1e-6 * np.abs(so[:, 7:]).max()


def test_adjoint96_vs_oracle(orc):
    """The 96^3 SIMP adjoint workload at full size (2,738,019 DOF): the mixed
    forward solve and the mixed adjoint solve K^T lambda = f behind
    torch.autograd, checked by the oracle on the same mesh and field:
      - the oracle's FP64 residuals ||F - K u|| / ||F|| and ||f - K lambda|| /
        ||f|| below 2e-10 (both solves stop at 1e-10 on their recurrence
        residuals; NM_Adjoint_Solve solves K^T lambda = dJ/du with K^T = K,
        NM_Solv_GMRESTranspose.f90:229-291);
      - lambda = u for compliance (J = f.u, f the external load) to 1e-8;
      - the element sensitivities the oracle forms from the GPU's u and
        lambda within 1e-10 of the autograd gradient."""
    import torch
    import ufc_hip
    from ufc_hip import problems as P
    from ufc_hip.autograd import StaticModel, compliance
    pa = P.simp_cube(96)
    model = StaticModel.from_problem(pa, precision=ufc_hip.PREC_MIXED, rtol=1e-10)
    try:
        rho = torch.as_tensor(pa.extra["rho"], device="cuda:0").requires_grad_(True)
        young = 1e-3 + rho ** 3 * 1e9
        say("adjoint 96^3: forward + backward")
        c, u = compliance(model, young)
        c.backward()
        g_gpu = rho.grad.detach().cpu().numpy()
        lam = torch.zeros_like(u)
        _, info = model.ctx.adjoint_solve(model.f, precision=ufc_hip.PREC_MIXED, rtol=1e-10,
                                          x=lam)
        assert info.converged
        u_h, lam_h = u.detach().cpu().numpy(), lam.cpu().numpy()
        f_ext = model.f.cpu().numpy()
        e_h = young.detach().cpu().numpy()
    finally:
        model.close()
    say("adjoint 96^3: oracle pattern + assembly")
    rp, ci = orc.csr_pattern(pa.coords.shape[0], pa.conn)
    val = orc.assemble(pa.conn, pa.coords, e_h, pa.nu, rp, ci, nthreads=_threads())
    F = np.zeros(pa.n_dof)
    orc.face_loads_pressure(pa.load_elem, pa.load_face, pa.conn, pa.coords,
                            pa.load_value[0], F)
    orc.apply_penalty(pa.bc_eq, pa.bc_val, P.PENALTY, rp, ci, val, F)
    ru = np.linalg.norm(F - orc.spmv(rp, ci, val, u_h, _threads())) / np.linalg.norm(F)
    rl = np.linalg.norm(f_ext - orc.spmv(rp, ci, val, lam_h, _threads())) / np.linalg.norm(f_ext)
    d = np.linalg.norm(lam_h - u_h) / np.linalg.norm(u_h)
    g_o = orc.sensitivity(pa.conn, pa.coords, pa.nu, 3.0 * pa.extra["rho"] ** 2 * 1e9, u_h,
                          lam_h, nthreads=_threads())
    eg = np.abs(g_o - g_gpu).max() / np.abs(g_o).max()
    say("adjoint 96^3: residual u", ru, "lambda", rl, "lambda vs u", d, "gradient", eg)
    assert ru < 2e-10 and rl < 2e-10
    assert d < 1e-8
    assert eg < 1e-10


def test_j2_cube64_oracle_replay(ctx, orc):
    """The 64^3 J2 job itself (config 5: jittered mesh, 10 increments to 0.5%
    strain, inexact Newton with the mixed inner solve), replayed by the
    oracle: from the GPU's converged displacement of each increment the
    oracle recomputes the J2 state from the previous increment's
    (orc_j2_internal_force: strain increment, radial return,
    PH_Mat_Plast_J2_Iso_Core.f90:254-456) and the residual R = P ubar -
    F_int - P u (RT_Asm_ComputeResidual, RT_Asm_Solv.f90:1311-1426):
      - ||R|| / ||F_int|| below 2e-8 at every increment (the GPU's Newton
        stops at 1e-8 by its own FP64 residual);
      - no local failure, and the GPU's committed state after the last
        increment (stress, equivalent and plastic strains of all 2,097,152
        Gauss points) within 1e-9 of the oracle's."""
    from ufc_hip import problems as P
    from ufc_hip.nonlinear import j2_static, to_host
    pr = P.j2_cube(64)
    props = pr.extra["props"]
    target = np.zeros(pr.bc_eq.size)
    target[pr.extra["n_clamp"]:] = pr.extra["strain"]
    ctx.build_pattern(pr.coords, pr.conn)
    say("j2 64^3 gpu job (mixed, inexact Newton)")
    res = j2_static(ctx, pr.bc_eq, target, props, n_increments=10, precision=2,
                    keep_increments=True)
    assert res.converged and len(res.u_increments) == 10 and res.cutbacks == 0
    sg = ctx.j2_get_state().reshape(-1, 13)
    nn = pr.coords.shape[0]
    state = np.zeros(pr.n_elems * 8 * 13)
    u_prev = np.zeros(pr.n_dof)
    rels = []
    for k, uk in enumerate(res.u_increments, start=1):
        uk = to_host(uk)
        nf, state, _, fi = orc.j2_internal_force(nn, pr.conn, pr.coords, props, uk, u_prev,
                                                 state, nthreads=_threads())
        assert nf == 0
        r = -fi
        r[pr.bc_eq] += P.PENALTY * (target * (k / 10.0) - uk[pr.bc_eq])
        rels.append(np.linalg.norm(r) / np.linalg.norm(fi))
        u_prev = uk
    so = state.reshape(-1, 13)
    es = np.abs(sg[:, :6] - so[:, :6]).max() / np.abs(so[:, :6]).max()
    ep = np.abs(sg[:, 6:] - so[:, 6:]).max() / np.abs(so[:, 6:]).max()
    say("j2 64^3 oracle residuals", ["%.2e" % x for x in rels], "state", es, ep,
        "yielded", int(np.count_nonzero(so[:, 6])))
    assert max(rels) < 2e-8
    assert es < 1e-9 and ep < 1e-9


def test_nnz_beyond_int32():
    """A 208 x 206 x 206 box (8.96M nodes, 26.9M DOF, nnz 2,155,283,649 >
    2^31): value positions are 64-bit, so assembly, both SpMVs and both
    solves run; the CSR export, whose row_ptr is the reference's int32
    rowPtr, refuses (UFC_STATUS_UNSUPPORTED).  Size-independent checks: the
    FP64 K annihilates rigid translations and an infinitesimal rotation, the
    FP32 symmetric view applies K to FP32 rounding, the mixed and FP64
    solutions agree to 1e-8, and the element-by-element FP64 residual of the
    mixed solution (no CSR involved) is below 2e-10."""
    import ufc_hip
    from ufc_hip import _abi as A
    from ufc_hip import problems as P
    nx, ny, nz = 208, 206, 206
    coords, conn = ufc_hip.mesh_box(nx, ny, nz, 1.0, 1.0, 1.0)
    young = ufc_hip.field("lognormal", conn.shape[0], P.SEED, np.log(1e9), 0.5)
    z0 = np.nonzero(coords[:, 2] == 0.0)[0]
    clamp = (3 * z0[:, None] + np.arange(3)).ravel().astype(np.int32)
    top = np.arange(nx * ny * nz, dtype=np.int32).reshape(nz, ny, nx)[nz - 1].ravel()
    with ufc_hip.Context(0) as ctx:
        say("nnz > 2^31: pattern")
        info = ctx.build_pattern(coords, conn)
        assert info.nnz == 9 * (3 * nx + 1) * (3 * ny + 1) * (3 * nz + 1) > 2 ** 31 - 1
        assert ctx.fp32_view == A.FP32_VIEW_SYMDIAG
        with pytest.raises(A.UfcError) as e:
            ctx.get_pattern()
        assert e.value.code == A.STATUS_UNSUPPORTED
        ctx.assemble_k(young, 0.3, precision=0)
        n = ctx.n_dof
        rnd = np.random.default_rng(5).normal(size=n)
        y_r = ctx.spmv(rnd, precision=0)
        for d in range(3):
            t = np.zeros(n)
            t[d::3] = 1.0
            y_t = ctx.spmv(t, precision=0)
            say("translation", d, np.linalg.norm(y_t) / np.linalg.norm(y_r))
            assert np.linalg.norm(y_t) <= 1e-12 * np.linalg.norm(y_r) * np.sqrt(n)
        rot = np.zeros((coords.shape[0], 3))
        rot[:, 0], rot[:, 1] = -coords[:, 1], coords[:, 0]  # rotation about z
        y_rot = ctx.spmv(rot.ravel(), precision=0)
        assert np.linalg.norm(y_rot) <= 1e-10 * np.linalg.norm(y_r)
        ctx.load_pressure(top, np.full(top.size, 1, np.int32), 1.0e6)
        ctx.apply_bc(clamp, None, P.PENALTY)
        ctx.assemble_k(young, 0.3, precision=0)
        y64 = ctx.spmv(rnd, precision=0)
        say("nnz > 2^31: FP64 solve")
        u64, i64 = ctx.pcg(precision=0, rtol=1e-10)
        assert i64.converged
        ctx.assemble_k(young, 0.3, precision=2)
        y32 = ctx.spmv(rnd, precision=1)
        assert np.abs(y32 - y64).max() <= 1e-5 * np.abs(y64).max()
        say("nnz > 2^31: mixed solve")
        um, im = ctx.pcg(precision=2, rtol=1e-10, replace_rtol=0.1)
        assert im.converged
        d = np.linalg.norm(um - u64) / np.linalg.norm(u64)
        r, norms = ctx.residual(um)
        F = ctx.get_rhs()
        rel = norms[0] / np.linalg.norm(F)
        say("nnz > 2^31: iterations", i64.niter, im.niter, "mixed vs FP64", d, "residual", rel)
        assert d < 1e-8 and rel < 2e-10
