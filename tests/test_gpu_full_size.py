"""Parity at the benchmark sizes, through the C-ABI.

128^3 lognormal cube (2,097,152 C3D8, 6,440,067 DOF, nnz 513,599,625; the
plain-FP64 and mixed-precision workloads), against the FP64 oracle on the
same seeded mesh and field:
  - row_ptr / col_idx bit-exact with orc_csr_pattern (RT_CSR_FromTripletMerged
    order, RT_Solv_Sparse.f90:336-447);
  - FP64 values within 1e-12 max|K| of orc_assemble;
  - the FP32 SpMV of the symmetric block-diagonal view (66 zero tiles in
    front at this size, int32 address arithmetic over 1e9 entries) within
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
