"""Transpose SpMV and the non-symmetric adjoint (GMRES on K^T) on the GPU.

A non-symmetric operator on the stiffness pattern is set with
ufc_rt_asm_set_values (K plus a skew-symmetric part, so the system stays
solvable); K^T x is checked against the oracle's CSR transpose product
(1e-12 of max|y|), and GMRES(m) on K^T against a dense solve of the same
system (lambda 1e-8 relative, both driven to a tight residual), for the
reference's restart 30 and a short restart that forces restarts.  On a
symmetric K the GMRES and PCG adjoints agree to 1e-8."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _system(ctx, orc, n=4, skew=0.2, seed=0):
    from ufc_hip import problems as P
    pr = P.cube_pressure(n, name=f"cube{n}", jitter=0.1)
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    rp, ci = ctx.get_pattern()
    K = ctx.get_values(0)
    rng = np.random.default_rng(seed)
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    # skew-symmetric part on the pattern: S_ij = r_ij - r_ji
    R = rng.standard_normal(ci.size)
    pos = {(int(i), int(j)): q for q, (i, j) in enumerate(zip(rows, ci))}
    mirror = np.array([pos[(int(j), int(i))] for i, j in zip(rows, ci)])
    offdiag = np.abs(K[rows != ci]).max()
    S = skew * offdiag * (R - R[mirror])
    return pr, rp, ci, rows, K + S


def _dense(rp, ci, rows, val):
    n = rp.size - 1
    A = np.zeros((n, n))
    A[rows, ci] = val
    return A


def test_spmv_t_vs_oracle(ctx, orc):
    pr, rp, ci, rows, M = _system(ctx, orc, n=5)
    say("spmv_t dof", pr.n_dof)
    ctx.set_values(M)
    x = np.random.default_rng(1).standard_normal(pr.n_dof)
    y = ctx.spmv_t(x)
    yo = orc.spmv_t(rp, ci, M, x)
    assert np.abs(y - yo).max() <= 1e-12 * np.abs(yo).max()
    # the set values are what the forward product uses, too
    yf = ctx.spmv(x, precision=0)
    assert np.abs(yf - orc.spmv(rp, ci, M, x)).max() <= 1e-12 * np.abs(yf).max()
    assert np.abs(yf - y).max() > 1e-6 * np.abs(y).max()  # really non-symmetric


@pytest.mark.parametrize("restart", [30, 5])
def test_gmres_adjoint_nonsymmetric_vs_dense(ctx, orc, restart):
    from ufc_hip._abi import KRYLOV_GMRES
    pr, rp, ci, rows, M = _system(ctx, orc, n=4)
    ctx.set_values(M)
    g = np.random.default_rng(2).standard_normal(pr.n_dof)
    lo = np.linalg.solve(_dense(rp, ci, rows, M).T, g)
    say("gmres adjoint restart", restart, "dof", pr.n_dof)
    lam, info = ctx.adjoint_solve(g, rtol=1e-12, max_iter=20000, krylov=KRYLOV_GMRES,
                                  restart=restart)
    assert info.status == 0 and info.converged and info.rnorm < 1e-11
    if restart == 5:
        assert info.n_outer > 0
    assert np.linalg.norm(lam - lo) / np.linalg.norm(lo) < 1e-8


def test_gmres_matches_pcg_on_symmetric_k(ctx):
    from ufc_hip import problems as P
    from ufc_hip._abi import KRYLOV_GMRES
    pr = P.cube_pressure(6, name="cube6", jitter=0.1)
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    g = np.random.default_rng(3).standard_normal(pr.n_dof)
    say("gmres vs pcg dof", pr.n_dof)
    l_pcg, i1 = ctx.adjoint_solve(g, rtol=1e-12, max_iter=100000)
    l_gm, i2 = ctx.adjoint_solve(g, rtol=1e-12, max_iter=100000, krylov=KRYLOV_GMRES)
    assert i1.converged and i2.converged
    assert np.linalg.norm(l_gm - l_pcg) / np.linalg.norm(l_pcg) < 1e-8


def test_set_values_round_trip_and_guards(ctx):
    from ufc_hip import problems as P, UfcError
    from ufc_hip._abi import KRYLOV_GMRES, STATUS_INVALID, STATUS_UNSUPPORTED
    pr = P.cube_pressure(4, name="cube4")
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    u0, _ = ctx.pcg(precision=0, rtol=1e-12)
    v = ctx.get_values(0)
    ctx.set_values(v)
    u1, _ = ctx.pcg(precision=0, rtol=1e-12)
    assert np.array_equal(u0, u1)
    with pytest.raises(UfcError) as e:
        ctx.residual(u1)
    assert e.value.code == STATUS_INVALID
    _, info = ctx.pcg(precision=2, rtol=1e-10)  # no FP32 values after set_values
    assert info.status == STATUS_INVALID
    with pytest.raises(UfcError) as e:
        ctx.adjoint_solve(np.ones(pr.n_dof), precision=2, krylov=KRYLOV_GMRES)
    assert e.value.code == STATUS_UNSUPPORTED
    # too few iterations: the reference's non-convergence code
    _, info = ctx.adjoint_solve(np.ones(pr.n_dof), rtol=1e-14, max_iter=3, krylov=KRYLOV_GMRES)
    assert info.status == STATUS_INVALID and not info.converged and info.niter == 3
    ctx.assemble_k(pr.young, pr.nu, precision=0)
    ctx.residual(u1)  # the element operator is back


@pytest.mark.parametrize("mesh", ["1x1x1", "2x1x1", "3x2x2", "20x4x4"])
def test_spmv_and_spmv_t_match_reference_golden(ctx, golden_dir, mesh):
    """The golden matrix of the reference's own assembler (non-symmetric
    integer element matrices) set on the GPU pattern: K x equals the
    reference's NM_CSR_MatVec_Optimized and K^T x its NM_CSR_Transpose
    product, bit for bit (integer arithmetic, exact in FP64)."""
    import os
    import ufc_hip
    g = np.load(os.path.join(golden_dir, f"ref_csr_{mesh}.npz"))
    nx, ny, nz = (int(v) for v in g["mesh"])
    coords, conn = ufc_hip.mesh_box(nx, ny, nz, index_base=1)
    say("golden K^T x", mesh)
    ctx.build_pattern(coords, conn, index_base=1)
    rp, ci = ctx.get_pattern(index_base=1)
    assert np.array_equal(rp, g["ia"]) and np.array_equal(ci, g["ja"])
    ctx.set_values(g["a"])
    x = 1.0 + (5 * np.arange(1, rp.size)) % 13
    assert np.array_equal(ctx.spmv(x, precision=0), g["y"])
    assert np.array_equal(ctx.spmv_t(x), g["yt"])
