"""C3D8 variants on the GPU (UFC_VARIANT_*) against the oracle's
orc_c3d8_ke_variant: batched Ke (FP64 1e-12, FP32 2e-5 of max|Ke|), the
fused CSR assembly (FP64 1e-12 of max|K|), FP64 and FP32-in-FP64 solves
(displacements 1e-8, both driven to a tight residual; the mixed form's FP64
residual is the matrix-free element apply of the same variant), the
residual at the solution, and the variant sensitivities lambda_e^T K_e u_e."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

VARIANTS = [1, 2, 3, 4, 6]
NAMES = {1: "standard", 2: "reduced", 3: "selective", 4: "incompat", 6: "gauss27"}


def say(*a):
    print("[gpu-test]", *a, flush=True)


def _xe(n, seed=0):
    unit = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    rng = np.random.default_rng(seed)
    return unit[None] * rng.uniform(0.5, 2.0, (n, 1, 3)) + 0.15 * rng.uniform(-1, 1, (n, 8, 3))


@pytest.mark.parametrize("variant", VARIANTS, ids=[NAMES[v] for v in VARIANTS])
def test_ke_batch_vs_oracle(ctx, orc, variant):
    n = 96
    xe = _xe(n)
    E = np.random.default_rng(3).uniform(1e9, 2e11, n)
    say("eval_ke", NAMES[variant], "n", n)
    K = ctx.eval_ke(xe, E, 0.3, variant=variant)
    Ko = np.stack([orc.ke_variant(xe[i], E[i], 0.3, variant) for i in range(n)])
    scale = np.abs(Ko).max(axis=(1, 2))
    assert (np.abs(K - Ko).max(axis=(1, 2)) / scale).max() < 1e-12
    K32 = ctx.eval_ke(xe, E, 0.3, precision=1, variant=variant)
    assert (np.abs(K32 - Ko).max(axis=(1, 2)) / scale).max() < 2e-5


def _problem(kind="jitter5"):
    from ufc_hip import problems as P
    if kind == "irregular":  # notched, renumbered: node degrees 1..8, ragged rows
        from test_gpu_parity import _irregular
        return _irregular()
    pr = P.cube_pressure(5, name=kind, jitter=0.1)
    pr.young = np.random.default_rng(7).uniform(0.5e9, 2e9, pr.n_elems)
    return pr


def _oracle_system(orc, pr, variant, penalty=1e12):
    rp, ci = orc.csr_pattern(pr.coords.shape[0], pr.conn)
    val = orc.assemble_variant(pr.conn, pr.coords, pr.young, pr.nu, variant, rp, ci)
    F = np.zeros(pr.n_dof)
    orc.face_loads_pressure(pr.load_elem, pr.load_face, pr.conn, pr.coords, pr.load_value[0], F)
    k_only = val.copy()
    orc.apply_penalty(pr.bc_eq, pr.bc_val, penalty, rp, ci, val, F)
    return rp, ci, k_only, val, F


@pytest.mark.parametrize("mesh", ["jitter5", "irregular"])
@pytest.mark.parametrize("variant", VARIANTS, ids=[NAMES[v] for v in VARIANTS])
def test_assembly_vs_oracle(ctx, orc, variant, mesh):
    from ufc_hip import problems as P
    pr = _problem(mesh)
    rp, ci, k_only, val, F = _oracle_system(orc, pr, variant)
    say("assembly", NAMES[variant], mesh, "dof", pr.n_dof)
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0, variant=variant)
    v64 = ctx.get_values(0)
    scale = np.abs(val).max()
    assert np.abs(v64 - val).max() <= 1e-12 * scale
    # block rows of K sum to zero (translations carry no force), which the
    # FP32 difference form relies on
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    rowsum = np.zeros((rp.size - 1, 3))
    np.add.at(rowsum, (rows, ci % 3), k_only)
    assert np.abs(rowsum).max() <= 1e-12 * scale


@pytest.mark.parametrize("mesh", ["jitter5", "irregular"])
@pytest.mark.parametrize("prec", [0, 2], ids=["fp64", "mixed"])
@pytest.mark.parametrize("variant", VARIANTS, ids=[NAMES[v] for v in VARIANTS])
def test_solve_vs_oracle(ctx, orc, variant, prec, mesh):
    from ufc_hip import problems as P
    pr = _problem(mesh)
    rp, ci, _, val, F = _oracle_system(orc, pr, variant)
    uo, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-14, max_iter=200000)
    assert st == 0
    say("solve", NAMES[variant], mesh, "precision", prec)
    P.setup_context(ctx, pr, prec)
    ctx.assemble_k(pr.young, pr.nu, precision=prec, variant=variant)
    u, info = ctx.pcg(precision=prec, rtol=1e-12, max_iter=200000, replace_rtol=0.1)
    assert info.status == 0 and info.converged
    assert np.linalg.norm(u - uo) / np.linalg.norm(uo) < 1e-8
    # residual through the matrix-free element apply of the same variant
    _, norms = ctx.residual(u)
    assert norms[0] <= 1e-9 * np.linalg.norm(F)


@pytest.mark.parametrize("variant", [2, 3, 4, 6], ids=[NAMES[v] for v in [2, 3, 4, 6]])
def test_sensitivity_vs_oracle(ctx, orc, variant):
    from ufc_hip import problems as P
    pr = _problem()
    P.setup_context(ctx, pr, 0)
    ctx.assemble_k(pr.young, pr.nu, precision=0, variant=variant)
    rng = np.random.default_rng(11)
    u, lam = rng.standard_normal(pr.n_dof), rng.standard_normal(pr.n_dof)
    dedrho = rng.uniform(0.5, 2.0, pr.n_elems) * 1e9
    say("sensitivity", NAMES[variant])
    g = ctx.sensitivity(u, lam, dedrho, pr.nu)
    go = np.empty(pr.n_elems)
    for e in range(pr.n_elems):
        d = (3 * pr.conn[e][:, None] + np.arange(3)).ravel()
        K0 = orc.ke_variant(pr.coords[pr.conn[e]], 1.0, pr.nu, variant)
        go[e] = -dedrho[e] * lam[d] @ K0 @ u[d]
    assert np.abs(g - go).max() <= 1e-10 * np.abs(go).max()


def test_variant_arguments(ctx):
    from ufc_hip import problems as P, UfcError
    from ufc_hip._abi import MAT_J2, STATUS_INVALID, STATUS_UNSUPPORTED
    pr = P.cube_pressure(2)
    P.setup_context(ctx, pr, 0)
    with pytest.raises(UfcError) as e:
        ctx.assemble_k(pr.young, pr.nu, precision=0, variant=9)
    assert e.value.code == STATUS_INVALID
    ctx.j2_init(200e9, 0.3, 250e6, 2e9)
    with pytest.raises(UfcError) as e:
        ctx.assemble_k(None, 0.3, precision=0, material=MAT_J2, variant=2)
    assert e.value.code == STATUS_UNSUPPORTED
    # 0 and 5 are the standard element, as the reference routes them
    ctx.assemble_k(pr.young, pr.nu, precision=0, variant=1)
    v1 = ctx.get_values(0)
    for v in (0, 5):
        ctx.assemble_k(pr.young, pr.nu, precision=0, variant=v)
        assert np.array_equal(ctx.get_values(0), v1)
