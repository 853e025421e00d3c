"""torch.autograd.Function over the C-ABI: compliance gradients of the SIMP
cube against the oracle's adjoint gradient, the reference's adjoint CG
algorithm and central finite differences (h = 1e-6 on rho, both solves
driven to 1e-14)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(n, precision, rtol):
    import torch
    from ufc_hip import problems as P
    from ufc_hip.autograd import StaticModel
    pr = P.simp_cube(n)
    model = StaticModel.from_problem(pr, precision=precision, rtol=rtol)
    rho = torch.tensor(pr.extra["rho"], device="cuda:0", requires_grad=True)
    return pr, model, rho


def _compliance(model, rho):
    import torch
    from ufc_hip.autograd import compliance
    E = 1e-3 + rho ** 3 * 1e9
    c, u = compliance(model, E)
    return c, u


def test_compliance_gradient_vs_oracle_and_fd(orc):
    import torch
    pr, model, rho = _setup(6, 0, 1e-14)
    print("[gpu-test] autograd simp 6^3", flush=True)
    c, u = _compliance(model, rho)
    c.backward()
    g = rho.grad.detach().cpu().numpy()
    # oracle: lambda = u for compliance; dc/drho = -3 rho^2 1e9 u_e^T K_e0 u_e
    rp, ci = orc.csr_pattern(pr.coords.shape[0], pr.conn)
    r = pr.extra["rho"]
    val = orc.assemble(pr.conn, pr.coords, 1e-3 + r ** 3 * 1e9, pr.nu, rp, ci)
    F = np.zeros(pr.n_dof)
    orc.face_loads_pressure(pr.load_elem, pr.load_face, pr.conn, pr.coords, 1e6, F)
    f0 = F.copy()
    orc.apply_penalty(pr.bc_eq, pr.bc_val, 1e12, rp, ci, val, F)
    uo, st, _, _ = orc.pcg(rp, ci, val, F, rtol=1e-14, max_iter=100000)
    lo, st2, _, _ = orc.cg_transpose(rp, ci, val, f0, tol=1e-13, max_iter=100000)
    go = orc.sensitivity(pr.conn, pr.coords, pr.nu, 3 * r ** 2 * 1e9, uo, lo)
    assert st == 0 and st2 == 0
    assert abs(float(c) - f0 @ uo) < 1e-12 * abs(f0 @ uo)
    assert np.abs(g - go).max() < 1e-8 * np.abs(go).max()
    # central finite differences on seeded elements, h = 1e-6
    picks = np.random.default_rng(20251206).choice(pr.n_elems, 8, replace=False)
    h = 1e-6
    with torch.no_grad():
        for e in picks:
            rp_ = rho.detach().clone()
            rm_ = rho.detach().clone()
            rp_[e] += h
            rm_[e] -= h
            cp, _ = _compliance(model, rp_)
            cm, _ = _compliance(model, rm_)
            fd = (float(cp) - float(cm)) / (2 * h)
            assert abs(fd - g[e]) < 2e-4 * abs(g[e]), (e, fd, g[e])
    model.close()


def test_mixed_precision_gradient_matches_fp64():
    import torch
    pr, m64, rho = _setup(8, 0, 1e-12)
    c, _ = _compliance(m64, rho)
    c.backward()
    g64 = rho.grad.detach().clone()
    m64.close()
    _, mmx, rho2 = _setup(8, 2, 1e-10)
    c2, _ = _compliance(mmx, rho2)
    c2.backward()
    assert mmx.last_forward.n_outer >= 1 and mmx.last_adjoint.converged
    rel = float((rho2.grad - g64).abs().max() / g64.abs().max())
    print(f"[gpu-test] mixed vs fp64 gradient rel {rel:.2e}", flush=True)
    assert rel < 1e-6
    mmx.close()
