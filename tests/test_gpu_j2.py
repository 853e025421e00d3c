"""J2 plasticity (jittered cube, prescribed top displacement, small mesh): the GPU Newton loop over the
C-ABI against the oracle's Newton loop, increment by increment; FP64 and
FP32-in-FP64 inner solves converge to the same state."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _problem(n):
    from ufc_hip import problems as P
    pr = P.j2_cube(n, increments=10)
    top = pr.extra["top_eq"]
    target = np.zeros(pr.bc_eq.size)
    target[pr.extra["n_clamp"]:] = pr.extra["strain"]  # u_z on z = 1 (unit height)
    assert np.array_equal(pr.bc_eq[pr.extra["n_clamp"]:], top)
    return pr, target


@pytest.mark.parametrize("precision", [0, 2])
def test_j2_newton_matches_oracle(ctx, orc, precision):
    from ufc_hip.nonlinear import j2_static
    pr, target = _problem(4)
    props = pr.extra["props"]
    ctx.build_pattern(pr.coords, pr.conn)
    print(f"[gpu-test] j2 4^3 precision {precision}", flush=True)
    # exact Newton (linear solves to 1e-10, as the oracle's) ...
    res = j2_static(ctx, pr.bc_eq, target, props, n_increments=10, precision=precision,
                    lin_rtol=1e-10)
    assert res.converged and len(res.increments) == 10
    assert max(i.newton_iters for i in res.increments) <= 8
    assert res.increments[-1].n_yield > 0
    sg_exact = ctx.j2_get_state().ravel()
    # ... and inexact Newton (forcing terms) reach the same state
    ctx.build_pattern(pr.coords, pr.conn)
    res_in = j2_static(ctx, pr.bc_eq, target, props, n_increments=10, precision=precision)
    assert res_in.converged and len(res_in.increments) == 10
    print(f"[gpu-test] j2 linear iterations exact {sum(i.linear_iters for i in res.increments)} "
          f"inexact {sum(i.linear_iters for i in res_in.increments)}", flush=True)
    assert np.linalg.norm(res_in.u - res.u) / np.linalg.norm(res.u) < 1e-7
    # oracle, same increments
    n_gp = pr.n_elems * 8
    state = np.zeros(n_gp * 13)
    uo = np.zeros(pr.n_dof)
    for k in range(1, 11):
        st, uo, state, nn, rn = orc.j2_newton_step(pr.coords.shape[0], pr.conn, pr.coords,
                                                   props, pr.bc_eq, target * (k / 10.0), 1e12,
                                                   state, uo)
        assert st == 0
    err = np.linalg.norm(res.u - uo) / np.linalg.norm(uo)
    print(f"[gpu-test] j2 u err {err:.2e}", flush=True)
    assert err < 1e-7
    sg = sg_exact
    peeq_g, peeq_o = sg.reshape(-1, 13)[:, 6], state.reshape(-1, 13)[:, 6]
    assert np.abs(peeq_g - peeq_o).max() < 1e-6 * peeq_o.max()
    sig_g, sig_o = sg.reshape(-1, 13)[:, :6], state.reshape(-1, 13)[:, :6]
    assert np.abs(sig_g - sig_o).max() < 1e-6 * np.abs(sig_o).max()
