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


VOCE = dict(young=200e9, nu=0.3, sigma_y0=250e6, hardening_type=3, voce_q=150e6,
            voce_delta=30.0)
SWIFT = dict(young=200e9, nu=0.3, sigma_y0=500e6 * 0.01 ** 0.2, hardening_type=2,
             swift_k=500e6, swift_n=0.2, swift_eps0=0.01)


def _oracle_job(orc, pr, props, target, load_factors):
    """orc_j2_newton_step over the given load factors (state committed after
    each); returns u and the state."""
    state = np.zeros(pr.n_elems * 8 * 13)
    uo = np.zeros(pr.n_dof)
    for lf in load_factors:
        st, uo, state, nn, rn = orc.j2_newton_step(pr.coords.shape[0], pr.conn, pr.coords,
                                                   props, pr.bc_eq, target * lf, 1e12, state,
                                                   uo)
        assert st == 0, lf
    return uo, state.reshape(-1, 13)


@pytest.mark.parametrize("precision", [0, 2])
def test_j2_predictor_same_increments(ctx, orc, precision):
    """The extrapolated first Newton iterate (predictor=True) changes the
    Newton path only: on the 12^3 jittered J2 cube every converged increment
    equals the plain job's (1e-7), the oracle's residual of each increment,
    with the state replayed from the GPU's increments, is below 2e-8 of
    ||F_int||, and the linear iterations drop."""
    from ufc_hip.nonlinear import j2_static, to_host
    pr, target = _problem(12)
    props = pr.extra["props"]
    jobs = {}
    for pred in (False, True):
        ctx.build_pattern(pr.coords, pr.conn)
        print(f"[gpu-test] j2 12^3 precision {precision} predictor {pred}", flush=True)
        res = j2_static(ctx, pr.bc_eq, target, props, n_increments=10, precision=precision,
                        keep_increments=True, predictor=pred)
        assert res.converged and len(res.u_increments) == 10 and res.cutbacks == 0
        jobs[pred] = res
    lin = {k: sum(i.linear_iters for i in r.increments) for k, r in jobs.items()}
    newton = {k: [i.newton_iters for i in r.increments] for k, r in jobs.items()}
    print(f"[gpu-test] j2 12^3 linear iterations {lin}, Newton {newton}", flush=True)
    assert lin[True] < lin[False]
    for a, b in zip(jobs[False].u_increments, jobs[True].u_increments):
        a, b = to_host(a), to_host(b)
        assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-7
    nn = pr.coords.shape[0]
    state = np.zeros(pr.n_elems * 8 * 13)
    u_prev = np.zeros(pr.n_dof)
    for k, uk in enumerate(jobs[True].u_increments, start=1):
        uk = to_host(uk)
        nf, state, _, fi = orc.j2_internal_force(nn, pr.conn, pr.coords, props, uk, u_prev, state)
        assert nf == 0
        r = -fi
        r[pr.bc_eq] += 1e12 * (target * (k / 10.0) - uk[pr.bc_eq])
        assert np.linalg.norm(r) / np.linalg.norm(fi) < 2e-8
        u_prev = uk


@pytest.mark.parametrize("law", ["swift", "voce"])
@pytest.mark.parametrize("precision", [0, 2])
def test_j2_hardening_laws_match_oracle(ctx, orc, law, precision):
    """Swift and Voce hardening (PH_J2_ComputeHardening,
    PH_Mat_Plast_J2_Iso_Core.f90:549-605; the local Newton iterates): the GPU
    Newton loop over 10 increments against the oracle's, displacements 1e-7,
    stresses and equivalent plastic strains 1e-6."""
    from ufc_hip.nonlinear import j2_static
    pr, target = _problem(4)
    props = SWIFT if law == "swift" else VOCE
    ctx.build_pattern(pr.coords, pr.conn)
    print(f"[gpu-test] j2 {law} 4^3 precision {precision}", flush=True)
    res = j2_static(ctx, pr.bc_eq, target, props, n_increments=10, precision=precision,
                    lin_rtol=1e-12)
    assert res.converged and len(res.increments) == 10 and res.cutbacks == 0
    assert res.increments[-1].n_yield > 0
    sg = ctx.j2_get_state().reshape(-1, 13)
    uo, so = _oracle_job(orc, pr, props, target, [k / 10.0 for k in range(1, 11)])
    err = np.linalg.norm(res.u - uo) / np.linalg.norm(uo)
    print(f"[gpu-test] j2 {law} u err {err:.2e} newton "
          f"{[i.newton_iters for i in res.increments]}", flush=True)
    assert err < 1e-7
    assert np.abs(sg[:, :6] - so[:, :6]).max() < 1e-6 * np.abs(so[:, :6]).max()
    assert np.abs(sg[:, 6] - so[:, 6]).max() < 1e-6 * so[:, 6].max()


def test_j2_unsupported_kinematic_and_bad_laws(ctx):
    """Armstrong-Frederick backstress is refused (the reference's never
    enters its yield function); invalid law parameters are errors."""
    from ufc_hip import _abi as A
    pr, _ = _problem(2)
    ctx.build_pattern(pr.coords, pr.conn)
    for kw, code in ((dict(hardening_type=4), A.STATUS_UNSUPPORTED),
                     (dict(af_c=1e9), A.STATUS_UNSUPPORTED),
                     (dict(hardening_type=2, swift_k=5e8, swift_n=0.2, swift_eps0=0.0),
                      A.STATUS_ERROR),
                     (dict(hardening_type=3, voce_q=1e8, voce_delta=-1.0), A.STATUS_ERROR),
                     (dict(hardening_type=7), A.STATUS_INVALID)):
        with pytest.raises(A.UfcError) as e:
            ctx.j2_init(200e9, 0.3, 250e6, **kw)
        assert e.value.code == code, kw


def test_j2_cutback_after_local_failure(ctx, orc):
    """A local-Newton limit of 2 with Voce hardening: the first increment to
    half the load fails at a Gauss point (the update returns
    UFC_STATUS_ERROR, the reference's pnewdt = 0.25 signal,
    PH_Mat_Plast_J2_Iso_Core.f90:419-427), the driver rolls back to the
    committed state and continues at a quarter of the step
    (RT_Step_Exec.f90:455-471, 575-590).  The cut job takes the load factors
    0.125 k, reports one cut-back, and ends equal to an uncut job on those
    load factors and to the oracle's."""
    from ufc_hip.nonlinear import j2_static
    pr, target = _problem(4)
    ctx.build_pattern(pr.coords, pr.conn)
    limited = dict(VOCE, max_local_iter=2)
    print("[gpu-test] j2 cutback job", flush=True)
    cut = j2_static(ctx, pr.bc_eq, target, limited, n_increments=2, precision=0,
                    lin_rtol=1e-12)
    assert cut.converged and cut.cutbacks == 1
    assert cut.failures[0][0] == 0.5 and "local Newton failure" in cut.failures[0][1]
    assert [i.load_factor for i in cut.increments] == [0.125 * k for k in range(1, 9)]
    s_cut = ctx.j2_get_state()
    ctx.build_pattern(pr.coords, pr.conn)
    print("[gpu-test] j2 uncut job", flush=True)
    ref = j2_static(ctx, pr.bc_eq, target, VOCE, n_increments=8, precision=0, lin_rtol=1e-12)
    assert ref.converged and ref.cutbacks == 0
    s_ref = ctx.j2_get_state()
    d = np.linalg.norm(cut.u - ref.u) / np.linalg.norm(ref.u)
    print(f"[gpu-test] cut vs uncut {d:.2e}", flush=True)
    assert d < 1e-12
    assert np.abs(s_cut - s_ref).max() <= 1e-12 * np.abs(s_ref).max()
    uo, so = _oracle_job(orc, pr, VOCE, target, [0.125 * k for k in range(1, 9)])
    assert np.linalg.norm(cut.u - uo) / np.linalg.norm(uo) < 1e-7
    assert np.abs(s_cut.reshape(-1, 13)[:, 6] - so[:, 6]).max() < 1e-6 * so[:, 6].max()


def test_j2_failed_attempt_keeps_committed_state(ctx):
    """The committed state is bitwise unchanged by an attempt whose update
    fails: three converged increments (Voce, local limit 2; the material
    has yielded), then a jump to twice the load whose second Newton
    iteration fails locally; ufc_ph_j2_get_state before and after is
    identical."""
    from ufc_hip import _abi as A
    pr, target = _problem(4)
    ctx.build_pattern(pr.coords, pr.conn)
    ctx.j2_init(**dict(VOCE, max_local_iter=2))

    def newton(lf, u_n, max_it=25):
        ctx.apply_bc(pr.bc_eq, target * lf, 1e12)
        u = u_n.copy()
        for it in range(1, max_it + 1):
            ctx.j2_update(u, u_n)
            r, norms = ctx.residual(u)
            if it > 1 and norms[0] <= 1e-8 * norms[1]:
                return u
            ctx.assemble_k(None, 0.3, precision=0, material=A.MAT_J2)
            du, info = ctx.pcg(precision=0, rtol=1e-12, b=r)
            u = u + du
        raise AssertionError("no convergence")

    u = np.zeros(pr.n_dof)
    for lf in (0.125, 0.25, 0.375):
        u = newton(lf, u)
        ctx.j2_commit()
    s1 = ctx.j2_get_state().copy()
    assert s1[..., 6].max() > 0.0  # yielded: a non-trivial committed state
    print("[gpu-test] j2 failing jump", flush=True)
    with pytest.raises(A.UfcError) as e:
        newton(0.75, u)
    assert e.value.code == A.STATUS_ERROR and "local Newton failed" in e.value.message
    assert np.array_equal(ctx.j2_get_state(), s1)


def test_j2_newton_failure_cuts_back_and_stops(ctx, tmp_path):
    """max_newton = 2 (no yielding increment converges in two iterations):
    increments that stay elastic converge, the others fail and are cut back
    by 0.25 until max_cutbacks consecutive cuts are exceeded; the job reports
    non-convergence, and its committed state and u are those of the last
    converged increment (its checkpoint, written after that commit, holds
    them bit for bit)."""
    from ufc_hip import restart as R
    from ufc_hip.nonlinear import j2_static
    pr, target = _problem(4)
    ctx.build_pattern(pr.coords, pr.conn)
    props = pr.extra["props"]
    ck = str(tmp_path / "last.rst")
    res = j2_static(ctx, pr.bc_eq, target, props, n_increments=10, precision=0,
                    lin_rtol=1e-12, max_newton=2, max_cutbacks=2, restart_file=ck)
    print(f"[gpu-test] j2 newton-failure job: {len(res.increments)} increments, "
          f"{res.cutbacks} cut-backs", flush=True)
    assert not res.converged and len(res.increments) >= 1 and res.cutbacks >= 3
    assert all("Newton iterations" in f[1] for f in res.failures)
    s_fail = ctx.j2_get_state().copy()
    u_ck, rs = R.load(ck, ctx=ctx, with_state=True)
    assert rs.increment == len(res.increments)
    assert rs.lambda_ == res.increments[-1].load_factor
    assert np.array_equal(u_ck, res.u)
    assert np.array_equal(ctx.j2_get_state(), s_fail)


@pytest.mark.parametrize("law", ["linear", "swift", "voce"])
@pytest.mark.parametrize("precision", [0, 2])
def test_j2_uniaxial_stress_closed_form(ctx, law, precision):
    """A boundary-value problem with a closed form through the GPU Newton
    loop: a 4x4x4 unit cube on rollers, u_z on top driven to 1% strain in ten
    increments.  The FE solution is homogeneous uniaxial stress, so at 0.5%
    and 1% every Gauss point's stress and plastic strain, and the lateral
    face, equal the law's closed form (bisection on sigma = sigma_y(peeq),
    not the radial return).  Newton stops at 1e-7 of the internal force (the
    penalty rows resolve 1e12 ulp(u))."""
    from ufc_hip import static as S
    from ufc_hip.nonlinear import j2_static
    from j2_cases import J2_LAWS_MPA, check_uniaxial_state, uniaxial_bvp
    coords, conn = S.mesh_box(4, 4, 4)
    eq, top, x1 = uniaxial_bvp(coords)
    props = J2_LAWS_MPA[law]
    print(f"[gpu-test] j2 uniaxial {law} precision {precision}", flush=True)
    for stop, e33 in ((5, 0.005), (None, 0.01)):
        ctx.build_pattern(coords, conn)
        res = j2_static(ctx, eq, 0.01 * top, props, n_increments=10, precision=precision,
                        tol_r=1e-7, lin_rtol=1e-12, stop_after=stop)
        assert res.converged and res.cutbacks == 0
        assert abs(res.increments[-1].load_factor - e33 / 0.01) < 1e-12
        check_uniaxial_state(props, e33, res.u, ctx.j2_get_state(), x1)
