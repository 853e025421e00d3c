"""Small-strain static Newton-Raphson with the J2 material over the C-ABI.

Mirrors the increment loop of RT_StepDriver_Execute (RT_Step_Exec.f90:453-600)
and RT_NLSolver_NewtonRaph (RT_Solv_Nonlin.f90:807-1000), paths relative to the
reference root, with the hot parts on the GPU:

  per Newton iteration
    ufc_ph_j2_update     strain increments B(u - u_n), radial return from the
                         committed state, consistent tangent  (PH_J2_ComputeStress)
    ufc_rt_residual      R = F + P ubar - F_int(sigma) - P u   (RT_Asm_ComputeResidual)
    converged when ||R|| <= tol_r ||F_int|| (after the first iteration)
    ufc_rt_asm_assemble_k(UFC_MAT_J2)  tangent K_t into the fixed CSR pattern
    ufc_nm_pcg           K_t du = R   (FP64 or FP32-in-FP64 refinement), to the
                         inexact-Newton forcing term (Eisenstat-Walker)
    u += du
  per increment: ufc_ph_j2_commit on convergence; a failed increment
  (local-Newton failure or no convergence) is cut back by 0.25
  (pnewdt, PH_Mat_Plast_J2_Iso_Core.f90:419-427) from the committed state.

The reference converges on absolute norms ||R|| < 1e-6 and ||du|| < 1e-5
(RT_Step_Exec.f90:380-382); in SI units those are unattainable, so the test
here is relative to the internal force.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _abi as A
from . import restart as R
from .static import Context


@dataclass
class IncrementInfo:
    load_factor: float
    newton_iters: int
    linear_iters: int
    rnorm_rel: float
    n_yield: int
    history: list = field(default_factory=list)  # ||R|| / ||F_int|| per Newton iteration


@dataclass
class NewtonResult:
    converged: bool
    increments: list = field(default_factory=list)
    u: np.ndarray = None
    cutbacks: int = 0          # failed attempts cut back by 0.25 over the job
    failures: list = field(default_factory=list)  # (load factor tried, reason) per cut
    j2_update_ms: list = field(default_factory=list)  # device time of every J2 update
    assemble_ms: list = field(default_factory=list)   # device time of every tangent assembly
    u_increments: list = field(default_factory=list)  # converged u per increment (keep_increments;
    #                                                   device tensors when the job ran on the GPU)


def forcing_term(r_now, r_prev, eta_prev, target, eta_first=1e-6, eta_max=1e-3,
                 eta_min=1e-10):
    """Linear-solve tolerance of an inexact Newton step.  The first iteration
    of an increment carries the jump of the prescribed values (a residual on
    the penalty rows only) and is solved to eta_first, so that the interior
    sees the new boundary displacement before plasticity is evaluated on it;
    later iterations use Eisenstat & Walker (1996) choice 2,
    eta = 0.9 (|R_k|/|R_k-1|)^2 (safeguarded by 0.9 eta_prev^2), capped at
    eta_max and never tighter than reaching `target` from |R_k| needs."""
    if r_prev is None:
        eta = eta_first
    else:
        eta = 0.9 * (r_now / r_prev) ** 2
        if eta_prev is not None and 0.9 * eta_prev ** 2 > 0.1:
            eta = max(eta, 0.9 * eta_prev ** 2)
        eta = min(eta, eta_max)
    eta = max(eta, 0.5 * target / max(r_now, 1e-300), eta_min)
    return min(eta, 0.5)


def j2_static(ctx: Context, bc_eq, bc_target, props, n_increments=10, precision=A.PREC_MIXED,
              tol_r=1e-8, lin_rtol=None, max_newton=25, penalty=1.0e12, max_cutbacks=4,
              restart_file=None, checkpoint_freq=1, resume_from=None, stop_after=None,
              keep_increments=False, predictor=False, replace_rtol=0.0):
    """Drives prescribed values bc_target (reached at load factor 1) on bc_eq
    in n_increments equal increments.  ctx must hold the mesh (build_pattern).
    props: (E, nu, sigma_y0, H) or a dict of Context.j2_init keywords (the
    hardening law, max_local_iter).  A failed attempt (local-Newton failure,
    no Newton convergence in max_newton iterations, a failed linear solve) is
    retried from the committed state with the step cut by 0.25, at most
    max_cutbacks times in a row; res.cutbacks counts them over the job.
    lin_rtol None: inexact Newton, each linear solve to the forcing term of
    forcing_term(); a number: every linear solve to that relative tolerance.

    restart_file: u, the step data and the committed J2 state are saved there
    after every checkpoint_freq-th converged increment (RT_Step_Exec.f90:551-558).
    resume_from: a restart file to continue from (load factor, increment count,
    u and the J2 state); increments then resume at the nominal size
    1/n_increments.  stop_after: return after that many converged increments
    in total (a job that ends early, for restarts).  keep_increments: the
    converged u of every increment in res.u_increments (the path a
    history-dependent material followed, e.g. to replay it elsewhere; device
    tensors when the job ran on the GPU, to_host() copies one to the host).

    predictor: the first Newton iterate of an increment is extrapolated from
    the previous converged increment, u_n + (dlam / dlam_prev) (u_n - u_n-1),
    instead of u_n (RT_NLSolver_NewtonRaph starts from u_n): it meets the new
    prescribed values (proportional to the load factor) and carries the
    previous increment's deformation, so the first linear solve is an
    ordinary inexact-Newton step (tolerance eta_max) instead of the
    boundary-jump solve to eta_first.  The stress update still goes from the
    committed state over the whole increment u - u_n, so the converged
    increment is the same to the Newton tolerance.  Not used for the first
    increment, after a resume, or on the retry of a failed attempt.
    replace_rtol: the mixed linear solves' FP64 residual replacement interval
    (Context.pcg: <= 0 the library's 0.1, >= 1 restarted refinement steps)."""
    bc_eq = np.ascontiguousarray(bc_eq, np.int32)
    bc_target = np.asarray(bc_target, np.float64)
    if isinstance(props, dict):
        ctx.j2_init(**props)
        nu = props["nu"]
    else:
        ctx.j2_init(*props)
        nu = props[1]
    n = ctx.n_dof
    # the Newton vectors live on the context's GPU (the C-ABI takes device
    # pointers): no host round trip of u, R and du per Newton iteration
    dev = _device(ctx)
    u = _zeros(n, dev)
    lam = 0.0
    count = 0
    if resume_from is not None:
        u_h, rs = R.load(resume_from, ctx=ctx, with_state=True)
        u = _to(u_h, dev)
        lam, count = rs.lambda_, rs.increment
    r_buf, du_buf = _zeros(n, dev), _zeros(n, dev)
    step = 1.0 / n_increments
    res = NewtonResult(converged=True)
    cutbacks = 0
    du_prev = dlam_prev = None  # the last converged increment (predictor)
    while lam < 1.0 - 1e-12 and (stop_after is None or count < stop_after):
        dlam = min(step, 1.0 - lam)
        target = lam + dlam
        ctx.apply_bc(bc_eq, bc_target * target, penalty)
        u_n = u.clone() if dev is not None else u.copy()
        ok, its, lin, rrel, ny, rabs = False, 0, 0, np.inf, 0, np.inf
        hist = []
        predicted = predictor and du_prev is not None
        u_try = u_n + (dlam / dlam_prev) * du_prev if predicted else u_n + 0.0
        r_prev = eta = None
        reason = f"no convergence in {max_newton} Newton iterations"
        for it in range(1, max_newton + 1):
            try:
                ny = ctx.j2_update(u_try, u_n)
            except A.UfcError as e:
                reason = f"local Newton failure ({e.message})"
                break  # cut back
            res.j2_update_ms.append(ctx.timing().j2_update_ms)
            r, norms = ctx.residual(u_try, out=r_buf)
            its = it
            rabs = norms[0]
            rrel = norms[0] / norms[1] if norms[1] > 0.0 else float("inf")
            hist.append(rrel)
            if it > 1 and norms[0] <= tol_r * max(norms[1], 1e-300):
                ok = True
                break
            if lin_rtol is None:
                eta = forcing_term(norms[0], r_prev, eta, tol_r * norms[1],
                                   eta_first=1e-3 if predicted else 1e-6)
            else:
                eta = lin_rtol
            r_prev = norms[0]
            ctx.assemble_k(None, nu, precision=precision, material=A.MAT_J2)
            res.assemble_ms.append(ctx.timing().assemble_ms)
            du, info = ctx.pcg(precision=precision, rtol=eta, b=r, x=du_buf,
                               replace_rtol=replace_rtol)
            lin += info.niter
            if not info.converged:
                reason = f"linear solve status {info.status}"
                break
            u_try = u_try + du
        if ok:
            ctx.j2_commit()
            du_prev, dlam_prev = u_try - u_n, dlam
            u = u_try
            lam = target
            count += 1
            res.increments.append(IncrementInfo(lam, its, lin, rrel, ny, hist))
            if keep_increments:
                # u is never modified in place (each Newton step makes a new
                # array), so keeping the reference costs nothing
                res.u_increments.append(u)
            cutbacks = 0
            if restart_file is not None and count % checkpoint_freq == 0:
                R.save(restart_file, to_host(u), R.StepState(lam, count, lam, rabs, its, True),
                       ctx=ctx, with_state=True)
        else:
            # the trial state is discarded: the next attempt starts again from
            # the committed state and u_n (RT_Step_Exec.f90:455-471)
            cutbacks += 1
            res.cutbacks += 1
            du_prev = None  # the retry starts from u_n
            res.failures.append((target, reason))
            if cutbacks > max_cutbacks:
                res.converged = False
                break
            step *= 0.25
    res.u = to_host(u)
    return res


def _device(ctx):
    """The torch device of ctx's GPU, or None (host arrays) without torch/CUDA."""
    try:
        import torch
    except ImportError:
        return None
    return torch.device("cuda", ctx.device) if torch.cuda.is_available() else None


def _zeros(n, dev):
    if dev is None:
        return np.zeros(n)
    import torch
    return torch.zeros(n, dtype=torch.float64, device=dev)


def _to(a, dev):
    if dev is None:
        return np.asarray(a, np.float64)
    import torch
    return torch.as_tensor(np.asarray(a, np.float64), device=dev)


def to_host(a):
    """A host float64 array of a numpy array or a (device) tensor (e.g. the
    entries of NewtonResult.u_increments)."""
    return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a, np.float64)
