! ufc_j2_step.f90 -- Fortran J2 static step over libufc_hip.so: the increment
! loop of RT_StepDriver_Execute (L5_RT/StepDriver/RT_Step_Exec.f90:453-600) and
! the Newton-Raphson iteration of RT_NLSolver_NewtonRaph
! (L5_RT/Solver/RT_Solv_Nonlin.f90:807-1000), paths relative to the reference
! root, with every per-Gauss-point and per-DOF operation on the GPU:
!
!   per increment  prescribed values lam * target (RT_Asm_ApplyBC, penalty)
!     per Newton iteration
!       ufc_ph_j2_update       strains B(u - u_n), radial return from the
!                              committed state, consistent tangent
!       ufc_rt_residual        R = F + P ubar - F_int - P u
!       converged when ||R|| <= 1e-8 ||F_int|| (after the first iteration)
!       ufc_rt_asm_assemble_k  tangent K_t (UFC_MAT_J2) into the fixed CSR
!       ufc_nm_pcg             K_t du = R to the inexact-Newton forcing term
!                              (first iteration 1e-6, then Eisenstat-Walker
!                              choice 2 capped at 1e-3), u += du
!     ufc_ph_j2_commit on convergence; otherwise the increment is cut back by
!     0.25 from the committed state (pnewdt, PH_Mat_Plast_J2_Iso_Core.f90:419-427).
!
! The reference converges on absolute norms (RT_Step_Exec.f90:380-382); the
! test here is relative to the internal force, which is scale-free.
!
! Usage: ufc_j2_step <n> <precision> <increments> <u_out> [restart [mode [stop]]]
!                    [key=value ...]
!   unit cube of n^3 C3D8, interior nodes jittered by U[-0.1, 0.1] h (seeded),
!   E = 200e9, nu = 0.3, sigma_y0 = 250e6, H = 2e9; z = 0 clamped, u_z on
!   z = 1 raised to 0.005 (0.5% strain) in <increments> equal increments.
!   key=value options (anywhere on the line): hard=linear|swift|voce (the
!   PH_MAT_J2_HARD_* law; swift: sigma_y = K (eps0 + e)^n, voce: sigma_y =
!   sigma_y0 + Q (1 - exp(-delta e))), sy0=, h=, k=, n=, eps0=, q=, delta=
!   (law parameters), local_iter= (local Newton limit, 0: 25), max_newton=
!   (Newton iterations per attempt, default 25), predict=1 (the first Newton
!   iterate of an increment extrapolated from the previous converged one,
!   u_n + (dlam / dlam_prev) (u_n - u_n-1), its linear solve at the capped
!   forcing term 1e-3; default 0: from u_n, as RT_NLSolver_NewtonRaph; not
!   used for the first increment, after a resume or on a cut-back retry).
!   precision fp64 | mixed (the linear solve); u_out: n_dof doubles, stream.
!   restart: checkpoint file written after every converged increment with u,
!   the step data and the J2 state (ufc_rt_restart_save; RT_Step_Exec.f90:
!   551-558); '-' for none.  mode: new (default) | resume (continue from the
!   checkpoint at its load factor; increments resume at the nominal size).
!   stop: end after that many converged increments in total (0: run to 1).
! Prints one line per increment, one per cut-back attempt (with its reason)
! and a summary with the number of cut-backs; exits 0 when every increment
! converged, 1 otherwise, 2 on a C-ABI error.
PROGRAM ufc_j2_step
  USE, INTRINSIC :: ISO_C_BINDING
  USE ufc_hip_iface
  IMPLICIT NONE

  INTEGER(C_INT64_T), PARAMETER :: mesh_seed = 20251205_C_INT64_T
  REAL(C_DOUBLE), PARAMETER :: strain = 0.005_C_DOUBLE, tol_r = 1.0E-8_C_DOUBLE
  INTEGER, PARAMETER :: max_cutbacks = 4
  INTEGER :: max_newton, n_cut, npos, ia, eqpos, predict
  CHARACTER(LEN=64) :: key, reason
  CHARACTER(LEN=512) :: pos(7)

  CHARACTER(LEN=64) :: arg, prec_name
  CHARACTER(LEN=512) :: outfile, rstfile
  CHARACTER(LEN=16) :: rstmode
  INTEGER :: stop_after
  TYPE(ufc_rt_restart_data) :: rst
  INTEGER :: n, n_inc, i, it, cutbacks, n_done, u_unit, total_newton, total_linear, lin
  INTEGER(C_INT32_T) :: nn, ne, prec, rc
  INTEGER(C_INT64_T) :: n_yield
  INTEGER(C_INT64_T) :: ndof
  REAL(C_DOUBLE) :: lam, dlam, step, target, rrel, eta, eta_prev, r_prev, dlam_prev
  REAL(C_DOUBLE), TARGET :: norms(2)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: coords(:, :), u(:), u_n(:), u_try(:), r(:), du(:)
  REAL(C_DOUBLE), ALLOCATABLE :: du_prev(:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: bc_target(:), bc_val(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: conn(:, :), bc_eq(:)
  LOGICAL :: ok, failed, have_prev, predicted
  TYPE(C_PTR) :: ctx
  TYPE(ufc_status) :: st
  TYPE(ufc_mesh_desc) :: mesh
  TYPE(ufc_rt_asm_info) :: info
  TYPE(ufc_rt_asm_algo) :: aalgo
  TYPE(ufc_rt_asm_args) :: aargs
  TYPE(ufc_rt_bc_args) :: bc
  TYPE(ufc_nm_solver_algo) :: salgo
  TYPE(ufc_nm_solver_state) :: sstate
  TYPE(ufc_ph_j2_props) :: props

  ! material of the step (PH_J2_Props); key=value options may change it
  props%young = 200.0E9_C_DOUBLE
  props%nu = 0.3_C_DOUBLE
  props%sigma_y0 = 250.0E6_C_DOUBLE
  props%hardening = 2.0E9_C_DOUBLE
  props%hardening_type = UFC_J2_HARD_LINEAR
  max_newton = 25
  predict = 0
  npos = 0
  pos = ' '
  DO ia = 1, COMMAND_ARGUMENT_COUNT()
    CALL GET_COMMAND_ARGUMENT(ia, outfile)
    eqpos = INDEX(outfile, '=')
    IF (eqpos == 0) THEN
      npos = MIN(npos + 1, 7)
      pos(npos) = outfile
      CYCLE
    END IF
    key = outfile(1:eqpos - 1)
    arg = outfile(eqpos + 1:)
    SELECT CASE (TRIM(key))
    CASE ('hard')
      SELECT CASE (TRIM(arg))
      CASE ('linear'); props%hardening_type = UFC_J2_HARD_LINEAR
      CASE ('swift'); props%hardening_type = UFC_J2_HARD_SWIFT
      CASE ('voce'); props%hardening_type = UFC_J2_HARD_VOCE
      CASE DEFAULT
        WRITE(*, '(A,A)') 'unknown hardening law ', TRIM(arg)
        CALL EXIT(2)
      END SELECT
    CASE ('sy0'); READ(arg, *) props%sigma_y0
    CASE ('h'); READ(arg, *) props%hardening
    CASE ('k'); READ(arg, *) props%swift_k
    CASE ('n'); READ(arg, *) props%swift_n
    CASE ('eps0'); READ(arg, *) props%swift_eps0
    CASE ('q'); READ(arg, *) props%voce_q
    CASE ('delta'); READ(arg, *) props%voce_delta
    CASE ('local_iter'); READ(arg, *) props%max_local_iter
    CASE ('max_newton'); READ(arg, *) max_newton
    CASE ('predict'); READ(arg, *) predict
    CASE DEFAULT
      WRITE(*, '(A,A)') 'unknown option ', TRIM(key)
      CALL EXIT(2)
    END SELECT
  END DO
  IF (npos < 4) THEN
    WRITE(*, '(A)') 'usage: ufc_j2_step <n> <precision> <increments> <u_out> ' // &
                    '[restart [mode [stop]]] [key=value ...]'
    CALL EXIT(2)
  END IF
  READ(pos(1), *) n
  prec_name = pos(2)
  READ(pos(3), *) n_inc
  outfile = pos(4)
  rstfile = '-'
  rstmode = 'new'
  stop_after = 0
  IF (npos >= 5) rstfile = pos(5)
  IF (npos >= 6) rstmode = pos(6)
  IF (npos >= 7) READ(pos(7), *) stop_after
  prec = UFC_PREC_FP64
  IF (TRIM(prec_name) == 'mixed') prec = UFC_PREC_MIXED

  nn = INT((n + 1)**3, C_INT32_T)
  ne = INT(n**3, C_INT32_T)
  ndof = 3_C_INT64_T * nn
  ALLOCATE(coords(3, nn), conn(8, ne))
  ALLOCATE(u(ndof), u_n(ndof), u_try(ndof), r(ndof), du(ndof), du_prev(ndof))
  rc = ufc_mesh_box(INT(n, C_INT32_T), INT(n, C_INT32_T), INT(n, C_INT32_T), 1.0_C_DOUBLE, &
                    1.0_C_DOUBLE, 1.0_C_DOUBLE, 0.1_C_DOUBLE, mesh_seed, 1_C_INT32_T, &
                    C_LOC(coords), C_LOC(conn), st)
  CALL check('ufc_mesh_box')

  ! z = 0: all three equations clamped; z = 1: u_z prescribed (1-based eqs)
  ALLOCATE(bc_eq(3 * 2 * (n + 1)**2), bc_target(3 * 2 * (n + 1)**2))
  bc%n_bc = 0
  DO i = 1, nn
    IF (coords(3, i) == 0.0_C_DOUBLE) THEN
      bc_eq(bc%n_bc + 1:bc%n_bc + 3) = [3 * (i - 1) + 1, 3 * (i - 1) + 2, 3 * (i - 1) + 3]
      bc_target(bc%n_bc + 1:bc%n_bc + 3) = 0.0_C_DOUBLE
      bc%n_bc = bc%n_bc + 3
    END IF
  END DO
  DO i = 1, nn
    IF (coords(3, i) == 1.0_C_DOUBLE) THEN
      bc%n_bc = bc%n_bc + 1
      bc_eq(bc%n_bc) = 3 * (i - 1) + 3
      bc_target(bc%n_bc) = strain
    END IF
  END DO
  ALLOCATE(bc_val(bc%n_bc))

  rc = ufc_ctx_create(0_C_INT32_T, ctx, st)
  CALL check('ufc_ctx_create')
  mesh%n_nodes = nn
  mesh%n_elems = ne
  mesh%index_base = 1
  mesh%coords = C_LOC(coords)
  mesh%conn = C_LOC(conn)
  rc = ufc_rt_asm_build_pattern(ctx, mesh, info, st)
  CALL check('ufc_rt_asm_build_pattern')
  rc = ufc_ph_j2_init(ctx, props, st)
  CALL check('ufc_ph_j2_init')

  aalgo%precision = prec
  aalgo%material = UFC_MAT_J2
  aargs%young = C_NULL_PTR
  aargs%nu = props%nu
  bc%index_base = 1
  bc%eq = C_LOC(bc_eq)
  bc%penalty = 1.0E12_C_DOUBLE
  salgo%precision = prec
  salgo%inner_rtol = 1.0E-5_C_DOUBLE

  u = 0.0_C_DOUBLE
  lam = 0.0_C_DOUBLE
  step = 1.0_C_DOUBLE / REAL(n_inc, C_DOUBLE)
  cutbacks = 0
  n_cut = 0
  n_done = 0
  total_newton = 0
  total_linear = 0
  failed = .FALSE.
  have_prev = .FALSE.
  dlam_prev = 0.0_C_DOUBLE
  IF (TRIM(rstmode) == 'resume') THEN
    rst%u = C_LOC(u)
    rc = ufc_rt_restart_load(ctx, TRIM(rstfile) // C_NULL_CHAR, rst, 1_C_INT32_T, st)
    CALL check('ufc_rt_restart_load')
    lam = rst%lambda
    n_done = rst%increment
    WRITE(*, '(A,I0,A,F8.5)') 'resumed after increment ', n_done, ' lambda ', lam
  END IF
  DO WHILE (lam < 1.0_C_DOUBLE - 1.0E-12_C_DOUBLE .AND. &
            (stop_after <= 0 .OR. n_done < stop_after))
    dlam = MIN(step, 1.0_C_DOUBLE - lam)
    target = lam + dlam
    ! element-wise into the fixed array: a whole-array assignment from the
    ! longer bc_target would reallocate bc_val and leave bc%value dangling
    bc_val(1:bc%n_bc) = bc_target(1:bc%n_bc) * target
    bc%value = C_LOC(bc_val)
    rc = ufc_rt_asm_apply_bc(ctx, bc, st)
    CALL check('ufc_rt_asm_apply_bc')
    u_n = u
    predicted = predict /= 0 .AND. have_prev
    IF (predicted) THEN
      u_try = u_n + (dlam / dlam_prev) * du_prev
    ELSE
      u_try = u_n
    END IF
    ok = .FALSE.
    lin = 0
    rrel = HUGE(1.0_C_DOUBLE)
    r_prev = -1.0_C_DOUBLE
    eta_prev = -1.0_C_DOUBLE
    reason = 'newton'
    DO it = 1, max_newton
      rc = ufc_ph_j2_update(ctx, C_LOC(u_try), C_LOC(u_n), n_yield, st)
      IF (rc /= UFC_STATUS_OK) THEN            ! local Newton failure: cut back
        WRITE(*, '(A,I0,A,A)') 'ufc_ph_j2_update status ', rc, ': ', TRIM(ufc_status_message(st))
        reason = 'local'
        EXIT
      END IF
      rc = ufc_rt_residual(ctx, C_LOC(u_try), C_LOC(r), C_LOC(norms), st)
      CALL check('ufc_rt_residual')
      IF (norms(2) > 0.0_C_DOUBLE) rrel = norms(1) / norms(2)
      IF (it > 1 .AND. norms(1) <= tol_r * MAX(norms(2), TINY(1.0_C_DOUBLE))) THEN
        ok = .TRUE.
        EXIT
      END IF
      ! forcing term: the first iteration carries the jump of the prescribed
      ! values and is solved to 1e-6 (1e-3 from a predicted iterate, which
      ! meets them already); then eta = 0.9 (|R_k|/|R_k-1|)^2
      ! (Eisenstat & Walker 1996), safeguarded, capped at 1e-3, and never
      ! tighter than reaching the Newton target needs
      IF (r_prev < 0.0_C_DOUBLE) THEN
        eta = 1.0E-6_C_DOUBLE
        IF (predicted) eta = 1.0E-3_C_DOUBLE
      ELSE
        eta = 0.9_C_DOUBLE * (norms(1) / r_prev)**2
        IF (eta_prev > 0.0_C_DOUBLE .AND. 0.9_C_DOUBLE * eta_prev**2 > 0.1_C_DOUBLE) &
          eta = MAX(eta, 0.9_C_DOUBLE * eta_prev**2)
        eta = MIN(eta, 1.0E-3_C_DOUBLE)
      END IF
      eta = MAX(eta, 0.5_C_DOUBLE * tol_r * norms(2) / MAX(norms(1), TINY(1.0_C_DOUBLE)), &
                1.0E-10_C_DOUBLE)
      eta = MIN(eta, 0.5_C_DOUBLE)
      r_prev = norms(1)
      eta_prev = eta
      salgo%rtol = eta
      rc = ufc_rt_asm_assemble_k(ctx, aalgo, aargs, st)
      CALL check('ufc_rt_asm_assemble_k')
      sstate%x = C_LOC(du)
      sstate%b = C_LOC(r)
      sstate%x0_zero = 1
      rc = ufc_nm_pcg(ctx, salgo, sstate, st)
      lin = lin + sstate%niter
      IF (rc /= UFC_STATUS_OK) THEN
        WRITE(*, '(A,I0,A,A)') 'ufc_nm_pcg status ', rc, ': ', TRIM(ufc_status_message(st))
        reason = 'linear'
        EXIT
      END IF
      u_try = u_try + du
    END DO
    total_linear = total_linear + lin
    IF (ok) THEN
      rc = ufc_ph_j2_commit(ctx, st)
      CALL check('ufc_ph_j2_commit')
      du_prev = u_try - u_n
      dlam_prev = dlam
      have_prev = .TRUE.
      u = u_try
      lam = target
      n_done = n_done + 1
      total_newton = total_newton + it
      cutbacks = 0
      WRITE(*, '(A,I0,A,F8.5,A,I0,A,I0,A,ES10.3,A,I0)') 'increment ', n_done, ' lambda ', lam, &
           ' newton ', it, ' linear ', lin, ' rrel ', rrel, ' n_yield ', n_yield
      IF (TRIM(rstfile) /= '-') THEN
        rst%time = lam
        rst%lambda = lam
        rst%increment = INT(n_done, C_INT32_T)
        rst%iterations = INT(it, C_INT32_T)
        rst%residual_norm = norms(1)
        rst%converged = 1
        rst%u = C_LOC(u)
        rc = ufc_rt_restart_save(ctx, TRIM(rstfile) // C_NULL_CHAR, rst, 1_C_INT32_T, st)
        CALL check('ufc_rt_restart_save')
      END IF
    ELSE
      ! the trial state is discarded: the next attempt starts again from the
      ! committed state and u_n (RT_Step_Exec.f90:455-471)
      WRITE(*, '(A,F8.5,A,I0,A,ES10.3,A,A)') 'cutback at lambda ', target, ' after newton ', &
           MIN(it, max_newton), ' rrel ', rrel, ' reason ', TRIM(reason)
      cutbacks = cutbacks + 1
      n_cut = n_cut + 1
      have_prev = .FALSE.  ! the retry starts from u_n
      IF (cutbacks > max_cutbacks) THEN
        failed = .TRUE.
        EXIT
      END IF
      step = step * 0.25_C_DOUBLE
    END IF
  END DO

  OPEN(NEWUNIT=u_unit, FILE=TRIM(outfile), ACCESS='STREAM', FORM='UNFORMATTED', &
       STATUS='REPLACE', ACTION='WRITE')
  WRITE(u_unit) u
  CLOSE(u_unit)
  WRITE(*, '(A,I0,A,I0,A,I0,A,I0,A,I0,A,L1)') 'ufc_j2_step n_dof ', info%n_dof, &
       ' increments ', n_done, ' newton ', total_newton, ' linear ', total_linear, &
       ' cutbacks ', n_cut, ' converged ', .NOT. failed
  rc = ufc_ctx_destroy(ctx, st)
  IF (failed) CALL EXIT(1)

CONTAINS

  SUBROUTINE check(what)
    CHARACTER(LEN=*), INTENT(IN) :: what
    IF (rc /= UFC_STATUS_OK) THEN
      WRITE(*, '(A,A,I0,A,A)') what, ' failed, status ', rc, ': ', TRIM(ufc_status_message(st))
      CALL EXIT(2)
    END IF
  END SUBROUTINE check

END PROGRAM ufc_j2_step
