! ufc_inp_step.f90 -- static C3D8 step of an input deck over libufc_hip.so.
!
! The deck is read by ufc_ap_read_inp (the keyword reading of
! AP_Input_Domain_ParseKeyword, L6_AP/Input/Parser/AP_Inp_Mgr.f90:99-125, for
! the keywords of this path); the static sequence then follows
! RT_StepDriver_Execute (L5_RT/StepDriver/RT_Step_Exec.f90:267-676):
!   sparse pattern       RT_Solv_Sparse.f90:336-447
!   global stiffness     RT_Asm_GlobalStiffness    RT_Asm_Solv.f90:1762-1980
!   global load          RT_Asm_GlobalLoad         RT_Asm_Solv.f90:1443-1760
!                        (consistent face pressures, then concentrated loads)
!   boundary conditions  RT_Asm_ApplyBC (penalty)  RT_Asm_Solv.f90:318-344
!   linear solve         NM_Solver_PCG             NM_Solv_Core.f90:363-423
! All arrays stay where ufc_ap_read_inp put them (0-based, library-owned).
!
! Usage: ufc_inp_step <deck.inp> <precision> <u_out> [rtol]
!   precision fp64 | mixed
!   u_out     displacements, n_dof doubles in deck node order, unformatted stream
!   rtol      stop at ||r||/||b|| < rtol (default 1e-10); ||b|| holds the
!             penalty terms of prescribed displacements
! Prints one summary line; exits 0 when converged, 1 when not, 2 on an error
! (the message names the deck line for a deck the reader rejects).
PROGRAM ufc_inp_step
  USE, INTRINSIC :: ISO_C_BINDING
  USE ufc_hip_iface
  IMPLICIT NONE

  CHARACTER(LEN=512) :: deck, outfile
  CHARACTER(LEN=64) :: prec_name, arg
  INTEGER :: i, j, ng, u_unit
  INTEGER(C_INT32_T) :: prec, rc
  REAL(C_DOUBLE) :: rtol
  TYPE(ufc_ap_model) :: model
  TYPE(C_PTR) :: ctx
  TYPE(ufc_status) :: st
  TYPE(ufc_mesh_desc) :: mesh
  TYPE(ufc_rt_asm_info) :: info
  TYPE(ufc_rt_asm_algo) :: aalgo
  TYPE(ufc_rt_asm_args) :: aargs
  TYPE(ufc_rt_bc_args) :: bc
  TYPE(ufc_rt_load_args) :: ld
  TYPE(ufc_nm_solver_algo) :: salgo
  TYPE(ufc_nm_solver_state) :: sstate
  TYPE(ufc_timing) :: tm
  INTEGER(C_INT32_T), POINTER :: face_elem(:), face_id(:), cload_eq(:)
  REAL(C_DOUBLE), POINTER :: face_p(:), cload_val(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: g_elem(:), g_face(:)
  LOGICAL, ALLOCATABLE :: done(:)
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: u(:), f(:)

  CALL GET_COMMAND_ARGUMENT(1, deck)
  CALL GET_COMMAND_ARGUMENT(2, prec_name)
  CALL GET_COMMAND_ARGUMENT(3, outfile)
  rtol = 1.0E-10_C_DOUBLE
  IF (COMMAND_ARGUMENT_COUNT() >= 4) THEN
    CALL GET_COMMAND_ARGUMENT(4, arg)
    READ(arg, *) rtol
  END IF
  prec = UFC_PREC_FP64
  IF (TRIM(prec_name) == 'mixed') prec = UFC_PREC_MIXED

  rc = ufc_ap_read_inp(TRIM(deck) // C_NULL_CHAR, model, st)
  CALL check('ufc_ap_read_inp')

  rc = ufc_ctx_create(0_C_INT32_T, ctx, st)
  CALL check('ufc_ctx_create')
  mesh%n_nodes = model%n_nodes
  mesh%n_elems = model%n_elems
  mesh%index_base = 0
  mesh%coords = model%coords
  mesh%conn = model%conn
  rc = ufc_rt_asm_build_pattern(ctx, mesh, info, st)
  CALL check('ufc_rt_asm_build_pattern')

  aalgo%precision = prec
  aalgo%material = UFC_MAT_ELASTIC
  aalgo%variant = model%variant            ! C3D8R, C3D8I, ... from *ELEMENT, TYPE=
  aargs%young = model%young
  aargs%nu = model%nu
  rc = ufc_rt_asm_assemble_k(ctx, aalgo, aargs, st)
  CALL check('ufc_rt_asm_assemble_k')

  ! ---- F: one library call per distinct face pressure, accumulated
  ALLOCATE(u(info%n_dof), f(info%n_dof))
  f = 0.0_C_DOUBLE
  rc = ufc_rt_set_rhs(ctx, C_LOC(f), st)
  CALL check('ufc_rt_set_rhs')
  IF (model%n_face > 0) THEN
    CALL C_F_POINTER(model%face_elem, face_elem, [model%n_face])
    CALL C_F_POINTER(model%face_id, face_id, [model%n_face])
    CALL C_F_POINTER(model%face_p, face_p, [model%n_face])
    ALLOCATE(g_elem(model%n_face), g_face(model%n_face), done(model%n_face))
    done = .FALSE.
    DO i = 1, model%n_face
      IF (done(i)) CYCLE
      ng = 0
      DO j = i, model%n_face
        IF (.NOT. done(j) .AND. face_p(j) == face_p(i)) THEN
          ng = ng + 1
          g_elem(ng) = face_elem(j)
          g_face(ng) = face_id(j)
          done(j) = .TRUE.
        END IF
      END DO
      ld%n_faces = INT(ng, C_INT32_T)
      ld%kind = UFC_LOAD_PRESSURE
      ld%index_base = 0
      ld%accumulate = 1
      ld%elem = C_LOC(g_elem)
      ld%face = C_LOC(g_face)
      ld%magnitude = face_p(i)
      rc = ufc_rt_asm_load(ctx, ld, st)
      CALL check('ufc_rt_asm_load')
    END DO
  END IF
  IF (model%n_cload > 0) THEN
    CALL C_F_POINTER(model%cload_eq, cload_eq, [model%n_cload])
    CALL C_F_POINTER(model%cload_val, cload_val, [model%n_cload])
    rc = ufc_rt_get_rhs(ctx, C_LOC(f), st)
    CALL check('ufc_rt_get_rhs')
    DO i = 1, model%n_cload
      f(cload_eq(i) + 1) = f(cload_eq(i) + 1) + cload_val(i)
    END DO
    rc = ufc_rt_set_rhs(ctx, C_LOC(f), st)
    CALL check('ufc_rt_set_rhs')
  END IF

  bc%n_bc = model%n_bc
  bc%index_base = 0
  bc%eq = model%bc_eq
  bc%value = model%bc_val
  bc%penalty = 1.0E12_C_DOUBLE
  rc = ufc_rt_asm_apply_bc(ctx, bc, st)
  CALL check('ufc_rt_asm_apply_bc')

  salgo%precision = prec
  salgo%rtol = rtol
  salgo%inner_rtol = 1.0E-5_C_DOUBLE
  sstate%x = C_LOC(u)
  sstate%b = C_NULL_PTR
  sstate%x0_zero = 1
  rc = ufc_nm_pcg(ctx, salgo, sstate, st)
  IF (rc /= UFC_STATUS_OK) THEN
    WRITE(*, '(A,I0,A,A)') 'ufc_nm_pcg status ', rc, ': ', TRIM(ufc_status_message(st))
  END IF
  rc = ufc_ctx_timing(ctx, tm, st)

  OPEN(NEWUNIT=u_unit, FILE=TRIM(outfile), ACCESS='STREAM', FORM='UNFORMATTED', &
       STATUS='REPLACE', ACTION='WRITE')
  WRITE(u_unit) u
  CLOSE(u_unit)
  WRITE(*, '(A,I0,A,I0,A,I0,A,I0,A,I0,A,I0,A,ES12.5,A,F10.3)') 'ufc_inp_step nodes ', &
       model%n_nodes, ' elems ', model%n_elems, ' bc ', model%n_bc, ' faces ', &
       model%n_face, ' cloads ', model%n_cload, ' niter ', sstate%niter, ' rnorm ', &
       sstate%rnorm, ' solve_ms ', tm%solve_ms
  rc = ufc_ap_free(model, st)
  rc = ufc_ctx_destroy(ctx, st)
  IF (sstate%converged /= 1) CALL EXIT(1)

CONTAINS

  SUBROUTINE check(what)
    CHARACTER(LEN=*), INTENT(IN) :: what
    IF (rc /= UFC_STATUS_OK) THEN
      WRITE(*, '(A,A,I0,A,A)') what, ' failed, status ', rc, ': ', TRIM(ufc_status_message(st))
      CALL EXIT(2)
    END IF
  END SUBROUTINE check

END PROGRAM ufc_inp_step
