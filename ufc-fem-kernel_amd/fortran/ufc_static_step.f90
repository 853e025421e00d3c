! ufc_static_step.f90 -- thin Fortran static-step driver over libufc_hip.so.
!
! The static sequence of RT_StepDriver_Execute (L5_RT/StepDriver/RT_Step_Exec.f90:267-676,
! paths relative to the reference root) with the hot path delegated through
! ISO_C_BINDING (ufc_hip_iface):
!   DOF numbering        eq = 3(node-1)+d          RT_Asm_DofMap.f90:199-205
!   sparse pattern       merged, column-sorted CSR RT_Solv_Sparse.f90:336-447
!   global stiffness     RT_Asm_GlobalStiffness    RT_Asm_Solv.f90:1762-1980
!   global load          RT_Asm_GlobalLoad         RT_Asm_Solv.f90:1443-1760
!   boundary conditions  RT_Asm_ApplyBC (penalty)  RT_Asm_Solv.f90:318-344
!   linear solve         NM_Solver_PCG             NM_Solv_Core.f90:363-423
! The mesh (Fortran-owned arrays, 1-based) is a structured C3D8 box.
!
! Usage: ufc_static_step <problem> <nx> <ny> <nz> <precision> <u_out>
!   problem   cantilever : box 10 x 2 x 2, E = 210e9, nu = 0.3, x = 0 clamped,
!                          uniform traction on x = 10 summing to -1e3 in z
!             cube       : unit cube, E = 1e9, nu = 0.3, z = 0 clamped,
!                          pressure 1e6 on z = 1
!   precision fp64 | mixed
!   u_out     displacements, n_dof doubles, unformatted stream
! Prints one summary line; exits with the C-ABI status of the solve.
PROGRAM ufc_static_step
  USE, INTRINSIC :: ISO_C_BINDING
  USE ufc_hip_iface
  IMPLICIT NONE

  CHARACTER(LEN=64) :: arg, problem, prec_name
  CHARACTER(LEN=512) :: outfile
  INTEGER :: nx, ny, nz, sx, sy, i, j, k, e, n0, a, nf, nb, u_unit
  INTEGER(C_INT32_T) :: nn, ne, prec, rc
  REAL(C_DOUBLE) :: lx, ly, lz, young, nu
  REAL(C_DOUBLE), ALLOCATABLE, TARGET :: coords(:, :), e_field(:), u(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: conn(:, :), bc_eq(:), f_elem(:), f_face(:)
  INTEGER(C_INT32_T), ALLOCATABLE, TARGET :: row_ptr(:), col_idx(:)
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
  INTEGER(C_INT64_T) :: colsum

  CALL GET_COMMAND_ARGUMENT(1, problem)
  CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg, *) nx
  CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg, *) ny
  CALL GET_COMMAND_ARGUMENT(4, arg); READ(arg, *) nz
  CALL GET_COMMAND_ARGUMENT(5, prec_name)
  CALL GET_COMMAND_ARGUMENT(6, outfile)

  IF (TRIM(problem) == 'cantilever') THEN
    lx = 10.0_C_DOUBLE; ly = 2.0_C_DOUBLE; lz = 2.0_C_DOUBLE; young = 210.0E9_C_DOUBLE
  ELSE
    lx = 1.0_C_DOUBLE; ly = 1.0_C_DOUBLE; lz = 1.0_C_DOUBLE; young = 1.0E9_C_DOUBLE
  END IF
  nu = 0.3_C_DOUBLE
  prec = UFC_PREC_FP64
  IF (TRIM(prec_name) == 'mixed') prec = UFC_PREC_MIXED

  ! ---- mesh: node id 1 + i + (nx+1)(j + (ny+1)k), Abaqus C3D8 node order
  sx = nx + 1
  sy = ny + 1
  nn = INT(sx * sy * (nz + 1), C_INT32_T)
  ne = INT(nx * ny * nz, C_INT32_T)
  ALLOCATE(coords(3, nn), conn(8, ne), e_field(ne), u(3 * nn))
  DO k = 0, nz
    DO j = 0, ny
      DO i = 0, nx
        n0 = 1 + i + sx * (j + sy * k)
        coords(1, n0) = lx * REAL(i, C_DOUBLE) / REAL(nx, C_DOUBLE)
        coords(2, n0) = ly * REAL(j, C_DOUBLE) / REAL(ny, C_DOUBLE)
        coords(3, n0) = lz * REAL(k, C_DOUBLE) / REAL(nz, C_DOUBLE)
      END DO
    END DO
  END DO
  DO k = 0, nz - 1
    DO j = 0, ny - 1
      DO i = 0, nx - 1
        e = 1 + i + nx * (j + ny * k)
        n0 = 1 + i + sx * (j + sy * k)
        conn(:, e) = [n0, n0 + 1, n0 + 1 + sx, n0 + sx, n0 + sx * sy, n0 + 1 + sx * sy, &
                      n0 + 1 + sx + sx * sy, n0 + sx + sx * sy]
      END DO
    END DO
  END DO
  e_field = young

  ! ---- boundary conditions and loaded faces (1-based equations and faces)
  IF (TRIM(problem) == 'cantilever') THEN
    nb = 3 * sy * (nz + 1)
    nf = ny * nz
  ELSE
    nb = 3 * sx * sy
    nf = nx * ny
  END IF
  ALLOCATE(bc_eq(nb), f_elem(nf), f_face(nf))
  nb = 0
  DO n0 = 1, nn
    IF ((TRIM(problem) == 'cantilever' .AND. coords(1, n0) == 0.0_C_DOUBLE) .OR. &
        (TRIM(problem) /= 'cantilever' .AND. coords(3, n0) == 0.0_C_DOUBLE)) THEN
      DO a = 1, 3
        nb = nb + 1
        bc_eq(nb) = INT(3 * (n0 - 1) + a, C_INT32_T)
      END DO
    END IF
  END DO
  nf = 0
  DO k = 0, nz - 1
    DO j = 0, ny - 1
      DO i = 0, nx - 1
        e = 1 + i + nx * (j + ny * k)
        IF (TRIM(problem) == 'cantilever' .AND. i == nx - 1) THEN
          nf = nf + 1; f_elem(nf) = e; f_face(nf) = 6        ! xi = +1 face
        ELSE IF (TRIM(problem) /= 'cantilever' .AND. k == nz - 1) THEN
          nf = nf + 1; f_elem(nf) = e; f_face(nf) = 2        ! zeta = +1 face
        END IF
      END DO
    END DO
  END DO

  ! ---- drive the C-ABI
  rc = ufc_ctx_create(0_C_INT32_T, ctx, st)
  CALL check('ufc_ctx_create')
  mesh%n_nodes = nn
  mesh%n_elems = ne
  mesh%index_base = 1
  mesh%coords = C_LOC(coords)
  mesh%conn = C_LOC(conn)
  rc = ufc_rt_asm_build_pattern(ctx, mesh, info, st)
  CALL check('ufc_rt_asm_build_pattern')

  aalgo%precision = prec
  aalgo%material = UFC_MAT_ELASTIC
  aargs%young = C_LOC(e_field)
  aargs%nu = nu
  rc = ufc_rt_asm_assemble_k(ctx, aalgo, aargs, st)
  CALL check('ufc_rt_asm_assemble_k')

  ld%n_faces = INT(nf, C_INT32_T)
  ld%index_base = 1
  ld%accumulate = 0
  ld%elem = C_LOC(f_elem)
  ld%face = C_LOC(f_face)
  IF (TRIM(problem) == 'cantilever') THEN
    ld%kind = UFC_LOAD_TRACTION
    ld%traction = [0.0_C_DOUBLE, 0.0_C_DOUBLE, -1.0E3_C_DOUBLE / (ly * lz)]
  ELSE
    ld%kind = UFC_LOAD_PRESSURE
    ld%magnitude = 1.0E6_C_DOUBLE
  END IF
  rc = ufc_rt_asm_load(ctx, ld, st)
  CALL check('ufc_rt_asm_load')

  bc%n_bc = INT(nb, C_INT32_T)
  bc%index_base = 1
  bc%eq = C_LOC(bc_eq)
  bc%value = C_NULL_PTR
  bc%penalty = 1.0E12_C_DOUBLE
  rc = ufc_rt_asm_apply_bc(ctx, bc, st)
  CALL check('ufc_rt_asm_apply_bc')

  salgo%precision = prec
  salgo%rtol = 1.0E-10_C_DOUBLE
  salgo%inner_rtol = 1.0E-5_C_DOUBLE
  sstate%x = C_LOC(u)
  sstate%b = C_NULL_PTR
  sstate%x0_zero = 1
  rc = ufc_nm_pcg(ctx, salgo, sstate, st)
  IF (rc /= UFC_STATUS_OK) THEN
    WRITE(*, '(A,I0,A,A)') 'ufc_nm_pcg status ', rc, ': ', TRIM(ufc_status_message(st))
  END IF
  rc = ufc_ctx_timing(ctx, tm, st)

  ! exported pattern in the caller's base (RT_CSRMatrix rowPtr/colInd)
  ALLOCATE(row_ptr(info%n_dof + 1), col_idx(info%nnz))
  rc = ufc_rt_asm_get_pattern(ctx, 1_C_INT32_T, C_LOC(row_ptr), C_LOC(col_idx), st)
  CALL check('ufc_rt_asm_get_pattern')
  colsum = SUM(INT(col_idx, C_INT64_T))

  OPEN(NEWUNIT=u_unit, FILE=TRIM(outfile), ACCESS='STREAM', FORM='UNFORMATTED', &
       STATUS='REPLACE', ACTION='WRITE')
  WRITE(u_unit) u
  CLOSE(u_unit)
  WRITE(*, '(A,I0,A,I0,A,I0,A,I0,A,ES12.5,A,I0,A,I0,A,F10.3)') 'ufc_static_step n_dof ', &
       info%n_dof, ' nnz ', info%nnz, ' row_ptr_end ', row_ptr(info%n_dof + 1), &
       ' niter ', sstate%niter, ' rnorm ', sstate%rnorm, ' n_outer ', sstate%n_outer, &
       ' colsum ', colsum, ' solve_ms ', tm%solve_ms
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

END PROGRAM ufc_static_step
