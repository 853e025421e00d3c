! ref_cantilever.f90 -- second-opinion CPU timing of the reference's small
! configuration: the 20x4x4 C3D8 cantilever on a 10x2x2 box (E = 210e9,
! nu = 0.3, x = 0 clamped, tip traction summing to -1e3 in z; 1575 DOF),
! assembled and solved with Jacobi-PCG to 1e-10.  TEST INFRASTRUCTURE ONLY:
! a CPU timing and cross-check, never linked into the HIP library.
!
! Built by oracle/ref.mk against the reference's own compilable sparse module
! (NM_Mtx_Sparse.f90, compiled where it lies in the reference tree), which
! supplies the two sparse steps:
!   - CSR assembly: NM_COO_Init, NM_COO_AddElementMatrix per element and
!     NM_COO_to_CSR (NM_Mtx_Sparse.f90:870-996, 164-222), the steps of
!     NM_CSR_AssembleFromElements (:1044-1091) without its final
!     NM_CSR_OptimizeStorage (:381-433): that pass drops |a| <= 1e-15 and, when
!     it drops any entry (an assembled C3D8 K has exact zeros: 1270 on this
!     mesh), rewrites ia(i) before reading row i from it, so it re-reads the
!     previous row's tail and writes past its new arrays (a heap corruption
!     here).  The zeros kept cost the SpMV a few percent.
!   - SpMV inside the PCG: NM_CSR_MatVec_Optimized (NM_Mtx_Sparse.f90:341-365).
! The reference modules of the other steps do not compile with this flang, so
! they are restated here:
!   - element matrices, 2x2x2 Gauss B^T D B (PH_Elem_C3D8_StiffMatrix,
!     PH_Elem_C3D8.f90:759-780; trilinear 1/8 shape functions, J^-T),
!   - traction lumped area/4 onto the face nodes (RT_Asm_GlobalLoad),
!   - diagonal penalty 1e12 (RT_Asm_ApplyBC, RT_Asm_Solv.f90:318-344),
!   - Jacobi-PCG from x = 0 with ||r||/||b|| < rtol (NM_Solver_PCG,
!     NM_Solv_Core.f90:363-423).
!
! Usage: ref_cantilever repeats out.txt
! Times `repeats` runs of assembly (element matrices + CSR) and of the solve
! (SYSTEM_CLOCK wall time) and writes to out.txt:
!   ndof nnz iterations repeats
!   the assembly seconds of each repeat, the solve seconds of each repeat
!   u (ndof values, 0-based equation order eq = 3 node + d)
PROGRAM ref_cantilever
  USE IF_Prec_Core, ONLY: wp, i4
  USE IF_Err_Brg, ONLY: ErrorStatusType
  USE NM_Mtx_Sparse, ONLY: NM_CSR_Type, NM_COO_Type, NM_COO_Init, NM_COO_AddElementMatrix, &
                           NM_COO_to_CSR, NM_CSR_MatVec_Optimized
  IMPLICIT NONE
  INTEGER(i4), PARAMETER :: nx = 20, ny = 4, nz = 4
  REAL(wp), PARAMETER :: lx = 10.0_wp, ly = 2.0_wp, lz = 2.0_wp
  REAL(wp), PARAMETER :: young = 210.0e9_wp, nu = 0.3_wp, total = -1.0e3_wp
  REAL(wp), PARAMETER :: penalty = 1.0e12_wp, rtol = 1.0e-10_wp
  INTEGER(i4) :: sx, sy, nn, ne, ndof, i, j, k, e, a, d, rep, repeats, iters, u_out
  INTEGER(i4), ALLOCATABLE :: conn(:,:), elem_dofs(:,:), bc(:)
  REAL(wp), ALLOCATABLE :: xyz(:,:), elem_K(:,:,:), f(:), u(:), t_asm(:), t_sol(:)
  INTEGER(i4) :: nbc
  CHARACTER(len=256) :: arg, outfile
  TYPE(NM_CSR_Type) :: Kc
  TYPE(NM_COO_Type) :: Kcoo
  TYPE(ErrorStatusType) :: st
  INTEGER(8) :: c0, c1, rate

  CALL GET_COMMAND_ARGUMENT(1, arg); READ(arg, *) repeats
  CALL GET_COMMAND_ARGUMENT(2, outfile)
  sx = nx + 1
  sy = ny + 1
  nn = sx * sy * (nz + 1)
  ne = nx * ny * nz
  ndof = 3 * nn
  ALLOCATE(xyz(3, nn), conn(8, ne), elem_K(24, 24, ne), elem_dofs(ne, 24), f(ndof), u(ndof))
  ALLOCATE(t_asm(repeats), t_sol(repeats), bc(3 * sy * (nz + 1)))
  DO k = 0, nz
    DO j = 0, ny
      DO i = 0, nx
        xyz(:, 1 + i + sx * (j + sy * k)) = [lx * i / nx, ly * j / ny, lz * k / nz]
      END DO
    END DO
  END DO
  DO k = 0, nz - 1
    DO j = 0, ny - 1
      DO i = 0, nx - 1
        e = 1 + i + nx * (j + ny * k)
        a = 1 + i + sx * (j + sy * k)
        conn(:, e) = [a, a + 1, a + 1 + sx, a + sx, a + sx * sy, a + 1 + sx * sy, &
                      a + 1 + sx + sx * sy, a + sx + sx * sy]
        DO a = 1, 8
          DO d = 1, 3
            elem_dofs(e, 3 * (a - 1) + d) = 3 * (conn(a, e) - 1) + d
          END DO
        END DO
      END DO
    END DO
  END DO
  ! x = 0 clamped; traction t = total / (ly lz) in z on the x = lx faces,
  ! area / 4 of every element face to each of its four nodes
  nbc = 0
  DO k = 0, nz
    DO j = 0, ny
      DO d = 1, 3
        nbc = nbc + 1
        bc(nbc) = 3 * (sx * (j + sy * k)) + d
      END DO
    END DO
  END DO
  f = 0.0_wp
  DO k = 0, nz - 1
    DO j = 0, ny - 1
      e = 1 + (nx - 1) + nx * (j + ny * k)
      DO a = 2, 7, 1
        IF (a == 2 .OR. a == 3 .OR. a == 6 .OR. a == 7) THEN
          f(3 * (conn(a, e) - 1) + 3) = f(3 * (conn(a, e) - 1) + 3) + &
              (total / (ly * lz)) * (ly / ny) * (lz / nz) / 4.0_wp
        END IF
      END DO
    END DO
  END DO
  DO i = 1, nbc
    f(bc(i)) = f(bc(i)) + penalty * 0.0_wp
  END DO

  DO rep = 1, repeats
    CALL SYSTEM_CLOCK(c0, rate)
    DO e = 1, ne
      CALL c3d8_ke(xyz(:, conn(:, e)), young, nu, elem_K(:, :, e))
    END DO
    CALL NM_COO_Init(Kcoo, ndof, ndof, 64, st)
    DO e = 1, ne
      CALL NM_COO_AddElementMatrix(Kcoo, elem_K(:, :, e), elem_dofs(e, :), st)
    END DO
    CALL NM_COO_to_CSR(Kcoo, Kc, st)
    CALL add_penalty(Kc, bc(1:nbc), penalty)
    CALL SYSTEM_CLOCK(c1)
    t_asm(rep) = REAL(c1 - c0, wp) / REAL(rate, wp)
    CALL SYSTEM_CLOCK(c0)
    CALL jacobi_pcg(Kc, f, u, rtol, 100000, iters)
    CALL SYSTEM_CLOCK(c1)
    t_sol(rep) = REAL(c1 - c0, wp) / REAL(rate, wp)
  END DO

  OPEN(NEWUNIT=u_out, FILE=TRIM(outfile), STATUS='REPLACE', ACTION='WRITE')
  WRITE(u_out, '(4I12)') ndof, Kc%nnz, iters, repeats
  WRITE(u_out, '(ES25.16)') t_asm
  WRITE(u_out, '(ES25.16)') t_sol
  WRITE(u_out, '(ES25.16)') u
  CLOSE(u_out)

CONTAINS

  ! Ke (24 x 24, dof 3(a-1)+d) of a C3D8 with node coordinates x(3, 8):
  ! sum over the 2x2x2 Gauss points (weights 1) of detJ B^T D B
  SUBROUTINE c3d8_ke(x, e_mod, pr, ke)
    REAL(wp), INTENT(IN) :: x(3, 8), e_mod, pr
    REAL(wp), INTENT(OUT) :: ke(24, 24)
    INTEGER(i4), PARAMETER :: xs(8) = [-1, 1, 1, -1, -1, 1, 1, -1]
    INTEGER(i4), PARAMETER :: ys(8) = [-1, -1, 1, 1, -1, -1, 1, 1]
    INTEGER(i4), PARAMETER :: zs(8) = [-1, -1, -1, -1, 1, 1, 1, 1]
    REAL(wp) :: g, xi, et, ze, dn(3, 8), jac(3, 3), ji(3, 3), det, gx(3, 8)
    REAL(wp) :: b(6, 24), dm(6, 6), lam, mu
    INTEGER(i4) :: p, q, r, n
    g = 1.0_wp / SQRT(3.0_wp)
    lam = e_mod * pr / ((1.0_wp + pr) * (1.0_wp - 2.0_wp * pr))
    mu = e_mod / (2.0_wp * (1.0_wp + pr))
    dm = 0.0_wp
    dm(1:3, 1:3) = lam
    DO n = 1, 3
      dm(n, n) = lam + 2.0_wp * mu
      dm(n + 3, n + 3) = mu
    END DO
    ke = 0.0_wp
    DO p = 1, 2
      DO q = 1, 2
        DO r = 1, 2
          xi = MERGE(-g, g, p == 1)
          et = MERGE(-g, g, q == 1)
          ze = MERGE(-g, g, r == 1)
          DO n = 1, 8
            dn(1, n) = 0.125_wp * xs(n) * (1.0_wp + ys(n) * et) * (1.0_wp + zs(n) * ze)
            dn(2, n) = 0.125_wp * (1.0_wp + xs(n) * xi) * ys(n) * (1.0_wp + zs(n) * ze)
            dn(3, n) = 0.125_wp * (1.0_wp + xs(n) * xi) * (1.0_wp + ys(n) * et) * zs(n)
          END DO
          jac = MATMUL(dn, TRANSPOSE(x))      ! jac(k, j) = d x_j / d xi_k
          det = jac(1,1) * (jac(2,2) * jac(3,3) - jac(2,3) * jac(3,2)) &
              - jac(1,2) * (jac(2,1) * jac(3,3) - jac(2,3) * jac(3,1)) &
              + jac(1,3) * (jac(2,1) * jac(3,2) - jac(2,2) * jac(3,1))
          IF (ABS(det) <= 1.0e-12_wp) CYCLE
          ji(1,1) = (jac(2,2) * jac(3,3) - jac(2,3) * jac(3,2)) / det
          ji(1,2) = -(jac(1,2) * jac(3,3) - jac(1,3) * jac(3,2)) / det
          ji(1,3) = (jac(1,2) * jac(2,3) - jac(1,3) * jac(2,2)) / det
          ji(2,1) = -(jac(2,1) * jac(3,3) - jac(2,3) * jac(3,1)) / det
          ji(2,2) = (jac(1,1) * jac(3,3) - jac(1,3) * jac(3,1)) / det
          ji(2,3) = -(jac(1,1) * jac(2,3) - jac(1,3) * jac(2,1)) / det
          ji(3,1) = (jac(2,1) * jac(3,2) - jac(2,2) * jac(3,1)) / det
          ji(3,2) = -(jac(1,1) * jac(3,2) - jac(1,2) * jac(3,1)) / det
          ji(3,3) = (jac(1,1) * jac(2,2) - jac(1,2) * jac(2,1)) / det
          gx = MATMUL(ji, dn)                 ! gx(j, n) = dN_n / dx_j
          b = 0.0_wp
          DO n = 1, 8
            b(1, 3 * n - 2) = gx(1, n)
            b(2, 3 * n - 1) = gx(2, n)
            b(3, 3 * n) = gx(3, n)
            b(4, 3 * n - 2) = gx(2, n)
            b(4, 3 * n - 1) = gx(1, n)
            b(5, 3 * n - 1) = gx(3, n)
            b(5, 3 * n) = gx(2, n)
            b(6, 3 * n - 2) = gx(3, n)
            b(6, 3 * n) = gx(1, n)
          END DO
          ke = ke + det * MATMUL(TRANSPOSE(b), MATMUL(dm, b))
        END DO
      END DO
    END DO
  END SUBROUTINE c3d8_ke

  SUBROUTINE add_penalty(kc_, eqs, pval)
    TYPE(NM_CSR_Type), INTENT(INOUT) :: kc_
    INTEGER(i4), INTENT(IN) :: eqs(:)
    REAL(wp), INTENT(IN) :: pval
    INTEGER(i4) :: ii, q
    DO ii = 1, SIZE(eqs)
      DO q = kc_%ia(eqs(ii)), kc_%ia(eqs(ii) + 1) - 1
        IF (kc_%ja(q) == eqs(ii)) kc_%a(q) = kc_%a(q) + pval
      END DO
    END DO
  END SUBROUTINE add_penalty

  SUBROUTINE jacobi_pcg(kc_, bvec, x, tol, max_iter, niter)
    TYPE(NM_CSR_Type), INTENT(IN) :: kc_
    REAL(wp), INTENT(IN) :: bvec(:), tol
    REAL(wp), INTENT(OUT) :: x(:)
    INTEGER(i4), INTENT(IN) :: max_iter
    INTEGER(i4), INTENT(OUT) :: niter
    REAL(wp), ALLOCATABLE :: r(:), z(:), pv(:), ap(:), dg(:)
    REAL(wp) :: rz, rz_new, alpha, beta, bn, rn
    INTEGER(i4) :: ii, q, it
    TYPE(ErrorStatusType) :: st_
    ALLOCATE(r(SIZE(bvec)), z(SIZE(bvec)), pv(SIZE(bvec)), ap(SIZE(bvec)), dg(SIZE(bvec)))
    DO ii = 1, SIZE(bvec)
      dg(ii) = 0.0_wp
      DO q = kc_%ia(ii), kc_%ia(ii + 1) - 1
        IF (kc_%ja(q) == ii) dg(ii) = kc_%a(q)
      END DO
    END DO
    x = 0.0_wp
    r = bvec
    z = r / dg
    pv = z
    rz = DOT_PRODUCT(r, z)
    bn = SQRT(DOT_PRODUCT(bvec, bvec))
    niter = max_iter
    DO it = 1, max_iter
      CALL NM_CSR_MatVec_Optimized(kc_, pv, ap, st_)
      alpha = rz / DOT_PRODUCT(pv, ap)
      x = x + alpha * pv
      r = r - alpha * ap
      rn = SQRT(DOT_PRODUCT(r, r)) / bn
      IF (rn < tol) THEN
        niter = it
        EXIT
      END IF
      z = r / dg
      rz_new = DOT_PRODUCT(r, z)
      beta = rz_new / rz
      rz = rz_new
      pv = z + beta * pv
    END DO
  END SUBROUTINE jacobi_pcg

END PROGRAM ref_cantilever
