! ref_golden.f90 -- golden-vector generator built against the reference's own
! compilable sparse module (NM_Mtx_Sparse.f90, compiled from the reference
! tree by oracle/ref.mk).  TEST INFRASTRUCTURE ONLY: run in the build
! container to write tests/golden fixtures; never shipped or linked into the
! HIP library.
!
! For a structured nx x ny x nz C3D8 box it:
!   1. numbers equations eq = 3(node-1)+d (RT_Asm_DofMap.f90:199-205),
!   2. gives every element an integer-valued 24x24 matrix
!      Ke(i,j) = 1 + MOD(7*e + 3*i + j, 11) (positive, exact in FP64, so
!      NM_CSR_OptimizeStorage drops nothing -- NM_Mtx_Sparse.f90:381-433),
!   3. assembles with NM_CSR_AssembleFromElements (NM_Mtx_Sparse.f90:1044-1091),
!   4. applies NM_CSR_MatVec_Optimized (NM_Mtx_Sparse.f90:341-365) to
!      x(i) = 1 + MOD(5*i, 13) (integers: the product is exact),
!   5. transposes the (non-symmetric) matrix with NM_CSR_Transpose
!      (NM_Mtx_SpMVCSRTranspose.f90:46-145) and applies the transposed rows to
!      the same x: yt = A^T x,
! and writes n, nnz, ia, ja, a, y and yt as text to the file named on the
! command line.  Usage: ref_golden nx ny nz out.txt
PROGRAM ref_golden
  USE IF_Prec_Core, ONLY: wp, i4
  USE IF_Err_Brg, ONLY: ErrorStatusType
  USE NM_Mtx_Sparse, ONLY: NM_CSR_Type, NM_CSR_AssembleFromElements, NM_CSR_MatVec_Optimized
  USE NM_Mtx_SpMVCSRTranspose, ONLY: NM_SparseMatrix_CSR, NM_CSR_Transpose
  IMPLICIT NONE
  INTEGER(i4) :: nx, ny, nz, nn, ne, ndof, i, j, k, e, a, d, n0, sx, sy, u
  INTEGER(i4) :: nodes(8)
  CHARACTER(len=256) :: arg, outfile
  REAL(wp), ALLOCATABLE :: elem_K(:,:,:), x(:), y(:), yt(:)
  INTEGER(i4), ALLOCATABLE :: elem_dofs(:,:)
  TYPE(NM_CSR_Type) :: Kc
  TYPE(NM_SparseMatrix_CSR) :: As, At
  TYPE(ErrorStatusType) :: st

  CALL GET_COMMAND_ARGUMENT(1, arg); READ(arg, *) nx
  CALL GET_COMMAND_ARGUMENT(2, arg); READ(arg, *) ny
  CALL GET_COMMAND_ARGUMENT(3, arg); READ(arg, *) nz
  CALL GET_COMMAND_ARGUMENT(4, outfile)

  sx = nx + 1
  sy = ny + 1
  nn = sx * sy * (nz + 1)
  ne = nx * ny * nz
  ndof = 3 * nn
  ALLOCATE(elem_K(24, 24, ne), elem_dofs(ne, 24))

  DO k = 0, nz - 1
    DO j = 0, ny - 1
      DO i = 0, nx - 1
        e = 1 + i + nx * (j + ny * k)
        n0 = 1 + i + sx * (j + sy * k)
        nodes = [n0, n0 + 1, n0 + 1 + sx, n0 + sx, &
                 n0 + sx * sy, n0 + 1 + sx * sy, n0 + 1 + sx + sx * sy, n0 + sx + sx * sy]
        DO a = 1, 8
          DO d = 1, 3
            elem_dofs(e, 3 * (a - 1) + d) = 3 * (nodes(a) - 1) + d
          END DO
        END DO
        DO a = 1, 24
          DO d = 1, 24
            elem_K(a, d, e) = REAL(1 + MOD(7 * e + 3 * a + d, 11), wp)
          END DO
        END DO
      END DO
    END DO
  END DO

  CALL NM_CSR_AssembleFromElements(Kc, elem_K, elem_dofs, ndof, 64, st)

  ALLOCATE(x(ndof), y(ndof))
  DO i = 1, ndof
    x(i) = REAL(1 + MOD(5 * i, 13), wp)
  END DO
  CALL NM_CSR_MatVec_Optimized(Kc, x, y, st)

  As%nrows = Kc%n
  As%ncols = Kc%n
  As%nnz = Kc%nnz
  ALLOCATE(As%row_ptr(Kc%n + 1), As%col_ind(Kc%nnz), As%values(Kc%nnz))
  As%row_ptr = Kc%ia(1:Kc%n + 1)
  As%col_ind = Kc%ja(1:Kc%nnz)
  As%values = Kc%a(1:Kc%nnz)
  CALL NM_CSR_Transpose(As, At, st)
  ALLOCATE(yt(ndof))
  DO i = 1, At%nrows
    yt(i) = 0.0_wp
    DO k = At%row_ptr(i), At%row_ptr(i + 1) - 1
      yt(i) = yt(i) + At%values(k) * x(At%col_ind(k))
    END DO
  END DO

  OPEN(NEWUNIT=u, FILE=TRIM(outfile), STATUS='REPLACE', ACTION='WRITE')
  WRITE(u, '(I0,1X,I0)') Kc%n, Kc%nnz
  DO i = 1, Kc%n + 1
    WRITE(u, '(I0)') Kc%ia(i)
  END DO
  DO i = 1, Kc%nnz
    WRITE(u, '(I0,1X,F0.1)') Kc%ja(i), Kc%a(i)
  END DO
  DO i = 1, Kc%n
    WRITE(u, '(F0.1)') y(i)
  END DO
  DO i = 1, Kc%n
    WRITE(u, '(F0.1)') yt(i)
  END DO
  CLOSE(u)
END PROGRAM ref_golden
