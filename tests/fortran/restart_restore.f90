! restart_restore.f90 -- reads a restart file with the READ sequence of
! RT_Out_RestartRestore (L5_RT/Output/RT_Out_Restart.f90:45-187): the file
! opened as IF_FileHandle opens IF_IO_FORMAT_BINARY files (FORM='UNFORMATTED',
! ACCESS='SEQUENTIAL', L1_IF/IO/IF_IO_File.f90:219-221) and one READ of
! (n_bytes + 3) / 4 INTEGER(4) words per ReadBinary call (IF_IO_File.f90:
! ReadBinary).  Test program: prints what it read, one value per line.
!
! Usage: restart_restore <file>
PROGRAM restart_restore
  IMPLICIT NONE
  INTEGER, PARAMETER :: i4 = SELECTED_INT_KIND(9), i8 = SELECTED_INT_KIND(18)
  INTEGER, PARAMETER :: wp = SELECTED_REAL_KIND(15, 307)
  CHARACTER(LEN=512) :: fname
  INTEGER :: unit, ios, k
  INTEGER(i4) :: hdr(3), sw(5), cw(4), checksum(1)
  INTEGER(i4), ALLOCATABLE :: words(:)
  REAL(wp), ALLOCATABLE :: u(:), v(:), a(:)
  REAL(wp) :: time, lambda, rnorm
  INTEGER(i4) :: increment, iterations, converged

  CALL GET_COMMAND_ARGUMENT(1, fname)
  OPEN(NEWUNIT=unit, FILE=TRIM(fname), ACTION='READ', FORM='UNFORMATTED', &
       ACCESS='SEQUENTIAL', STATUS='OLD', IOSTAT=ios)
  IF (ios /= 0) STOP 2
  ! header: 12 bytes
  READ(unit, IOSTAT=ios) hdr(1:3)
  IF (ios /= 0) STOP 3
  PRINT '(A,I0)', 'version ', hdr(1)
  PRINT '(A,I0)', 'stamp ', hdr(2)
  PRINT '(A,I0)', 'n ', hdr(3)
  ALLOCATE(words(2 * hdr(3)), u(hdr(3)), v(hdr(3)), a(hdr(3)))
  ! u, v, a: 8 n bytes each
  READ(unit, IOSTAT=ios) words
  IF (ios /= 0) STOP 4
  u = TRANSFER(words, u)
  READ(unit, IOSTAT=ios) words
  IF (ios /= 0) STOP 5
  v = TRANSFER(words, v)
  READ(unit, IOSTAT=ios) words
  IF (ios /= 0) STOP 6
  a = TRANSFER(words, a)
  ! step state: 20 bytes = 5 words: time (2), increment (1), lambda (2)
  READ(unit, IOSTAT=ios) sw
  IF (ios /= 0) STOP 7
  time = TRANSFER(sw(1:2), time)
  increment = sw(3)
  lambda = TRANSFER(sw(4:5), lambda)
  ! convergence: 16 bytes = 4 words: residual_norm (2), iterations, converged
  READ(unit, IOSTAT=ios) cw
  IF (ios /= 0) STOP 8
  rnorm = TRANSFER(cw(1:2), rnorm)
  iterations = cw(3)
  converged = cw(4)
  READ(unit, IOSTAT=ios) checksum
  IF (ios /= 0) STOP 9
  CLOSE(unit)
  PRINT '(A,ES25.17)', 'time ', time
  PRINT '(A,I0)', 'increment ', increment
  PRINT '(A,ES25.17)', 'lambda ', lambda
  PRINT '(A,ES25.17)', 'residual_norm ', rnorm
  PRINT '(A,I0)', 'iterations ', iterations
  PRINT '(A,I0)', 'converged ', converged
  PRINT '(A,I0)', 'checksum ', checksum(1)
  PRINT '(A,ES25.17)', 'max_abs_va ', MAX(MAXVAL(ABS(v)), MAXVAL(ABS(a)))
  DO k = 1, hdr(3)
    PRINT '(A,ES25.17)', 'u ', u(k)
  END DO
END PROGRAM restart_restore
