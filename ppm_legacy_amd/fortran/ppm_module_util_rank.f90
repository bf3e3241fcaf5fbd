!-------------------------------------------------------------------------------
! ppm_module_util_rank -- ppm_util_rank3d / ppm_util_rank2d (ppm_module_util_rank.f:
! 40-47; arguments as ppm_util_rank3d.f:147-151): the stable cell ranking, run
! on the GPU (counting sort).  lhbx(1:nbox+1) holds the 1-based cell heads and
! lpdx(1:icount) the 1-based particle indices in cell order; both POINTER
! arrays are (re)allocated here as the reference's ppm_alloc fit does.
!-------------------------------------------------------------------------------
MODULE ppm_module_util_rank
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  USE ppmg_fcore
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: ppm_util_rank3d, ppm_util_rank2d
  INTERFACE ppm_util_rank3d
    MODULE PROCEDURE ppm_util_rank3d_s
    MODULE PROCEDURE ppm_util_rank3d_d
  END INTERFACE
  INTERFACE ppm_util_rank2d
    MODULE PROCEDURE ppm_util_rank2d_s
    MODULE PROCEDURE ppm_util_rank2d_d
  END INTERFACE
CONTAINS

  SUBROUTINE ppm_util_rank3d_s(xp, Np, xmin, xmax, Nm, Ngl, lpdx, lhbx, info)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), INTENT(IN) :: xp
    INTEGER, INTENT(IN) :: Np
    REAL(MK), DIMENSION(:), INTENT(IN) :: xmin, xmax
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, Ngl
    INTEGER, DIMENSION(:), POINTER :: lpdx, lhbx
    INTEGER, INTENT(INOUT) :: info
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_util_rank3d'
    INTEGER :: nbox, d, icount, n
    INTEGER, DIMENSION(:), ALLOCATABLE, TARGET :: lp
    CALL rank_check(caller, 3, SHAPE(xp), Np, SIZE(xmin), SIZE(xmax), Nm, Ngl, nbox, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    ALLOCATE(lp(MAX(n, 1)))
    ! lhbx(nbox+1), lpdx(icount) as ppm_util_rank3d.f (ppm_alloc fit)
    IF (ASSOCIATED(lhbx)) DEALLOCATE(lhbx)
    ALLOCATE(lhbx(nbox + 1))
    CALL rank_core3_s(xp(1:3, 1:n), n, xmin(1:3), xmax(1:3), Nm, Ngl, lp, lhbx, icount, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(lpdx)) DEALLOCATE(lpdx)
    ALLOCATE(lpdx(icount))
    DO d = 1, icount
      lpdx(d) = lp(d)
    END DO
  END SUBROUTINE ppm_util_rank3d_s

  SUBROUTINE rank_core3_s(xp, n, xmin, xmax, Nm, Ngl, lp, lh, icount, info)
    REAL(C_FLOAT), TARGET, INTENT(IN) :: xp(*), xmin(*), xmax(*)
    INTEGER, INTENT(IN) :: n
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, Ngl
    INTEGER(C_INT), TARGET :: lp(*), lh(*)
    INTEGER, INTENT(OUT) :: icount, info
    INTEGER(C_INT) :: cnm(3), cngl(6), ic
    INTEGER :: d
    cnm = 1
    cngl = 0
    DO d = 1, SIZE(Nm)
      cnm(d) = Nm(d)
    END DO
    DO d = 1, SIZE(Ngl)
      cngl(d) = Ngl(d)
    END DO
    info = ppmg_util_rank_reg(INT(3, C_INT), ploc_s(xp, n), INT(n, C_INT), C_LOC(xmin(1)), C_LOC(xmax(1)), cnm, &
                              cngl, C_LOC(lp(1)), C_LOC(lh(1)), ic, ppmg_prec_f32, ppmg_mem_host)
    icount = ic
    CALL ppmg_report('ppm_util_rank3d', info)
  END SUBROUTINE rank_core3_s

  SUBROUTINE ppm_util_rank3d_d(xp, Np, xmin, xmax, Nm, Ngl, lpdx, lhbx, info)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), INTENT(IN) :: xp
    INTEGER, INTENT(IN) :: Np
    REAL(MK), DIMENSION(:), INTENT(IN) :: xmin, xmax
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, Ngl
    INTEGER, DIMENSION(:), POINTER :: lpdx, lhbx
    INTEGER, INTENT(INOUT) :: info
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_util_rank3d'
    INTEGER :: nbox, d, icount, n
    INTEGER, DIMENSION(:), ALLOCATABLE, TARGET :: lp
    CALL rank_check(caller, 3, SHAPE(xp), Np, SIZE(xmin), SIZE(xmax), Nm, Ngl, nbox, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    ALLOCATE(lp(MAX(n, 1)))
    ! lhbx(nbox+1), lpdx(icount) as ppm_util_rank3d.f (ppm_alloc fit)
    IF (ASSOCIATED(lhbx)) DEALLOCATE(lhbx)
    ALLOCATE(lhbx(nbox + 1))
    CALL rank_core3_d(xp(1:3, 1:n), n, xmin(1:3), xmax(1:3), Nm, Ngl, lp, lhbx, icount, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(lpdx)) DEALLOCATE(lpdx)
    ALLOCATE(lpdx(icount))
    DO d = 1, icount
      lpdx(d) = lp(d)
    END DO
  END SUBROUTINE ppm_util_rank3d_d

  SUBROUTINE rank_core3_d(xp, n, xmin, xmax, Nm, Ngl, lp, lh, icount, info)
    REAL(C_DOUBLE), TARGET, INTENT(IN) :: xp(*), xmin(*), xmax(*)
    INTEGER, INTENT(IN) :: n
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, Ngl
    INTEGER(C_INT), TARGET :: lp(*), lh(*)
    INTEGER, INTENT(OUT) :: icount, info
    INTEGER(C_INT) :: cnm(3), cngl(6), ic
    INTEGER :: d
    cnm = 1
    cngl = 0
    DO d = 1, SIZE(Nm)
      cnm(d) = Nm(d)
    END DO
    DO d = 1, SIZE(Ngl)
      cngl(d) = Ngl(d)
    END DO
    info = ppmg_util_rank_reg(INT(3, C_INT), ploc_d(xp, n), INT(n, C_INT), C_LOC(xmin(1)), C_LOC(xmax(1)), cnm, &
                              cngl, C_LOC(lp(1)), C_LOC(lh(1)), ic, ppmg_prec_f64, ppmg_mem_host)
    icount = ic
    CALL ppmg_report('ppm_util_rank3d', info)
  END SUBROUTINE rank_core3_d

  SUBROUTINE ppm_util_rank2d_s(xp, Np, xmin, xmax, Nm, Ngl, lpdx, lhbx, info)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), INTENT(IN) :: xp
    INTEGER, INTENT(IN) :: Np
    REAL(MK), DIMENSION(:), INTENT(IN) :: xmin, xmax
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, Ngl
    INTEGER, DIMENSION(:), POINTER :: lpdx, lhbx
    INTEGER, INTENT(INOUT) :: info
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_util_rank2d'
    INTEGER :: nbox, d, icount, n
    INTEGER, DIMENSION(:), ALLOCATABLE, TARGET :: lp
    CALL rank_check(caller, 2, SHAPE(xp), Np, SIZE(xmin), SIZE(xmax), Nm, Ngl, nbox, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    ALLOCATE(lp(MAX(n, 1)))
    ! lhbx(nbox+1), lpdx(icount) as ppm_util_rank2d.f (ppm_alloc fit)
    IF (ASSOCIATED(lhbx)) DEALLOCATE(lhbx)
    ALLOCATE(lhbx(nbox + 1))
    CALL rank_core2_s(xp(1:2, 1:n), n, xmin(1:2), xmax(1:2), Nm, Ngl, lp, lhbx, icount, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(lpdx)) DEALLOCATE(lpdx)
    ALLOCATE(lpdx(icount))
    DO d = 1, icount
      lpdx(d) = lp(d)
    END DO
  END SUBROUTINE ppm_util_rank2d_s

  SUBROUTINE rank_core2_s(xp, n, xmin, xmax, Nm, Ngl, lp, lh, icount, info)
    REAL(C_FLOAT), TARGET, INTENT(IN) :: xp(*), xmin(*), xmax(*)
    INTEGER, INTENT(IN) :: n
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, Ngl
    INTEGER(C_INT), TARGET :: lp(*), lh(*)
    INTEGER, INTENT(OUT) :: icount, info
    INTEGER(C_INT) :: cnm(3), cngl(6), ic
    INTEGER :: d
    cnm = 1
    cngl = 0
    DO d = 1, SIZE(Nm)
      cnm(d) = Nm(d)
    END DO
    DO d = 1, SIZE(Ngl)
      cngl(d) = Ngl(d)
    END DO
    info = ppmg_util_rank_reg(INT(2, C_INT), ploc_s(xp, n), INT(n, C_INT), C_LOC(xmin(1)), C_LOC(xmax(1)), cnm, &
                              cngl, C_LOC(lp(1)), C_LOC(lh(1)), ic, ppmg_prec_f32, ppmg_mem_host)
    icount = ic
    CALL ppmg_report('ppm_util_rank2d', info)
  END SUBROUTINE rank_core2_s

  SUBROUTINE ppm_util_rank2d_d(xp, Np, xmin, xmax, Nm, Ngl, lpdx, lhbx, info)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), INTENT(IN) :: xp
    INTEGER, INTENT(IN) :: Np
    REAL(MK), DIMENSION(:), INTENT(IN) :: xmin, xmax
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, Ngl
    INTEGER, DIMENSION(:), POINTER :: lpdx, lhbx
    INTEGER, INTENT(INOUT) :: info
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_util_rank2d'
    INTEGER :: nbox, d, icount, n
    INTEGER, DIMENSION(:), ALLOCATABLE, TARGET :: lp
    CALL rank_check(caller, 2, SHAPE(xp), Np, SIZE(xmin), SIZE(xmax), Nm, Ngl, nbox, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    ALLOCATE(lp(MAX(n, 1)))
    ! lhbx(nbox+1), lpdx(icount) as ppm_util_rank2d.f (ppm_alloc fit)
    IF (ASSOCIATED(lhbx)) DEALLOCATE(lhbx)
    ALLOCATE(lhbx(nbox + 1))
    CALL rank_core2_d(xp(1:2, 1:n), n, xmin(1:2), xmax(1:2), Nm, Ngl, lp, lhbx, icount, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(lpdx)) DEALLOCATE(lpdx)
    ALLOCATE(lpdx(icount))
    DO d = 1, icount
      lpdx(d) = lp(d)
    END DO
  END SUBROUTINE ppm_util_rank2d_d

  SUBROUTINE rank_core2_d(xp, n, xmin, xmax, Nm, Ngl, lp, lh, icount, info)
    REAL(C_DOUBLE), TARGET, INTENT(IN) :: xp(*), xmin(*), xmax(*)
    INTEGER, INTENT(IN) :: n
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, Ngl
    INTEGER(C_INT), TARGET :: lp(*), lh(*)
    INTEGER, INTENT(OUT) :: icount, info
    INTEGER(C_INT) :: cnm(3), cngl(6), ic
    INTEGER :: d
    cnm = 1
    cngl = 0
    DO d = 1, SIZE(Nm)
      cnm(d) = Nm(d)
    END DO
    DO d = 1, SIZE(Ngl)
      cngl(d) = Ngl(d)
    END DO
    info = ppmg_util_rank_reg(INT(2, C_INT), ploc_d(xp, n), INT(n, C_INT), C_LOC(xmin(1)), C_LOC(xmax(1)), cnm, &
                              cngl, C_LOC(lp(1)), C_LOC(lh(1)), ic, ppmg_prec_f64, ppmg_mem_host)
    icount = ic
    CALL ppmg_report('ppm_util_rank2d', info)
  END SUBROUTINE rank_core2_d

  SUBROUTINE rank_check(caller, dim, xshape, np, nmin, nmax, Nm, Ngl, nbox, info)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    INTEGER, INTENT(IN) :: dim, np, nmin, nmax
    INTEGER, DIMENSION(:), INTENT(IN) :: xshape, Nm, Ngl
    INTEGER, INTENT(OUT) :: nbox, info
    INTEGER :: d
    info = 0
    nbox = 1
    IF (np < 0) THEN
      CALL ppmg_argfail(caller, 'Np must be >= 0', info)
      RETURN
    END IF
    IF (SIZE(Nm) < dim .OR. SIZE(Ngl) < 2 * dim .OR. nmin < dim .OR. nmax < dim) THEN
      CALL ppmg_argfail(caller, 'Nm, Ngl, xmin or xmax too short', info)
      RETURN
    END IF
    IF (np > 0 .AND. (xshape(1) < dim .OR. xshape(2) < np)) THEN
      CALL ppmg_argfail(caller, 'xp too small', info)
      RETURN
    END IF
    DO d = 1, dim
      IF (Nm(d) <= 0) THEN
        CALL ppmg_argfail(caller, 'Nm must be > 0', info)
        RETURN
      END IF
      IF (Ngl(d) < 0 .OR. Ngl(dim + d) < 0) THEN
        CALL ppmg_argfail(caller, 'Ngl must be >= 0', info)
        RETURN
      END IF
      nbox = nbox * (Nm(d) + Ngl(d) + Ngl(dim + d))
    END DO
  END SUBROUTINE rank_check

END MODULE ppm_module_util_rank
