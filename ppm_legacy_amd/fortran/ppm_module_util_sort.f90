!-------------------------------------------------------------------------------
! ppm_module_util_sort -- ppm_util_sort3d / ppm_util_sort2d (ppm_module_util_sort.f;
! arguments as ppm_util_sort3d.f:88-91): xp reordered in place so that the
! particles of each of the Nm boxes are consecutive (the stable order of
! ppm_util_rank3d with no ghost layers), npbx(1:nbox) (allocated here) the
! particles per box.  A particle outside the boxes is the reference's error
! ppm_err_part_lost.  Ranking and reordering run on the GPU.
!-------------------------------------------------------------------------------
MODULE ppm_module_util_sort
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  USE ppmg_fcore
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: ppm_util_sort3d, ppm_util_sort2d
  INTERFACE ppm_util_sort3d
    MODULE PROCEDURE ppm_util_sort3d_s
    MODULE PROCEDURE ppm_util_sort3d_d
  END INTERFACE
  INTERFACE ppm_util_sort2d
    MODULE PROCEDURE ppm_util_sort2d_s
    MODULE PROCEDURE ppm_util_sort2d_d
  END INTERFACE
CONTAINS

  SUBROUTINE sort_check(caller, dim, xshape, np, nmin, nmax, Nm, nbox, info)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    INTEGER, INTENT(IN) :: dim, np, nmin, nmax
    INTEGER, DIMENSION(:), INTENT(IN) :: xshape, Nm
    INTEGER, INTENT(OUT) :: nbox, info
    INTEGER :: d
    info = 0
    nbox = 1
    IF (np < 0) THEN
      CALL ppmg_argfail(caller, 'Np must be >0', info)
      RETURN
    END IF
    IF (SIZE(Nm) < dim .OR. nmin < dim .OR. nmax < dim) THEN
      CALL ppmg_argfail(caller, 'xmin, xmax and Nm need one entry per dimension', info)
      RETURN
    END IF
    DO d = 1, dim
      IF (Nm(d) <= 0) THEN
        CALL ppmg_argfail(caller, 'Nm must be >0', info)
        RETURN
      END IF
      nbox = nbox * Nm(d)
    END DO
    IF (np > 0 .AND. (xshape(1) < dim .OR. xshape(2) < np)) CALL ppmg_argfail(caller, 'xp too small', info)
  END SUBROUTINE sort_check

  SUBROUTINE ppm_util_sort3d_s(xp, Np, xmin, xmax, Nm, npbx, info)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), INTENT(INOUT) :: xp
    REAL(MK), DIMENSION(:), INTENT(IN) :: xmin, xmax
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm
    INTEGER, DIMENSION(:), POINTER :: npbx
    INTEGER, INTENT(IN) :: Np
    INTEGER, INTENT(INOUT) :: info
    INTEGER :: nbox, n
    CALL sort_check('ppm_util_sort3d', 3, SHAPE(xp), Np, SIZE(xmin), SIZE(xmax), Nm, nbox, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    ! npbx(nbox) as ppm_util_sort3d.f allocates it (ppm_alloc fit)
    IF (ASSOCIATED(npbx)) DEALLOCATE(npbx)
    ALLOCATE(npbx(nbox))
    CALL sort_core_s(3, xp(1:3, 1:MAX(n, 1)), n, xmin(1:3), xmax(1:3), Nm, npbx, info)
  END SUBROUTINE ppm_util_sort3d_s

  SUBROUTINE ppm_util_sort3d_d(xp, Np, xmin, xmax, Nm, npbx, info)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), INTENT(INOUT) :: xp
    REAL(MK), DIMENSION(:), INTENT(IN) :: xmin, xmax
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm
    INTEGER, DIMENSION(:), POINTER :: npbx
    INTEGER, INTENT(IN) :: Np
    INTEGER, INTENT(INOUT) :: info
    INTEGER :: nbox, n
    CALL sort_check('ppm_util_sort3d', 3, SHAPE(xp), Np, SIZE(xmin), SIZE(xmax), Nm, nbox, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    ! npbx(nbox) as ppm_util_sort3d.f allocates it (ppm_alloc fit)
    IF (ASSOCIATED(npbx)) DEALLOCATE(npbx)
    ALLOCATE(npbx(nbox))
    CALL sort_core_d(3, xp(1:3, 1:MAX(n, 1)), n, xmin(1:3), xmax(1:3), Nm, npbx, info)
  END SUBROUTINE ppm_util_sort3d_d

  SUBROUTINE ppm_util_sort2d_s(xp, Np, xmin, xmax, Nm, npbx, info)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), INTENT(INOUT) :: xp
    REAL(MK), DIMENSION(:), INTENT(IN) :: xmin, xmax
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm
    INTEGER, DIMENSION(:), POINTER :: npbx
    INTEGER, INTENT(IN) :: Np
    INTEGER, INTENT(INOUT) :: info
    INTEGER :: nbox, n
    CALL sort_check('ppm_util_sort2d', 2, SHAPE(xp), Np, SIZE(xmin), SIZE(xmax), Nm, nbox, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    ! npbx(nbox) as ppm_util_sort2d.f allocates it (ppm_alloc fit)
    IF (ASSOCIATED(npbx)) DEALLOCATE(npbx)
    ALLOCATE(npbx(nbox))
    CALL sort_core_s(2, xp(1:2, 1:MAX(n, 1)), n, xmin(1:2), xmax(1:2), Nm, npbx, info)
  END SUBROUTINE ppm_util_sort2d_s

  SUBROUTINE ppm_util_sort2d_d(xp, Np, xmin, xmax, Nm, npbx, info)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), INTENT(INOUT) :: xp
    REAL(MK), DIMENSION(:), INTENT(IN) :: xmin, xmax
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm
    INTEGER, DIMENSION(:), POINTER :: npbx
    INTEGER, INTENT(IN) :: Np
    INTEGER, INTENT(INOUT) :: info
    INTEGER :: nbox, n
    CALL sort_check('ppm_util_sort2d', 2, SHAPE(xp), Np, SIZE(xmin), SIZE(xmax), Nm, nbox, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    ! npbx(nbox) as ppm_util_sort2d.f allocates it (ppm_alloc fit)
    IF (ASSOCIATED(npbx)) DEALLOCATE(npbx)
    ALLOCATE(npbx(nbox))
    CALL sort_core_d(2, xp(1:2, 1:MAX(n, 1)), n, xmin(1:2), xmax(1:2), Nm, npbx, info)
  END SUBROUTINE ppm_util_sort2d_d

  SUBROUTINE sort_core_s(dim, xp, n, xmin, xmax, Nm, npbx, info)
    INTEGER, INTENT(IN) :: dim, n
    REAL(C_FLOAT), TARGET :: xp(*)
    REAL(C_FLOAT), TARGET, INTENT(IN) :: xmin(*), xmax(*)
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm
    INTEGER(C_INT), TARGET :: npbx(*)
    INTEGER, INTENT(OUT) :: info
    INTEGER(C_INT) :: cnm(3)
    INTEGER :: d
    cnm = 1
    DO d = 1, dim
      cnm(d) = Nm(d)
    END DO
    info = ppmg_util_sort_reg(INT(dim, C_INT), C_LOC(xp(1)), INT(n, C_INT), C_LOC(xmin(1)), C_LOC(xmax(1)), cnm, &
                              C_LOC(npbx(1)), ppmg_prec_f32, ppmg_mem_host)
    CALL ppmg_report('ppm_util_sort', info)
  END SUBROUTINE sort_core_s

  SUBROUTINE sort_core_d(dim, xp, n, xmin, xmax, Nm, npbx, info)
    INTEGER, INTENT(IN) :: dim, n
    REAL(C_DOUBLE), TARGET :: xp(*)
    REAL(C_DOUBLE), TARGET, INTENT(IN) :: xmin(*), xmax(*)
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm
    INTEGER(C_INT), TARGET :: npbx(*)
    INTEGER, INTENT(OUT) :: info
    INTEGER(C_INT) :: cnm(3)
    INTEGER :: d
    cnm = 1
    DO d = 1, dim
      cnm(d) = Nm(d)
    END DO
    info = ppmg_util_sort_reg(INT(dim, C_INT), C_LOC(xp(1)), INT(n, C_INT), C_LOC(xmin(1)), C_LOC(xmax(1)), cnm, &
                              C_LOC(npbx(1)), ppmg_prec_f64, ppmg_mem_host)
    CALL ppmg_report('ppm_util_sort', info)
  END SUBROUTINE sort_core_d

END MODULE ppm_module_util_sort
