!-------------------------------------------------------------------------------
! Cell lists of the drop-in: ppm_module_data_neighlist (the cell list type,
! ppm_module_data_neighlist.f:35-40), ppm_module_neighlist_clist
! (ppm_neighlist_clist.f:173-446), ppm_module_clist_destroy
! (ppm_clist_destroy.f:76-162) and the umbrella ppm_module_neighlist.  One sub
! covers the domain (the topology the last ppm_mktopo made), so clist and Nm
! have one entry: cells of at least the cutoff, one ghost layer on every side
! (or on the upper sides with lsymm), ranked by the GPU counting sort through
! ppm_util_rank3d / rank2d.  Verlet lists (ppm_neighlist_vlist) are not
! provided.
!-------------------------------------------------------------------------------
MODULE ppm_module_data_neighlist
  IMPLICIT NONE
  TYPE ppm_type_ptr_to_clist
    ! particle index list
    INTEGER, DIMENSION(:), POINTER :: lpdx => NULL()
    ! first particle in each cell
    INTEGER, DIMENSION(:), POINTER :: lhbx => NULL()
  END TYPE
END MODULE ppm_module_data_neighlist

MODULE ppm_module_clist_destroy
  USE ppm_module_data_neighlist
  IMPLICIT NONE
  INTERFACE ppm_clist_destroy
    MODULE PROCEDURE ppm_clist_destroy
  END INTERFACE
CONTAINS
  SUBROUTINE ppm_clist_destroy(clist, info)
    TYPE(ppm_type_ptr_to_clist), DIMENSION(:), POINTER :: clist
    INTEGER, INTENT(OUT) :: info
    INTEGER :: i
    info = 0
    IF (.NOT. ASSOCIATED(clist)) RETURN
    DO i = 1, SIZE(clist)
      IF (ASSOCIATED(clist(i)%lpdx)) DEALLOCATE(clist(i)%lpdx)
      IF (ASSOCIATED(clist(i)%lhbx)) DEALLOCATE(clist(i)%lhbx)
    END DO
    DEALLOCATE(clist)
    NULLIFY(clist)
  END SUBROUTINE ppm_clist_destroy
END MODULE ppm_module_clist_destroy

MODULE ppm_module_neighlist_clist
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  USE ppmg_fcore
  USE ppm_module_data_neighlist
  USE ppm_module_clist_destroy
  USE ppm_module_util_rank
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: ppm_neighlist_clist
  INTERFACE ppm_neighlist_clist
    MODULE PROCEDURE ppm_neighlist_clist_d
    MODULE PROCEDURE ppm_neighlist_clist_s
  END INTERFACE
CONTAINS

  ! Nm(dim, 1) and clist(1) (re)allocated as ppm_neighlist_clist.f:250-301 does.
  SUBROUTINE clist_setup(caller, ncut, np, clist, Nm, dim, info)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    INTEGER, INTENT(IN) :: ncut, np
    TYPE(ppm_type_ptr_to_clist), DIMENSION(:), POINTER :: clist
    INTEGER, DIMENSION(:,:), POINTER :: Nm
    INTEGER, INTENT(OUT) :: dim, info
    INTEGER :: tdim, nm3(3)
    info = 0
    CALL ppmg_topo_nm(ppmg_current_topo, tdim, nm3, info)
    IF (info /= 0) THEN
      CALL ppmg_argfail(caller, 'no topology has been defined so far', info)
      RETURN
    END IF
    dim = tdim
    IF (ncut < dim) THEN
      CALL ppmg_argfail(caller, 'cutoff must be given in all dimensions', info)
      RETURN
    END IF
    IF (np <= 0) THEN
      CALL ppmg_argfail(caller, 'Np must be >0', info)
      RETURN
    END IF
    IF (ASSOCIATED(Nm)) DEALLOCATE(Nm)
    ALLOCATE(Nm(dim, 1))
    Nm = 0
    IF (ASSOCIATED(clist)) THEN
      IF (SIZE(clist) /= 1) CALL ppm_clist_destroy(clist, info)
    END IF
    IF (.NOT. ASSOCIATED(clist)) ALLOCATE(clist(1))
  END SUBROUTINE clist_setup

  SUBROUTINE clist_ghosts(dim, lsymm, ngl)
    INTEGER, INTENT(IN) :: dim
    LOGICAL, INTENT(IN) :: lsymm
    INTEGER, INTENT(OUT) :: ngl(6)
    INTEGER :: i
    ngl = 0
    DO i = 1, dim
      ngl(dim + i) = 1
      IF (.NOT. lsymm) ngl(i) = 1
    END DO
  END SUBROUTINE clist_ghosts

  SUBROUTINE ppm_neighlist_clist_s(xp, Np, cutoff, lsymm, clist, Nm, info, pidx)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), INTENT(IN) :: xp
    INTEGER, INTENT(IN) :: Np
    REAL(MK), DIMENSION(:), INTENT(IN) :: cutoff
    LOGICAL, INTENT(IN) :: lsymm
    TYPE(ppm_type_ptr_to_clist), DIMENSION(:), POINTER :: clist
    INTEGER, DIMENSION(:,:), POINTER :: Nm
    INTEGER, INTENT(OUT) :: info
    INTEGER, DIMENSION(:), OPTIONAL :: pidx
    REAL(MK) :: xmin(3), xmax(3)
    INTEGER :: dim, ngl(6), d
    CALL clist_setup('ppm_neighlist_clist', SIZE(cutoff), Np, clist, Nm, dim, info)
    IF (info /= 0) RETURN
    CALL clist_extent_s(xmin, xmax)
    DO d = 1, dim
      IF (cutoff(d) <= 0.0_MK) THEN
        CALL ppmg_argfail('ppm_neighlist_clist', 'cutoff must be >0 in all dimensions', info)
        RETURN
      END IF
      ! cells of at least the cutoff, at least one (ppm_neighlist_clist.f:365-371)
      Nm(d, 1) = INT((xmax(d) - xmin(d)) / cutoff(d))
      IF (Nm(d, 1) < 1) Nm(d, 1) = 1
    END DO
    CALL clist_ghosts(dim, lsymm, ngl)
    IF (dim == 3) THEN
      IF (PRESENT(pidx)) THEN
        CALL ppm_util_rank3d(xp(1:3, pidx), SIZE(pidx), xmin(1:3), xmax(1:3), Nm(1:3, 1), ngl(1:6), &
                             clist(1)%lpdx, clist(1)%lhbx, info)
      ELSE
        CALL ppm_util_rank3d(xp, Np, xmin(1:3), xmax(1:3), Nm(1:3, 1), ngl(1:6), clist(1)%lpdx, clist(1)%lhbx, info)
      END IF
    ELSE
      IF (PRESENT(pidx)) THEN
        CALL ppm_util_rank2d(xp(1:2, pidx), SIZE(pidx), xmin(1:2), xmax(1:2), Nm(1:2, 1), ngl(1:4), &
                             clist(1)%lpdx, clist(1)%lhbx, info)
      ELSE
        CALL ppm_util_rank2d(xp, Np, xmin(1:2), xmax(1:2), Nm(1:2, 1), ngl(1:4), clist(1)%lpdx, clist(1)%lhbx, info)
      END IF
    END IF
    ! Nm counts the ghost layers as well (ppm_neighlist_clist.f:403-424)
    DO d = 1, dim
      Nm(d, 1) = Nm(d, 1) + ngl(d) + ngl(dim + d)
    END DO
    IF (info < 0) THEN
      info = ppm_error_error
      IF (ppm_debug > 0) WRITE(*, '(A)') 'ppm_neighlist_clist : ranking of particles failed!'
    END IF
  END SUBROUTINE ppm_neighlist_clist_s

  SUBROUTINE clist_extent_s(xmin, xmax)
    REAL(ppm_kind_single), INTENT(OUT) :: xmin(3), xmax(3)
    INTEGER(C_INT) :: cdim, cnm(3), cbc(6), rc
    REAL(C_DOUBLE) :: mn(3), mx(3)
    rc = ppmg_topo_info(INT(ppmg_current_topo, C_INT), cdim, cnm, mn, mx, cbc)
    xmin = REAL(mn, ppm_kind_single)
    xmax = REAL(mx, ppm_kind_single)
  END SUBROUTINE clist_extent_s

  SUBROUTINE ppm_neighlist_clist_d(xp, Np, cutoff, lsymm, clist, Nm, info, pidx)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), INTENT(IN) :: xp
    INTEGER, INTENT(IN) :: Np
    REAL(MK), DIMENSION(:), INTENT(IN) :: cutoff
    LOGICAL, INTENT(IN) :: lsymm
    TYPE(ppm_type_ptr_to_clist), DIMENSION(:), POINTER :: clist
    INTEGER, DIMENSION(:,:), POINTER :: Nm
    INTEGER, INTENT(OUT) :: info
    INTEGER, DIMENSION(:), OPTIONAL :: pidx
    REAL(MK) :: xmin(3), xmax(3)
    INTEGER :: dim, ngl(6), d
    CALL clist_setup('ppm_neighlist_clist', SIZE(cutoff), Np, clist, Nm, dim, info)
    IF (info /= 0) RETURN
    CALL clist_extent_d(xmin, xmax)
    DO d = 1, dim
      IF (cutoff(d) <= 0.0_MK) THEN
        CALL ppmg_argfail('ppm_neighlist_clist', 'cutoff must be >0 in all dimensions', info)
        RETURN
      END IF
      ! cells of at least the cutoff, at least one (ppm_neighlist_clist.f:365-371)
      Nm(d, 1) = INT((xmax(d) - xmin(d)) / cutoff(d))
      IF (Nm(d, 1) < 1) Nm(d, 1) = 1
    END DO
    CALL clist_ghosts(dim, lsymm, ngl)
    IF (dim == 3) THEN
      IF (PRESENT(pidx)) THEN
        CALL ppm_util_rank3d(xp(1:3, pidx), SIZE(pidx), xmin(1:3), xmax(1:3), Nm(1:3, 1), ngl(1:6), &
                             clist(1)%lpdx, clist(1)%lhbx, info)
      ELSE
        CALL ppm_util_rank3d(xp, Np, xmin(1:3), xmax(1:3), Nm(1:3, 1), ngl(1:6), clist(1)%lpdx, clist(1)%lhbx, info)
      END IF
    ELSE
      IF (PRESENT(pidx)) THEN
        CALL ppm_util_rank2d(xp(1:2, pidx), SIZE(pidx), xmin(1:2), xmax(1:2), Nm(1:2, 1), ngl(1:4), &
                             clist(1)%lpdx, clist(1)%lhbx, info)
      ELSE
        CALL ppm_util_rank2d(xp, Np, xmin(1:2), xmax(1:2), Nm(1:2, 1), ngl(1:4), clist(1)%lpdx, clist(1)%lhbx, info)
      END IF
    END IF
    ! Nm counts the ghost layers as well (ppm_neighlist_clist.f:403-424)
    DO d = 1, dim
      Nm(d, 1) = Nm(d, 1) + ngl(d) + ngl(dim + d)
    END DO
    IF (info < 0) THEN
      info = ppm_error_error
      IF (ppm_debug > 0) WRITE(*, '(A)') 'ppm_neighlist_clist : ranking of particles failed!'
    END IF
  END SUBROUTINE ppm_neighlist_clist_d

  SUBROUTINE clist_extent_d(xmin, xmax)
    REAL(ppm_kind_double), INTENT(OUT) :: xmin(3), xmax(3)
    INTEGER(C_INT) :: cdim, cnm(3), cbc(6), rc
    REAL(C_DOUBLE) :: mn(3), mx(3)
    rc = ppmg_topo_info(INT(ppmg_current_topo, C_INT), cdim, cnm, mn, mx, cbc)
    xmin = REAL(mn, ppm_kind_double)
    xmax = REAL(mx, ppm_kind_double)
  END SUBROUTINE clist_extent_d

END MODULE ppm_module_neighlist_clist

MODULE ppm_module_neighlist
  USE ppm_module_data_neighlist
  USE ppm_module_neighlist_clist
  USE ppm_module_clist_destroy
END MODULE ppm_module_neighlist
