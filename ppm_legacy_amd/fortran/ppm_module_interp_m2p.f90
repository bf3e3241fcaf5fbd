!-------------------------------------------------------------------------------
! ppm_module_interp_m2p -- generic ppm_interp_m2p with the eight specifics of
! the reference (ppm_module_interp_m2p.f; arguments as in ppm_interp_m2p.f:
! 129-159): up([lda,]Np) is overwritten with the field interpolated at xp.
! The ghost layers the reference fills by ghost_get are filled by the GPU
! library from the real nodes (periodic) before the gather.
!-------------------------------------------------------------------------------
MODULE ppm_module_interp_m2p
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  USE ppmg_fcore
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: ppm_interp_m2p
  INTERFACE ppm_interp_m2p
    MODULE PROCEDURE ppm_interp_m2p_ss_2d
    MODULE PROCEDURE ppm_interp_m2p_ds_2d
    MODULE PROCEDURE ppm_interp_m2p_sv_2d
    MODULE PROCEDURE ppm_interp_m2p_dv_2d
    MODULE PROCEDURE ppm_interp_m2p_ss_3d
    MODULE PROCEDURE ppm_interp_m2p_ds_3d
    MODULE PROCEDURE ppm_interp_m2p_sv_3d
    MODULE PROCEDURE ppm_interp_m2p_dv_3d
  END INTERFACE
CONTAINS

  SUBROUTINE ppm_interp_m2p_ss_2d(xp, Np, up, topo_id, mesh_id, kernel, ghostsize, field_up, info)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:), POINTER :: up
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    INTEGER :: lo(3), hi(3), n
    INTEGER(C_INT) :: cg(3)
    INTEGER, PARAMETER :: lda = 1
    CALL ppmg_box('ppm_interp_m2p', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'field_up is not associated', info)
      RETURN
    END IF
    IF (Np > 0 .AND. (.NOT. ASSOCIATED(xp) .OR. .NOT. ASSOCIATED(up))) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'xp or up is not associated', info)
      RETURN
    END IF
    IF (Np == 0) RETURN
    CALL ppmg_part_check('ppm_interp_m2p', Np, SHAPE(xp), 2, SHAPE(up), 0, info)
    IF (info /= 0) RETURN
    CALL ppmg_field_check('ppm_interp_m2p', LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    CALL m2p_core_s(xp(1:2, 1:n), up(1:n), field_up(lo(1):hi(1), lo(2):hi(2), 1), &
                    Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_interp_m2p_ss_2d

  SUBROUTINE ppm_interp_m2p_ds_2d(xp, Np, up, topo_id, mesh_id, kernel, ghostsize, field_up, info)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:), POINTER :: up
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    INTEGER :: lo(3), hi(3), n
    INTEGER(C_INT) :: cg(3)
    INTEGER, PARAMETER :: lda = 1
    CALL ppmg_box('ppm_interp_m2p', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'field_up is not associated', info)
      RETURN
    END IF
    IF (Np > 0 .AND. (.NOT. ASSOCIATED(xp) .OR. .NOT. ASSOCIATED(up))) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'xp or up is not associated', info)
      RETURN
    END IF
    IF (Np == 0) RETURN
    CALL ppmg_part_check('ppm_interp_m2p', Np, SHAPE(xp), 2, SHAPE(up), 0, info)
    IF (info /= 0) RETURN
    CALL ppmg_field_check('ppm_interp_m2p', LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    CALL m2p_core_d(xp(1:2, 1:n), up(1:n), field_up(lo(1):hi(1), lo(2):hi(2), 1), &
                    Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_interp_m2p_ds_2d

  SUBROUTINE ppm_interp_m2p_sv_2d(xp, Np, up, lda, topo_id, mesh_id, kernel, ghostsize, field_up, info)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    INTEGER :: lo(3), hi(3), n
    INTEGER(C_INT) :: cg(3)
    CALL ppmg_box('ppm_interp_m2p', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'field_up is not associated', info)
      RETURN
    END IF
    IF (Np > 0 .AND. (.NOT. ASSOCIATED(xp) .OR. .NOT. ASSOCIATED(up))) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'xp or up is not associated', info)
      RETURN
    END IF
    IF (Np == 0) RETURN
    CALL ppmg_part_check('ppm_interp_m2p', Np, SHAPE(xp), 2, SHAPE(up), lda, info)
    IF (info /= 0) RETURN
    CALL ppmg_field_check('ppm_interp_m2p', LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    CALL m2p_core_s(xp(1:2, 1:n), up(1:lda, 1:n), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
                    Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_interp_m2p_sv_2d

  SUBROUTINE ppm_interp_m2p_dv_2d(xp, Np, up, lda, topo_id, mesh_id, kernel, ghostsize, field_up, info)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    INTEGER :: lo(3), hi(3), n
    INTEGER(C_INT) :: cg(3)
    CALL ppmg_box('ppm_interp_m2p', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'field_up is not associated', info)
      RETURN
    END IF
    IF (Np > 0 .AND. (.NOT. ASSOCIATED(xp) .OR. .NOT. ASSOCIATED(up))) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'xp or up is not associated', info)
      RETURN
    END IF
    IF (Np == 0) RETURN
    CALL ppmg_part_check('ppm_interp_m2p', Np, SHAPE(xp), 2, SHAPE(up), lda, info)
    IF (info /= 0) RETURN
    CALL ppmg_field_check('ppm_interp_m2p', LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    CALL m2p_core_d(xp(1:2, 1:n), up(1:lda, 1:n), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
                    Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_interp_m2p_dv_2d

  SUBROUTINE ppm_interp_m2p_ss_3d(xp, Np, up, topo_id, mesh_id, kernel, ghostsize, field_up, info)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:), POINTER :: up
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    INTEGER :: lo(3), hi(3), n
    INTEGER(C_INT) :: cg(3)
    INTEGER, PARAMETER :: lda = 1
    CALL ppmg_box('ppm_interp_m2p', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'field_up is not associated', info)
      RETURN
    END IF
    IF (Np > 0 .AND. (.NOT. ASSOCIATED(xp) .OR. .NOT. ASSOCIATED(up))) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'xp or up is not associated', info)
      RETURN
    END IF
    IF (Np == 0) RETURN
    CALL ppmg_part_check('ppm_interp_m2p', Np, SHAPE(xp), 3, SHAPE(up), 0, info)
    IF (info /= 0) RETURN
    CALL ppmg_field_check('ppm_interp_m2p', LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    CALL m2p_core_s(xp(1:3, 1:n), up(1:n), field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), &
                    Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_interp_m2p_ss_3d

  SUBROUTINE ppm_interp_m2p_ds_3d(xp, Np, up, topo_id, mesh_id, kernel, ghostsize, field_up, info)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:), POINTER :: up
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    INTEGER :: lo(3), hi(3), n
    INTEGER(C_INT) :: cg(3)
    INTEGER, PARAMETER :: lda = 1
    CALL ppmg_box('ppm_interp_m2p', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'field_up is not associated', info)
      RETURN
    END IF
    IF (Np > 0 .AND. (.NOT. ASSOCIATED(xp) .OR. .NOT. ASSOCIATED(up))) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'xp or up is not associated', info)
      RETURN
    END IF
    IF (Np == 0) RETURN
    CALL ppmg_part_check('ppm_interp_m2p', Np, SHAPE(xp), 3, SHAPE(up), 0, info)
    IF (info /= 0) RETURN
    CALL ppmg_field_check('ppm_interp_m2p', LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    CALL m2p_core_d(xp(1:3, 1:n), up(1:n), field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), &
                    Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_interp_m2p_ds_3d

  SUBROUTINE ppm_interp_m2p_sv_3d(xp, Np, up, lda, topo_id, mesh_id, kernel, ghostsize, field_up, info)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    INTEGER :: lo(3), hi(3), n
    INTEGER(C_INT) :: cg(3)
    CALL ppmg_box('ppm_interp_m2p', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'field_up is not associated', info)
      RETURN
    END IF
    IF (Np > 0 .AND. (.NOT. ASSOCIATED(xp) .OR. .NOT. ASSOCIATED(up))) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'xp or up is not associated', info)
      RETURN
    END IF
    IF (Np == 0) RETURN
    CALL ppmg_part_check('ppm_interp_m2p', Np, SHAPE(xp), 3, SHAPE(up), lda, info)
    IF (info /= 0) RETURN
    CALL ppmg_field_check('ppm_interp_m2p', LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    CALL m2p_core_s(xp(1:3, 1:n), up(1:lda, 1:n), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), &
                    Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_interp_m2p_sv_3d

  SUBROUTINE ppm_interp_m2p_dv_3d(xp, Np, up, lda, topo_id, mesh_id, kernel, ghostsize, field_up, info)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    INTEGER :: lo(3), hi(3), n
    INTEGER(C_INT) :: cg(3)
    CALL ppmg_box('ppm_interp_m2p', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'field_up is not associated', info)
      RETURN
    END IF
    IF (Np > 0 .AND. (.NOT. ASSOCIATED(xp) .OR. .NOT. ASSOCIATED(up))) THEN
      CALL ppmg_argfail('ppm_interp_m2p', 'xp or up is not associated', info)
      RETURN
    END IF
    IF (Np == 0) RETURN
    CALL ppmg_part_check('ppm_interp_m2p', Np, SHAPE(xp), 3, SHAPE(up), lda, info)
    IF (info /= 0) RETURN
    CALL ppmg_field_check('ppm_interp_m2p', LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    n = MAX(Np, 0)
    CALL m2p_core_d(xp(1:3, 1:n), up(1:lda, 1:n), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), &
                    Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_interp_m2p_dv_3d

END MODULE ppm_module_interp_m2p
