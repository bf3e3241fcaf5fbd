!-------------------------------------------------------------------------------
! ppm_module_map_field_ghost -- generic ppm_map_field_ghost with the specifics
! of the reference (ppm_module_map_field_ghost.f:59-100: single, double,
! single and double complex, integer and logical fields, scalar and vector, 2D
! and 3D; arguments as in ppm_map_field_ghost.f:134-240: fv, [lda,] topo_id,
! mesh_id, ghostsize, maptype, info [, mask]) on one sub covering the domain.
! The call sequence is the reference's: init, then ghost_get or ghost_put,
! then push, send and pop.  On one process push and send move no data; pop
! applies the mapping on the GPU (ghost_get: every periodic image of a real
! node, in the ghost layers or on the duplicate face, receives the node's
! value; ghost_put: every image is added onto its real node, integers as
! integers, logicals by .AND.).  The optional mask (mask(i,j[,k],isub) over
! the same box as the field) is applied by the pop as in
! ppm_map_field_pop_3d.f:1818-1897: a node whose mask is .FALSE. keeps its
! value.
!-------------------------------------------------------------------------------
MODULE ppm_module_map_field_ghost
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  USE ppmg_fcore
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: ppm_map_field_ghost
  INTERFACE ppm_map_field_ghost
    MODULE PROCEDURE ppm_map_field_ghost_2d_sca_s
    MODULE PROCEDURE ppm_map_field_ghost_2d_sca_d
    MODULE PROCEDURE ppm_map_field_ghost_2d_sca_sc
    MODULE PROCEDURE ppm_map_field_ghost_2d_sca_dc
    MODULE PROCEDURE ppm_map_field_ghost_2d_sca_i
    MODULE PROCEDURE ppm_map_field_ghost_2d_sca_l
    MODULE PROCEDURE ppm_map_field_ghost_2d_vec_s
    MODULE PROCEDURE ppm_map_field_ghost_2d_vec_d
    MODULE PROCEDURE ppm_map_field_ghost_2d_vec_sc
    MODULE PROCEDURE ppm_map_field_ghost_2d_vec_dc
    MODULE PROCEDURE ppm_map_field_ghost_2d_vec_i
    MODULE PROCEDURE ppm_map_field_ghost_2d_vec_l
    MODULE PROCEDURE ppm_map_field_ghost_3d_sca_s
    MODULE PROCEDURE ppm_map_field_ghost_3d_sca_d
    MODULE PROCEDURE ppm_map_field_ghost_3d_sca_sc
    MODULE PROCEDURE ppm_map_field_ghost_3d_sca_dc
    MODULE PROCEDURE ppm_map_field_ghost_3d_sca_i
    MODULE PROCEDURE ppm_map_field_ghost_3d_sca_l
    MODULE PROCEDURE ppm_map_field_ghost_3d_vec_s
    MODULE PROCEDURE ppm_map_field_ghost_3d_vec_d
    MODULE PROCEDURE ppm_map_field_ghost_3d_vec_sc
    MODULE PROCEDURE ppm_map_field_ghost_3d_vec_dc
    MODULE PROCEDURE ppm_map_field_ghost_3d_vec_i
    MODULE PROCEDURE ppm_map_field_ghost_3d_vec_l
  END INTERFACE
  ! mapping requested by ghost_get / ghost_put (0: none), and whether it was pushed
  INTEGER, SAVE :: pending = 0
  LOGICAL, SAVE :: pushed = .FALSE.
CONTAINS

  ! Protocol of ppm_map_field_ghost.f:404-714: go = .TRUE. when this call is
  ! the pop that carries out the pending mapping.
  SUBROUTINE ghost_step(maptype, lda, go, info)
    INTEGER, INTENT(IN) :: maptype, lda
    LOGICAL, INTENT(OUT) :: go
    INTEGER, INTENT(OUT) :: info
    info = 0
    go = .FALSE.
    IF (lda < 1) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'lda must be >0', info)
      RETURN
    END IF
    SELECT CASE (maptype)
    CASE (ppm_param_map_init, ppm_param_map_cancel)
      pending = 0
      pushed = .FALSE.
    CASE (ppm_param_map_ghost_get, ppm_param_map_ghost_put)
      pending = maptype
      pushed = .FALSE.
    CASE (ppm_param_map_push)
      IF (pending == 0) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'push without a ghost_get or ghost_put mapping', info)
        RETURN
      END IF
      pushed = .TRUE.
    CASE (ppm_param_map_send)
      IF (.NOT. pushed) CALL ppmg_argfail('ppm_map_field_ghost', 'send without push', info)
    CASE (ppm_param_map_pop)
      IF (.NOT. pushed) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'pop without push', info)
        RETURN
      END IF
      pushed = .FALSE.
      go = .TRUE.
    CASE DEFAULT
      CALL ppmg_argfail('ppm_map_field_ghost', 'unknown maptype', info)
    END SELECT
  END SUBROUTINE ghost_step

  ! The mask's bounds must cover the box lo..hi and sub 1 (the reference's
  ! checks, ppm_map_field_pop_3d.f:345-385).
  SUBROUTINE mask_check(lb, ub, dim, lo, hi, info)
    INTEGER, DIMENSION(:), INTENT(IN) :: lb, ub
    INTEGER, INTENT(IN) :: dim, lo(3), hi(3)
    INTEGER, INTENT(OUT) :: info
    INTEGER :: d
    info = 0
    DO d = 1, dim
      IF (lb(d) > lo(d) .OR. ub(d) < hi(d)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask does not match the mesh and its ghost layers', info)
        RETURN
      END IF
    END DO
    IF (lb(dim + 1) > 1 .OR. ub(dim + 1) < 1) CALL ppmg_argfail('ppm_map_field_ghost', 'mask has no sub 1', info)
  END SUBROUTINE mask_check

  SUBROUTINE ppm_map_field_ghost_2d_sca_s(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    REAL(ppm_kind_single), DIMENSION(:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_s(fv(lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_sca_s

  SUBROUTINE ppm_map_field_ghost_2d_sca_d(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    REAL(ppm_kind_double), DIMENSION(:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_d(fv(lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_sca_d

  SUBROUTINE ppm_map_field_ghost_2d_sca_sc(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    COMPLEX(ppm_kind_single), DIMENSION(:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_sc(fv(lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_sca_sc

  SUBROUTINE ppm_map_field_ghost_2d_sca_dc(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    COMPLEX(ppm_kind_double), DIMENSION(:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_dc(fv(lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_sca_dc

  SUBROUTINE ppm_map_field_ghost_2d_sca_i(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    INTEGER, DIMENSION(:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_i(fv(lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_sca_i

  SUBROUTINE ppm_map_field_ghost_2d_sca_l(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    LOGICAL, DIMENSION(:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_l(fv(lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_sca_l

  SUBROUTINE ppm_map_field_ghost_2d_vec_s(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    REAL(ppm_kind_single), DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_s(fv(1:lda, lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_vec_s

  SUBROUTINE ppm_map_field_ghost_2d_vec_d(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    REAL(ppm_kind_double), DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_d(fv(1:lda, lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_vec_d

  SUBROUTINE ppm_map_field_ghost_2d_vec_sc(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    COMPLEX(ppm_kind_single), DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_sc(fv(1:lda, lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_vec_sc

  SUBROUTINE ppm_map_field_ghost_2d_vec_dc(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    COMPLEX(ppm_kind_double), DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_dc(fv(1:lda, lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_vec_dc

  SUBROUTINE ppm_map_field_ghost_2d_vec_i(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    INTEGER, DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_i(fv(1:lda, lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_vec_i

  SUBROUTINE ppm_map_field_ghost_2d_vec_l(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    LOGICAL, DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 2, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_l(fv(1:lda, lo(1):hi(1), lo(2):hi(2), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_2d_vec_l

  SUBROUTINE ppm_map_field_ghost_3d_sca_s(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    REAL(ppm_kind_single), DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_s(fv(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_sca_s

  SUBROUTINE ppm_map_field_ghost_3d_sca_d(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    REAL(ppm_kind_double), DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_d(fv(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_sca_d

  SUBROUTINE ppm_map_field_ghost_3d_sca_sc(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    COMPLEX(ppm_kind_single), DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_sc(fv(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_sca_sc

  SUBROUTINE ppm_map_field_ghost_3d_sca_dc(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    COMPLEX(ppm_kind_double), DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_dc(fv(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_sca_dc

  SUBROUTINE ppm_map_field_ghost_3d_sca_i(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    INTEGER, DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_i(fv(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_sca_i

  SUBROUTINE ppm_map_field_ghost_3d_sca_l(fv, topo_id, mesh_id, ghostsize, maptype, info, mask)
    LOGICAL, DIMENSION(:,:,:,:), POINTER :: fv
    INTEGER, PARAMETER :: lda = 1
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_l(fv(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_sca_l

  SUBROUTINE ppm_map_field_ghost_3d_vec_s(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    REAL(ppm_kind_single), DIMENSION(:,:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_s(fv(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_vec_s

  SUBROUTINE ppm_map_field_ghost_3d_vec_d(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    REAL(ppm_kind_double), DIMENSION(:,:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_d(fv(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_vec_d

  SUBROUTINE ppm_map_field_ghost_3d_vec_sc(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    COMPLEX(ppm_kind_single), DIMENSION(:,:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_sc(fv(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_vec_sc

  SUBROUTINE ppm_map_field_ghost_3d_vec_dc(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    COMPLEX(ppm_kind_double), DIMENSION(:,:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_dc(fv(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_vec_dc

  SUBROUTINE ppm_map_field_ghost_3d_vec_i(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    INTEGER, DIMENSION(:,:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_i(fv(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_vec_i

  SUBROUTINE ppm_map_field_ghost_3d_vec_l(fv, lda, topo_id, mesh_id, ghostsize, maptype, info, mask)
    LOGICAL, DIMENSION(:,:,:,:,:), POINTER :: fv
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: topo_id, mesh_id, maptype
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: info
    LOGICAL, DIMENSION(:,:,:,:), POINTER, OPTIONAL :: mask
    INTEGER :: lo(3), hi(3)
    INTEGER(C_INT) :: cg(3)
    INTEGER(C_INT), ALLOCATABLE :: mi(:)
    LOGICAL :: go
    CALL ghost_step(maptype, lda, go, info)
    IF (.NOT. go) RETURN
    CALL ppmg_box('ppm_map_field_ghost', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(fv)) THEN
      CALL ppmg_argfail('ppm_map_field_ghost', 'fv is not associated', info)
      RETURN
    END IF
    CALL ppmg_field_check('ppm_map_field_ghost', LBOUND(fv), UBOUND(fv), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (PRESENT(mask)) THEN
      IF (.NOT. ASSOCIATED(mask)) THEN
        CALL ppmg_argfail('ppm_map_field_ghost', 'mask is not associated', info)
        RETURN
      END IF
      CALL mask_check(LBOUND(mask), UBOUND(mask), 3, lo, hi, info)
      IF (info /= 0) RETURN
      mi = PACK(MERGE(1_C_INT, 0_C_INT, mask(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1)), .TRUE.)
    ELSE
      ALLOCATE(mi(0))
    END IF
    CALL ghost_core_l(fv(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), lda, topo_id, mesh_id, cg, pending, mi, PRESENT(mask), info)
    pending = 0
  END SUBROUTINE ppm_map_field_ghost_3d_vec_l

END MODULE ppm_module_map_field_ghost
