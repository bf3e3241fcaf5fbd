!-------------------------------------------------------------------------------
! Remeshing modules of the drop-in, same names as the reference:
!   ppm_module_data_rmsh          weights kept between comp_weights and remesh
!                                 (ppm_module_data_rmsh.f:53-58)
!   ppm_module_rmsh_comp_weights  ppm_rmsh_comp_weights_s/_d (ppm_rmsh_comp_weights.f:96-104)
!   ppm_module_rmsh_remesh        ppm_rmsh_remesh, eight specifics (ppm_rmsh_remesh.f:91-121)
!   ppm_module_rmsh_create_part   ppm_rmsh_create_part (ppm_rmsh_create_part.f:90-171)
!   ppm_module_rmsh               all of the above
! Weights are (1:2*ks, 1:nsublist, 1-ks:Np+ks) with ks = 1 (BSP2) or 2 (M'4),
! the reference's shape; the library reads wx(:, 1, 1:Np).
!-------------------------------------------------------------------------------
MODULE ppm_module_data_rmsh
  USE ppm_module_ppmg, ONLY: ppm_kind_single, ppm_kind_double
  IMPLICIT NONE
  REAL(ppm_kind_single), DIMENSION(:,:,:), POINTER :: rmsh_wx1_s => NULL(), rmsh_wx2_s => NULL(), rmsh_wx3_s => NULL()
  REAL(ppm_kind_double), DIMENSION(:,:,:), POINTER :: rmsh_wx1_d => NULL(), rmsh_wx2_d => NULL(), rmsh_wx3_d => NULL()
  ! half support of each kernel: BSP2 1, M'4 2 (ppm_rmsh_kernelsize, ppm_module_data_rmsh.f)
  INTEGER, DIMENSION(4), PARAMETER :: ppm_rmsh_kernelsize = (/1, 2, 2, 4/)
CONTAINS
  INTEGER FUNCTION ppmg_kernel_support(kernel)
    INTEGER, INTENT(IN) :: kernel
    ppmg_kernel_support = ppm_rmsh_kernelsize(kernel)
  END FUNCTION ppmg_kernel_support
  LOGICAL FUNCTION weights_hold(lb, ub, ks, np)
    INTEGER, DIMENSION(3), INTENT(IN) :: lb, ub
    INTEGER, INTENT(IN) :: ks, np
    weights_hold = lb(1) <= 1 .AND. ub(1) >= 2 * ks .AND. lb(2) <= 1 .AND. ub(2) >= 1 .AND. lb(3) <= 1 .AND. ub(3) >= np
  END FUNCTION weights_hold
END MODULE ppm_module_data_rmsh

MODULE ppm_module_rmsh_comp_weights
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  USE ppmg_fcore
  USE ppm_module_data_rmsh
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: ppm_rmsh_comp_weights, remesh_weights_s, remesh_weights_d
  INTERFACE ppm_rmsh_comp_weights
    MODULE PROCEDURE ppm_rmsh_comp_weights_s
    MODULE PROCEDURE ppm_rmsh_comp_weights_d
  END INTERFACE
CONTAINS
  SUBROUTINE ppm_rmsh_comp_weights_s(xp, Np, topo_id, mesh_id, kernel, info, wx1_user, wx2_user, wx3_user)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, INTENT(OUT) :: info
    REAL(MK), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    REAL(MK), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER :: dim, nm(3), ks, n
    LOGICAL :: user
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_comp_weights'
    CALL ppmg_topo_nm(topo_id, dim, nm, info)
    IF (info /= 0) THEN
      CALL ppmg_report(caller, info)
      RETURN
    END IF
    IF (kernel /= ppm_param_rmsh_kernel_bsp2 .AND. kernel /= ppm_param_rmsh_kernel_mp4) THEN
      CALL ppmg_argfail(caller, 'wrong kernel definition', info)
      RETURN
    END IF
    IF (dim == 3) THEN
      IF (PRESENT(wx1_user) .NEQV. PRESENT(wx2_user) .OR. PRESENT(wx1_user) .NEQV. PRESENT(wx3_user)) THEN
        CALL ppmg_argfail(caller, 'wx1_user, wx2_user and wx3_user go together', info)
        RETURN
      END IF
    ELSE
      IF (PRESENT(wx1_user) .NEQV. PRESENT(wx2_user)) THEN
        CALL ppmg_argfail(caller, 'a wx_user array is missing', info)
        RETURN
      END IF
    END IF
    IF (Np < 0) THEN
      CALL ppmg_argfail(caller, 'particles must be specified', info)
      RETURN
    END IF
    IF (Np == 0) RETURN
    IF (SIZE(xp, 1) < dim .OR. SIZE(xp, 2) < Np) THEN
      CALL ppmg_argfail(caller, 'xp too small', info)
      RETURN
    END IF
    ks = ppmg_kernel_support(kernel)
    ! weights (1:2ks, 1:nsublist, 1-ks:Np+ks) as ppm_rmsh_comp_weights.f:668-674
    user = .FALSE.
    IF (PRESENT(wx1_user)) user = ASSOCIATED(wx1_user) .AND. ASSOCIATED(wx2_user)
    IF (user .AND. dim == 3) user = ASSOCIATED(wx3_user)
    IF (user) THEN
      w1 => wx1_user
      w2 => wx2_user
      w3 => wx1_user
      IF (dim == 3) w3 => wx3_user
    ELSE
      CALL weights_fit_s(rmsh_wx1_s, ks, Np)
      CALL weights_fit_s(rmsh_wx2_s, ks, Np)
      CALL weights_fit_s(rmsh_wx3_s, ks, Np)
      w1 => rmsh_wx1_s
      w2 => rmsh_wx2_s
      w3 => rmsh_wx3_s
      IF (PRESENT(wx1_user)) THEN
        wx1_user => rmsh_wx1_s
        wx2_user => rmsh_wx2_s
        IF (dim == 3) wx3_user => rmsh_wx3_s
      END IF
    END IF
    IF (.NOT. weights_hold(LBOUND(w1), UBOUND(w1), ks, Np) .OR. .NOT. weights_hold(LBOUND(w2), UBOUND(w2), ks, Np) &
        .OR. .NOT. weights_hold(LBOUND(w3), UBOUND(w3), ks, Np)) THEN
      CALL ppmg_argfail(caller, 'weight arrays too small', info)
      RETURN
    END IF
    n = Np
    CALL cw_core_s(xp(1:dim, 1:n), w1(1:2*ks, 1, 1:n), w2(1:2*ks, 1, 1:n), w3(1:2*ks, 1, 1:n), &
                    Np, topo_id, mesh_id, kernel, info)
  END SUBROUTINE ppm_rmsh_comp_weights_s

  SUBROUTINE cw_core_s(xp, w1, w2, w3, np, topo_id, mesh_id, kernel, info)
    REAL(C_FLOAT), TARGET, INTENT(IN) :: xp(*)
    REAL(C_FLOAT), TARGET :: w1(*), w2(*), w3(*)
    INTEGER, INTENT(IN) :: np, topo_id, mesh_id, kernel
    INTEGER, INTENT(OUT) :: info
    info = ppmg_rmsh_comp_weights_reg(C_LOC(xp(1)), INT(np, C_INT), INT(topo_id, C_INT), INT(mesh_id, C_INT), &
                                      INT(kernel, C_INT), C_LOC(w1(1)), C_LOC(w2(1)), C_LOC(w3(1)), ppmg_prec_f32, &
                                          ppmg_mem_host)
    CALL ppmg_report('ppm_rmsh_comp_weights', info)
  END SUBROUTINE cw_core_s

  SUBROUTINE weights_fit_s(w, ks, np)
    REAL(ppm_kind_single), DIMENSION(:,:,:), POINTER :: w
    INTEGER, INTENT(IN) :: ks, np
    IF (ASSOCIATED(w)) THEN
      IF (SIZE(w, 1) == 2 * ks .AND. SIZE(w, 2) == 1 .AND. LBOUND(w, 3) == 1 - ks .AND. UBOUND(w, 3) == np + ks) RETURN
      DEALLOCATE(w)
    END IF
    ALLOCATE(w(1:2 * ks, 1:1, 1 - ks:np + ks))
    w = 0.0_ppm_kind_single
  END SUBROUTINE weights_fit_s

  ! Weights remesh uses: the caller's (all associated) or those comp_weights kept
  ! (ppm_rmsh_remesh.f:425-490).
  SUBROUTINE remesh_weights_s(caller, dim, kernel, np, w1, w2, w3, info, wx1_user, wx2_user, wx3_user)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    INTEGER, INTENT(IN) :: dim, kernel, np
    REAL(ppm_kind_single), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER, INTENT(OUT) :: info
    REAL(ppm_kind_single), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    INTEGER :: ks
    info = 0
    IF (kernel /= ppm_param_rmsh_kernel_bsp2 .AND. kernel /= ppm_param_rmsh_kernel_mp4) THEN
      CALL ppmg_argfail(caller, 'wrong kernel definition', info)
      RETURN
    END IF
    ks = ppmg_kernel_support(kernel)
    IF (PRESENT(wx1_user)) THEN
      IF (.NOT. PRESENT(wx2_user) .OR. (dim == 3 .AND. .NOT. PRESENT(wx3_user))) THEN
        CALL ppmg_argfail(caller, 'a wx_user array is missing', info)
        RETURN
      END IF
      IF (.NOT. ASSOCIATED(wx1_user) .OR. .NOT. ASSOCIATED(wx2_user)) THEN
        CALL ppmg_argfail(caller, 'cannot remesh with empty weights', info)
        RETURN
      END IF
      w1 => wx1_user
      w2 => wx2_user
      w3 => wx1_user
      IF (dim == 3) THEN
        IF (.NOT. ASSOCIATED(wx3_user)) THEN
          CALL ppmg_argfail(caller, 'cannot remesh with empty weights', info)
          RETURN
        END IF
        w3 => wx3_user
      END IF
    ELSE
      IF (.NOT. ASSOCIATED(rmsh_wx1_s)) THEN
        CALL ppmg_argfail(caller, 'compute weights before remeshing', info)
        RETURN
      END IF
      w1 => rmsh_wx1_s
      w2 => rmsh_wx2_s
      w3 => rmsh_wx3_s
    END IF
    IF (np <= 0) RETURN
    IF (.NOT. weights_hold(LBOUND(w1), UBOUND(w1), ks, np) .OR. .NOT. weights_hold(LBOUND(w2), UBOUND(w2), ks, np) &
        .OR. .NOT. weights_hold(LBOUND(w3), UBOUND(w3), ks, np)) THEN
      CALL ppmg_argfail(caller, 'weights do not cover the particles (kernel or Np changed since comp_weights?)', info)
      RETURN
    END IF
  END SUBROUTINE remesh_weights_s

  SUBROUTINE ppm_rmsh_comp_weights_d(xp, Np, topo_id, mesh_id, kernel, info, wx1_user, wx2_user, wx3_user)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, INTENT(OUT) :: info
    REAL(MK), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    REAL(MK), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER :: dim, nm(3), ks, n
    LOGICAL :: user
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_comp_weights'
    CALL ppmg_topo_nm(topo_id, dim, nm, info)
    IF (info /= 0) THEN
      CALL ppmg_report(caller, info)
      RETURN
    END IF
    IF (kernel /= ppm_param_rmsh_kernel_bsp2 .AND. kernel /= ppm_param_rmsh_kernel_mp4) THEN
      CALL ppmg_argfail(caller, 'wrong kernel definition', info)
      RETURN
    END IF
    IF (dim == 3) THEN
      IF (PRESENT(wx1_user) .NEQV. PRESENT(wx2_user) .OR. PRESENT(wx1_user) .NEQV. PRESENT(wx3_user)) THEN
        CALL ppmg_argfail(caller, 'wx1_user, wx2_user and wx3_user go together', info)
        RETURN
      END IF
    ELSE
      IF (PRESENT(wx1_user) .NEQV. PRESENT(wx2_user)) THEN
        CALL ppmg_argfail(caller, 'a wx_user array is missing', info)
        RETURN
      END IF
    END IF
    IF (Np < 0) THEN
      CALL ppmg_argfail(caller, 'particles must be specified', info)
      RETURN
    END IF
    IF (Np == 0) RETURN
    IF (SIZE(xp, 1) < dim .OR. SIZE(xp, 2) < Np) THEN
      CALL ppmg_argfail(caller, 'xp too small', info)
      RETURN
    END IF
    ks = ppmg_kernel_support(kernel)
    ! weights (1:2ks, 1:nsublist, 1-ks:Np+ks) as ppm_rmsh_comp_weights.f:668-674
    user = .FALSE.
    IF (PRESENT(wx1_user)) user = ASSOCIATED(wx1_user) .AND. ASSOCIATED(wx2_user)
    IF (user .AND. dim == 3) user = ASSOCIATED(wx3_user)
    IF (user) THEN
      w1 => wx1_user
      w2 => wx2_user
      w3 => wx1_user
      IF (dim == 3) w3 => wx3_user
    ELSE
      CALL weights_fit_d(rmsh_wx1_d, ks, Np)
      CALL weights_fit_d(rmsh_wx2_d, ks, Np)
      CALL weights_fit_d(rmsh_wx3_d, ks, Np)
      w1 => rmsh_wx1_d
      w2 => rmsh_wx2_d
      w3 => rmsh_wx3_d
      IF (PRESENT(wx1_user)) THEN
        wx1_user => rmsh_wx1_d
        wx2_user => rmsh_wx2_d
        IF (dim == 3) wx3_user => rmsh_wx3_d
      END IF
    END IF
    IF (.NOT. weights_hold(LBOUND(w1), UBOUND(w1), ks, Np) .OR. .NOT. weights_hold(LBOUND(w2), UBOUND(w2), ks, Np) &
        .OR. .NOT. weights_hold(LBOUND(w3), UBOUND(w3), ks, Np)) THEN
      CALL ppmg_argfail(caller, 'weight arrays too small', info)
      RETURN
    END IF
    n = Np
    CALL cw_core_d(xp(1:dim, 1:n), w1(1:2*ks, 1, 1:n), w2(1:2*ks, 1, 1:n), w3(1:2*ks, 1, 1:n), &
                    Np, topo_id, mesh_id, kernel, info)
  END SUBROUTINE ppm_rmsh_comp_weights_d

  SUBROUTINE cw_core_d(xp, w1, w2, w3, np, topo_id, mesh_id, kernel, info)
    REAL(C_DOUBLE), TARGET, INTENT(IN) :: xp(*)
    REAL(C_DOUBLE), TARGET :: w1(*), w2(*), w3(*)
    INTEGER, INTENT(IN) :: np, topo_id, mesh_id, kernel
    INTEGER, INTENT(OUT) :: info
    info = ppmg_rmsh_comp_weights_reg(C_LOC(xp(1)), INT(np, C_INT), INT(topo_id, C_INT), INT(mesh_id, C_INT), &
                                      INT(kernel, C_INT), C_LOC(w1(1)), C_LOC(w2(1)), C_LOC(w3(1)), ppmg_prec_f64, &
                                          ppmg_mem_host)
    CALL ppmg_report('ppm_rmsh_comp_weights', info)
  END SUBROUTINE cw_core_d

  SUBROUTINE weights_fit_d(w, ks, np)
    REAL(ppm_kind_double), DIMENSION(:,:,:), POINTER :: w
    INTEGER, INTENT(IN) :: ks, np
    IF (ASSOCIATED(w)) THEN
      IF (SIZE(w, 1) == 2 * ks .AND. SIZE(w, 2) == 1 .AND. LBOUND(w, 3) == 1 - ks .AND. UBOUND(w, 3) == np + ks) RETURN
      DEALLOCATE(w)
    END IF
    ALLOCATE(w(1:2 * ks, 1:1, 1 - ks:np + ks))
    w = 0.0_ppm_kind_double
  END SUBROUTINE weights_fit_d

  ! Weights remesh uses: the caller's (all associated) or those comp_weights kept
  ! (ppm_rmsh_remesh.f:425-490).
  SUBROUTINE remesh_weights_d(caller, dim, kernel, np, w1, w2, w3, info, wx1_user, wx2_user, wx3_user)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    INTEGER, INTENT(IN) :: dim, kernel, np
    REAL(ppm_kind_double), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER, INTENT(OUT) :: info
    REAL(ppm_kind_double), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    INTEGER :: ks
    info = 0
    IF (kernel /= ppm_param_rmsh_kernel_bsp2 .AND. kernel /= ppm_param_rmsh_kernel_mp4) THEN
      CALL ppmg_argfail(caller, 'wrong kernel definition', info)
      RETURN
    END IF
    ks = ppmg_kernel_support(kernel)
    IF (PRESENT(wx1_user)) THEN
      IF (.NOT. PRESENT(wx2_user) .OR. (dim == 3 .AND. .NOT. PRESENT(wx3_user))) THEN
        CALL ppmg_argfail(caller, 'a wx_user array is missing', info)
        RETURN
      END IF
      IF (.NOT. ASSOCIATED(wx1_user) .OR. .NOT. ASSOCIATED(wx2_user)) THEN
        CALL ppmg_argfail(caller, 'cannot remesh with empty weights', info)
        RETURN
      END IF
      w1 => wx1_user
      w2 => wx2_user
      w3 => wx1_user
      IF (dim == 3) THEN
        IF (.NOT. ASSOCIATED(wx3_user)) THEN
          CALL ppmg_argfail(caller, 'cannot remesh with empty weights', info)
          RETURN
        END IF
        w3 => wx3_user
      END IF
    ELSE
      IF (.NOT. ASSOCIATED(rmsh_wx1_d)) THEN
        CALL ppmg_argfail(caller, 'compute weights before remeshing', info)
        RETURN
      END IF
      w1 => rmsh_wx1_d
      w2 => rmsh_wx2_d
      w3 => rmsh_wx3_d
    END IF
    IF (np <= 0) RETURN
    IF (.NOT. weights_hold(LBOUND(w1), UBOUND(w1), ks, np) .OR. .NOT. weights_hold(LBOUND(w2), UBOUND(w2), ks, np) &
        .OR. .NOT. weights_hold(LBOUND(w3), UBOUND(w3), ks, np)) THEN
      CALL ppmg_argfail(caller, 'weights do not cover the particles (kernel or Np changed since comp_weights?)', info)
      RETURN
    END IF
  END SUBROUTINE remesh_weights_d
END MODULE ppm_module_rmsh_comp_weights


MODULE ppm_module_rmsh_remesh
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  USE ppmg_fcore
  USE ppm_module_data_rmsh
  USE ppm_module_rmsh_comp_weights
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: ppm_rmsh_remesh
  INTERFACE ppm_rmsh_remesh
    MODULE PROCEDURE ppm_rmsh_remesh_ss_2d
    MODULE PROCEDURE ppm_rmsh_remesh_ds_2d
    MODULE PROCEDURE ppm_rmsh_remesh_sv_2d
    MODULE PROCEDURE ppm_rmsh_remesh_dv_2d
    MODULE PROCEDURE ppm_rmsh_remesh_ss_3d
    MODULE PROCEDURE ppm_rmsh_remesh_ds_3d
    MODULE PROCEDURE ppm_rmsh_remesh_sv_3d
    MODULE PROCEDURE ppm_rmsh_remesh_dv_3d
  END INTERFACE
CONTAINS

  SUBROUTINE ppm_rmsh_remesh_ss_2d(xp, Np, up, topo_id, mesh_id, kernel, ghostsize, info, field_up, wx1_user, &
      wx2_user, wx3_user)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:), POINTER :: up
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    REAL(MK), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    REAL(MK), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER :: lo(3), hi(3), n, ks
    INTEGER(C_INT) :: cg(3)
    INTEGER, PARAMETER :: lda = 1
    CALL ppmg_box('ppm_rmsh_remesh', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    CALL remesh_weights_s('ppm_rmsh_remesh', 2, kernel, Np, w1, w2, w3, info, wx1_user, wx2_user, wx3_user)
    IF (info /= 0) RETURN
    CALL ppmg_part_check('ppm_rmsh_remesh', Np, SHAPE(xp), 2, SHAPE(up), 0, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(field_up)) THEN
      CALL ppmg_field_check('ppm_rmsh_remesh', LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 2, lo, hi, info)
      IF (info /= 0) THEN
        DEALLOCATE(field_up)
        NULLIFY(field_up)
      END IF
    END IF
    IF (.NOT. ASSOCIATED(field_up)) THEN
      ALLOCATE(field_up(lo(1):hi(1), lo(2):hi(2), 1:1), STAT=info)
      IF (info /= 0) THEN
        info = ppm_error_fatal
        RETURN
      END IF
    END IF
    field_up = 0.0_MK
    IF (Np == 0) RETURN
    n = Np
    ks = SIZE(w1, 1)
    CALL remesh_core_s(xp(1:2, 1:n), up(1:n), field_up(lo(1):hi(1), lo(2):hi(2), 1), &
                       w1(1:ks, 1:n, 1), w2(1:ks, 1:n, 1), w1(1:ks, 1:n, 1), &
                       Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_rmsh_remesh_ss_2d

  SUBROUTINE ppm_rmsh_remesh_ds_2d(xp, Np, up, topo_id, mesh_id, kernel, ghostsize, info, field_up, wx1_user, &
      wx2_user, wx3_user)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:), POINTER :: up
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    REAL(MK), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    REAL(MK), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER :: lo(3), hi(3), n, ks
    INTEGER(C_INT) :: cg(3)
    INTEGER, PARAMETER :: lda = 1
    CALL ppmg_box('ppm_rmsh_remesh', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    CALL remesh_weights_d('ppm_rmsh_remesh', 2, kernel, Np, w1, w2, w3, info, wx1_user, wx2_user, wx3_user)
    IF (info /= 0) RETURN
    CALL ppmg_part_check('ppm_rmsh_remesh', Np, SHAPE(xp), 2, SHAPE(up), 0, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(field_up)) THEN
      CALL ppmg_field_check('ppm_rmsh_remesh', LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 2, lo, hi, info)
      IF (info /= 0) THEN
        DEALLOCATE(field_up)
        NULLIFY(field_up)
      END IF
    END IF
    IF (.NOT. ASSOCIATED(field_up)) THEN
      ALLOCATE(field_up(lo(1):hi(1), lo(2):hi(2), 1:1), STAT=info)
      IF (info /= 0) THEN
        info = ppm_error_fatal
        RETURN
      END IF
    END IF
    field_up = 0.0_MK
    IF (Np == 0) RETURN
    n = Np
    ks = SIZE(w1, 1)
    CALL remesh_core_d(xp(1:2, 1:n), up(1:n), field_up(lo(1):hi(1), lo(2):hi(2), 1), &
                       w1(1:ks, 1:n, 1), w2(1:ks, 1:n, 1), w1(1:ks, 1:n, 1), &
                       Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_rmsh_remesh_ds_2d

  SUBROUTINE ppm_rmsh_remesh_sv_2d(xp, Np, up, lda, topo_id, mesh_id, kernel, ghostsize, info, field_up, wx1_user, &
      wx2_user, wx3_user)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    REAL(MK), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    REAL(MK), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER :: lo(3), hi(3), n, ks
    INTEGER(C_INT) :: cg(3)
    CALL ppmg_box('ppm_rmsh_remesh', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    CALL remesh_weights_s('ppm_rmsh_remesh', 2, kernel, Np, w1, w2, w3, info, wx1_user, wx2_user, wx3_user)
    IF (info /= 0) RETURN
    CALL ppmg_part_check('ppm_rmsh_remesh', Np, SHAPE(xp), 2, SHAPE(up), lda, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(field_up)) THEN
      CALL ppmg_field_check('ppm_rmsh_remesh', LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 2, lo, hi, info)
      IF (info /= 0) THEN
        DEALLOCATE(field_up)
        NULLIFY(field_up)
      END IF
    END IF
    IF (.NOT. ASSOCIATED(field_up)) THEN
      ALLOCATE(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1:1), STAT=info)
      IF (info /= 0) THEN
        info = ppm_error_fatal
        RETURN
      END IF
    END IF
    field_up = 0.0_MK
    IF (Np == 0) RETURN
    n = Np
    ks = SIZE(w1, 1)
    CALL remesh_core_s(xp(1:2, 1:n), up(1:lda, 1:n), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
                       w1(1:ks, 1:n, 1), w2(1:ks, 1:n, 1), w1(1:ks, 1:n, 1), &
                       Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_rmsh_remesh_sv_2d

  SUBROUTINE ppm_rmsh_remesh_dv_2d(xp, Np, up, lda, topo_id, mesh_id, kernel, ghostsize, info, field_up, wx1_user, &
      wx2_user, wx3_user)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    REAL(MK), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    REAL(MK), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER :: lo(3), hi(3), n, ks
    INTEGER(C_INT) :: cg(3)
    CALL ppmg_box('ppm_rmsh_remesh', topo_id, ghostsize, 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    CALL remesh_weights_d('ppm_rmsh_remesh', 2, kernel, Np, w1, w2, w3, info, wx1_user, wx2_user, wx3_user)
    IF (info /= 0) RETURN
    CALL ppmg_part_check('ppm_rmsh_remesh', Np, SHAPE(xp), 2, SHAPE(up), lda, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(field_up)) THEN
      CALL ppmg_field_check('ppm_rmsh_remesh', LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 2, lo, hi, info)
      IF (info /= 0) THEN
        DEALLOCATE(field_up)
        NULLIFY(field_up)
      END IF
    END IF
    IF (.NOT. ASSOCIATED(field_up)) THEN
      ALLOCATE(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1:1), STAT=info)
      IF (info /= 0) THEN
        info = ppm_error_fatal
        RETURN
      END IF
    END IF
    field_up = 0.0_MK
    IF (Np == 0) RETURN
    n = Np
    ks = SIZE(w1, 1)
    CALL remesh_core_d(xp(1:2, 1:n), up(1:lda, 1:n), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
                       w1(1:ks, 1:n, 1), w2(1:ks, 1:n, 1), w1(1:ks, 1:n, 1), &
                       Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_rmsh_remesh_dv_2d

  SUBROUTINE ppm_rmsh_remesh_ss_3d(xp, Np, up, topo_id, mesh_id, kernel, ghostsize, info, field_up, wx1_user, &
      wx2_user, wx3_user)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:), POINTER :: up
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    REAL(MK), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    REAL(MK), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER :: lo(3), hi(3), n, ks
    INTEGER(C_INT) :: cg(3)
    INTEGER, PARAMETER :: lda = 1
    CALL ppmg_box('ppm_rmsh_remesh', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    CALL remesh_weights_s('ppm_rmsh_remesh', 3, kernel, Np, w1, w2, w3, info, wx1_user, wx2_user, wx3_user)
    IF (info /= 0) RETURN
    CALL ppmg_part_check('ppm_rmsh_remesh', Np, SHAPE(xp), 3, SHAPE(up), 0, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(field_up)) THEN
      CALL ppmg_field_check('ppm_rmsh_remesh', LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 3, lo, hi, info)
      IF (info /= 0) THEN
        DEALLOCATE(field_up)
        NULLIFY(field_up)
      END IF
    END IF
    IF (.NOT. ASSOCIATED(field_up)) THEN
      ALLOCATE(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1:1), STAT=info)
      IF (info /= 0) THEN
        info = ppm_error_fatal
        RETURN
      END IF
    END IF
    field_up = 0.0_MK
    IF (Np == 0) RETURN
    n = Np
    ks = SIZE(w1, 1)
    CALL remesh_core_s(xp(1:3, 1:n), up(1:n), field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), &
                       w1(1:ks, 1:n, 1), w2(1:ks, 1:n, 1), w3(1:ks, 1:n, 1), &
                       Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_rmsh_remesh_ss_3d

  SUBROUTINE ppm_rmsh_remesh_ds_3d(xp, Np, up, topo_id, mesh_id, kernel, ghostsize, info, field_up, wx1_user, &
      wx2_user, wx3_user)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:), POINTER :: up
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    REAL(MK), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    REAL(MK), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER :: lo(3), hi(3), n, ks
    INTEGER(C_INT) :: cg(3)
    INTEGER, PARAMETER :: lda = 1
    CALL ppmg_box('ppm_rmsh_remesh', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    CALL remesh_weights_d('ppm_rmsh_remesh', 3, kernel, Np, w1, w2, w3, info, wx1_user, wx2_user, wx3_user)
    IF (info /= 0) RETURN
    CALL ppmg_part_check('ppm_rmsh_remesh', Np, SHAPE(xp), 3, SHAPE(up), 0, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(field_up)) THEN
      CALL ppmg_field_check('ppm_rmsh_remesh', LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 3, lo, hi, info)
      IF (info /= 0) THEN
        DEALLOCATE(field_up)
        NULLIFY(field_up)
      END IF
    END IF
    IF (.NOT. ASSOCIATED(field_up)) THEN
      ALLOCATE(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1:1), STAT=info)
      IF (info /= 0) THEN
        info = ppm_error_fatal
        RETURN
      END IF
    END IF
    field_up = 0.0_MK
    IF (Np == 0) RETURN
    n = Np
    ks = SIZE(w1, 1)
    CALL remesh_core_d(xp(1:3, 1:n), up(1:n), field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), &
                       w1(1:ks, 1:n, 1), w2(1:ks, 1:n, 1), w3(1:ks, 1:n, 1), &
                       Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_rmsh_remesh_ds_3d

  SUBROUTINE ppm_rmsh_remesh_sv_3d(xp, Np, up, lda, topo_id, mesh_id, kernel, ghostsize, info, field_up, wx1_user, &
      wx2_user, wx3_user)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    REAL(MK), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    REAL(MK), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER :: lo(3), hi(3), n, ks
    INTEGER(C_INT) :: cg(3)
    CALL ppmg_box('ppm_rmsh_remesh', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    CALL remesh_weights_s('ppm_rmsh_remesh', 3, kernel, Np, w1, w2, w3, info, wx1_user, wx2_user, wx3_user)
    IF (info /= 0) RETURN
    CALL ppmg_part_check('ppm_rmsh_remesh', Np, SHAPE(xp), 3, SHAPE(up), lda, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(field_up)) THEN
      CALL ppmg_field_check('ppm_rmsh_remesh', LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 3, lo, hi, info)
      IF (info /= 0) THEN
        DEALLOCATE(field_up)
        NULLIFY(field_up)
      END IF
    END IF
    IF (.NOT. ASSOCIATED(field_up)) THEN
      ALLOCATE(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1:1), STAT=info)
      IF (info /= 0) THEN
        info = ppm_error_fatal
        RETURN
      END IF
    END IF
    field_up = 0.0_MK
    IF (Np == 0) RETURN
    n = Np
    ks = SIZE(w1, 1)
    CALL remesh_core_s(xp(1:3, 1:n), up(1:lda, 1:n), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), &
                       w1(1:ks, 1:n, 1), w2(1:ks, 1:n, 1), w3(1:ks, 1:n, 1), &
                       Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_rmsh_remesh_sv_3d

  SUBROUTINE ppm_rmsh_remesh_dv_3d(xp, Np, up, lda, topo_id, mesh_id, kernel, ghostsize, info, field_up, wx1_user, &
      wx2_user, wx3_user)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    INTEGER, INTENT(IN) :: Np, topo_id, mesh_id, kernel
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(OUT) :: info
    REAL(MK), OPTIONAL, DIMENSION(:,:,:), POINTER :: wx1_user, wx2_user, wx3_user
    REAL(MK), DIMENSION(:,:,:), POINTER :: w1, w2, w3
    INTEGER :: lo(3), hi(3), n, ks
    INTEGER(C_INT) :: cg(3)
    CALL ppmg_box('ppm_rmsh_remesh', topo_id, ghostsize, 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    CALL remesh_weights_d('ppm_rmsh_remesh', 3, kernel, Np, w1, w2, w3, info, wx1_user, wx2_user, wx3_user)
    IF (info /= 0) RETURN
    CALL ppmg_part_check('ppm_rmsh_remesh', Np, SHAPE(xp), 3, SHAPE(up), lda, info)
    IF (info /= 0) RETURN
    IF (ASSOCIATED(field_up)) THEN
      CALL ppmg_field_check('ppm_rmsh_remesh', LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 3, lo, hi, info)
      IF (info /= 0) THEN
        DEALLOCATE(field_up)
        NULLIFY(field_up)
      END IF
    END IF
    IF (.NOT. ASSOCIATED(field_up)) THEN
      ALLOCATE(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1:1), STAT=info)
      IF (info /= 0) THEN
        info = ppm_error_fatal
        RETURN
      END IF
    END IF
    field_up = 0.0_MK
    IF (Np == 0) RETURN
    n = Np
    ks = SIZE(w1, 1)
    CALL remesh_core_d(xp(1:3, 1:n), up(1:lda, 1:n), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), &
                       w1(1:ks, 1:n, 1), w2(1:ks, 1:n, 1), w3(1:ks, 1:n, 1), &
                       Np, lda, topo_id, mesh_id, kernel, cg, info)
  END SUBROUTINE ppm_rmsh_remesh_dv_3d

END MODULE ppm_module_rmsh_remesh


MODULE ppm_module_rmsh_create_part
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  USE ppmg_fcore
  IMPLICIT NONE
  PRIVATE
  PUBLIC :: ppm_rmsh_create_part
  INTERFACE ppm_rmsh_create_part
    MODULE PROCEDURE ppm_rmsh_create_part_sss_2d
    MODULE PROCEDURE ppm_rmsh_create_part_ssv_2d
    MODULE PROCEDURE ppm_rmsh_create_part_svs_2d
    MODULE PROCEDURE ppm_rmsh_create_part_svv_2d
    MODULE PROCEDURE ppm_rmsh_create_part_dss_2d
    MODULE PROCEDURE ppm_rmsh_create_part_dsv_2d
    MODULE PROCEDURE ppm_rmsh_create_part_dvs_2d
    MODULE PROCEDURE ppm_rmsh_create_part_dvv_2d
    MODULE PROCEDURE ppm_rmsh_create_part_sss_3d
    MODULE PROCEDURE ppm_rmsh_create_part_ssv_3d
    MODULE PROCEDURE ppm_rmsh_create_part_svs_3d
    MODULE PROCEDURE ppm_rmsh_create_part_svv_3d
    MODULE PROCEDURE ppm_rmsh_create_part_dss_3d
    MODULE PROCEDURE ppm_rmsh_create_part_dsv_3d
    MODULE PROCEDURE ppm_rmsh_create_part_dvs_3d
    MODULE PROCEDURE ppm_rmsh_create_part_dvv_3d
  END INTERFACE
CONTAINS

  SUBROUTINE cp_core_s(field, fieldw, slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, maxnp, n, &
                       pxp, pup, pwp, info)
    REAL(C_FLOAT), TARGET, INTENT(IN) :: field(*), fieldw(*)
    LOGICAL, INTENT(IN) :: slave, reset
    INTEGER, INTENT(IN) :: lda, lda2, topo_id, mesh_id, maxnp
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    REAL(C_DOUBLE), INTENT(IN) :: cut(2)
    TYPE(C_PTR), INTENT(IN) :: cwp, pxp, pup, pwp
    INTEGER, INTENT(OUT) :: n, info
    INTEGER(C_INT) :: cn, rs
    TYPE(C_PTR) :: pfw
    REAL(C_DOUBLE) :: c2(2)
    pfw = C_NULL_PTR
    IF (slave) pfw = C_LOC(fieldw(1))
    rs = 0
    IF (reset) rs = 1
    c2 = cut
    info = ppmg_rmsh_create_part_reg(C_LOC(field(1)), INT(lda, C_INT), INT(topo_id, C_INT), INT(mesh_id, C_INT), cg, &
                                     c2, cwp, rs, INT(maxnp, C_INT), cn, pxp, pup, pfw, INT(lda2, C_INT), pwp, &
                                     ppmg_prec_f32, ppmg_mem_host)
    n = cn
    CALL ppmg_report('ppm_rmsh_create_part', info)
  END SUBROUTINE cp_core_s


  SUBROUTINE cp_core_d(field, fieldw, slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, maxnp, n, &
                       pxp, pup, pwp, info)
    REAL(C_DOUBLE), TARGET, INTENT(IN) :: field(*), fieldw(*)
    LOGICAL, INTENT(IN) :: slave, reset
    INTEGER, INTENT(IN) :: lda, lda2, topo_id, mesh_id, maxnp
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    REAL(C_DOUBLE), INTENT(IN) :: cut(2)
    TYPE(C_PTR), INTENT(IN) :: cwp, pxp, pup, pwp
    INTEGER, INTENT(OUT) :: n, info
    INTEGER(C_INT) :: cn, rs
    TYPE(C_PTR) :: pfw
    REAL(C_DOUBLE) :: c2(2)
    pfw = C_NULL_PTR
    IF (slave) pfw = C_LOC(fieldw(1))
    rs = 0
    IF (reset) rs = 1
    c2 = cut
    info = ppmg_rmsh_create_part_reg(C_LOC(field(1)), INT(lda, C_INT), INT(topo_id, C_INT), INT(mesh_id, C_INT), cg, &
                                     c2, cwp, rs, INT(maxnp, C_INT), cn, pxp, pup, pfw, INT(lda2, C_INT), pwp, &
                                     ppmg_prec_f64, ppmg_mem_host)
    n = cn
    CALL ppmg_report('ppm_rmsh_create_part', info)
  END SUBROUTINE cp_core_d


  SUBROUTINE ppm_rmsh_create_part_sss_2d(xp, np, up, field_up, topo_id, mesh_id, cutoff, info, resetpos, field_wp, wp)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:), POINTER :: up
    REAL(MK), DIMENSION(:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), DIMENSION(:,:,:), POINTER, OPTIONAL :: field_wp
    REAL(MK), DIMENSION(:), POINTER, OPTIONAL :: wp
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda = 1
    INTEGER, PARAMETER :: lda2 = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = PRESENT(field_wp)
    CALL ppmg_box(caller, topo_id, (/0, 0/), 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 1)
    lo(1) = LBOUND(field_up, 1)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 2)
    lo(2) = LBOUND(field_up, 2)
    hi(2) = hi(2) + cg(2)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .FALSE., lda2, 2, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (slave) THEN
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_wp(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_up(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(2, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1))
    IF (slave) pwp = C_LOC(wp(1))
    IF (slave) THEN
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_wp(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_up(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_sss_2d

  SUBROUTINE ppm_rmsh_create_part_ssv_2d(xp, np, up, field_up, topo_id, mesh_id, cutoff, info, resetpos, field_wp, wp, &
      lda2)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:), POINTER :: up
    REAL(MK), DIMENSION(:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_wp
    REAL(MK), DIMENSION(:,:), POINTER :: wp
    INTEGER, INTENT(IN) :: lda2
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = .TRUE.
    CALL ppmg_box(caller, topo_id, (/0, 0/), 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 1)
    lo(1) = LBOUND(field_up, 1)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 2)
    lo(2) = LBOUND(field_up, 2)
    hi(2) = hi(2) + cg(2)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .TRUE., lda2, 2, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (slave) THEN
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), 1), slave, lda, &
          lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_up(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(lda2, n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(2, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1))
    IF (slave) pwp = C_LOC(wp(1, 1))
    IF (slave) THEN
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), 1), slave, lda, &
          lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_up(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_ssv_2d

  SUBROUTINE ppm_rmsh_create_part_svs_2d(xp, np, up, lda, field_up, topo_id, mesh_id, cutoff, info, resetpos, &
      cutoff_weights, field_wp, wp)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), OPTIONAL, DIMENSION(:), POINTER :: cutoff_weights
    REAL(MK), DIMENSION(:,:,:), POINTER, OPTIONAL :: field_wp
    REAL(MK), DIMENSION(:), POINTER, OPTIONAL :: wp
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda2 = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = PRESENT(field_wp)
    CALL ppmg_box(caller, topo_id, (/0, 0/), 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 2)
    lo(1) = LBOUND(field_up, 2)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 3)
    lo(2) = LBOUND(field_up, 3)
    hi(2) = hi(2) + cg(2)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .FALSE., lda2, 2, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (PRESENT(cutoff_weights)) THEN
      IF (.NOT. ASSOCIATED(cutoff_weights)) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights is not allocated', info)
        RETURN
      END IF
      IF (SIZE(cutoff_weights) < lda) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights too short', info)
        RETURN
      END IF
      IF (ANY(cutoff_weights(1:lda) < 0.0_MK) .OR. ANY(cutoff_weights(1:lda) > 1.0_MK)) THEN
        CALL ppmg_argfail(caller, '0 <= cutoff_weights <= 1 is violated.', info)
        RETURN
      END IF
      ALLOCATE(cw(lda))
      cw = REAL(cutoff_weights(1:lda), C_DOUBLE)
      cwp = C_LOC(cw(1))
    END IF
    IF (slave) THEN
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_wp(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, &
          lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(lda, n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(2, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1, 1))
    IF (slave) pwp = C_LOC(wp(1))
    IF (slave) THEN
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_wp(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, &
          lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_svs_2d

  SUBROUTINE ppm_rmsh_create_part_svv_2d(xp, np, up, lda, field_up, topo_id, mesh_id, cutoff, info, resetpos, &
      cutoff_weights, field_wp, wp, lda2)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), OPTIONAL, DIMENSION(:), POINTER :: cutoff_weights
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_wp
    REAL(MK), DIMENSION(:,:), POINTER :: wp
    INTEGER, INTENT(IN) :: lda2
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = .TRUE.
    CALL ppmg_box(caller, topo_id, (/0, 0/), 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 2)
    lo(1) = LBOUND(field_up, 2)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 3)
    lo(2) = LBOUND(field_up, 3)
    hi(2) = hi(2) + cg(2)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .TRUE., lda2, 2, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (PRESENT(cutoff_weights)) THEN
      IF (.NOT. ASSOCIATED(cutoff_weights)) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights is not allocated', info)
        RETURN
      END IF
      IF (SIZE(cutoff_weights) < lda) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights too short', info)
        RETURN
      END IF
      IF (ANY(cutoff_weights(1:lda) < 0.0_MK) .OR. ANY(cutoff_weights(1:lda) > 1.0_MK)) THEN
        CALL ppmg_argfail(caller, '0 <= cutoff_weights <= 1 is violated.', info)
        RETURN
      END IF
      ALLOCATE(cw(lda))
      cw = REAL(cutoff_weights(1:lda), C_DOUBLE)
      cwp = C_LOC(cw(1))
    END IF
    IF (slave) THEN
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(lda, n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(lda2, n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(2, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1, 1))
    IF (slave) pwp = C_LOC(wp(1, 1))
    IF (slave) THEN
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_svv_2d

  SUBROUTINE ppm_rmsh_create_part_dss_2d(xp, np, up, field_up, topo_id, mesh_id, cutoff, info, resetpos, field_wp, wp)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:), POINTER :: up
    REAL(MK), DIMENSION(:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), DIMENSION(:,:,:), POINTER, OPTIONAL :: field_wp
    REAL(MK), DIMENSION(:), POINTER, OPTIONAL :: wp
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda = 1
    INTEGER, PARAMETER :: lda2 = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = PRESENT(field_wp)
    CALL ppmg_box(caller, topo_id, (/0, 0/), 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 1)
    lo(1) = LBOUND(field_up, 1)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 2)
    lo(2) = LBOUND(field_up, 2)
    hi(2) = hi(2) + cg(2)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .FALSE., lda2, 2, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (slave) THEN
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_wp(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_up(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(2, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1))
    IF (slave) pwp = C_LOC(wp(1))
    IF (slave) THEN
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_wp(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_up(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_dss_2d

  SUBROUTINE ppm_rmsh_create_part_dsv_2d(xp, np, up, field_up, topo_id, mesh_id, cutoff, info, resetpos, field_wp, wp, &
      lda2)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:), POINTER :: up
    REAL(MK), DIMENSION(:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_wp
    REAL(MK), DIMENSION(:,:), POINTER :: wp
    INTEGER, INTENT(IN) :: lda2
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = .TRUE.
    CALL ppmg_box(caller, topo_id, (/0, 0/), 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 1)
    lo(1) = LBOUND(field_up, 1)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 2)
    lo(2) = LBOUND(field_up, 2)
    hi(2) = hi(2) + cg(2)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .TRUE., lda2, 2, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (slave) THEN
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), 1), slave, lda, &
          lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_up(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(lda2, n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(2, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1))
    IF (slave) pwp = C_LOC(wp(1, 1))
    IF (slave) THEN
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), 1), slave, lda, &
          lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), 1), field_up(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, lda2, &
          topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_dsv_2d

  SUBROUTINE ppm_rmsh_create_part_dvs_2d(xp, np, up, lda, field_up, topo_id, mesh_id, cutoff, info, resetpos, &
      cutoff_weights, field_wp, wp)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), OPTIONAL, DIMENSION(:), POINTER :: cutoff_weights
    REAL(MK), DIMENSION(:,:,:), POINTER, OPTIONAL :: field_wp
    REAL(MK), DIMENSION(:), POINTER, OPTIONAL :: wp
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda2 = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = PRESENT(field_wp)
    CALL ppmg_box(caller, topo_id, (/0, 0/), 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 2)
    lo(1) = LBOUND(field_up, 2)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 3)
    lo(2) = LBOUND(field_up, 3)
    hi(2) = hi(2) + cg(2)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .FALSE., lda2, 2, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (PRESENT(cutoff_weights)) THEN
      IF (.NOT. ASSOCIATED(cutoff_weights)) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights is not allocated', info)
        RETURN
      END IF
      IF (SIZE(cutoff_weights) < lda) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights too short', info)
        RETURN
      END IF
      IF (ANY(cutoff_weights(1:lda) < 0.0_MK) .OR. ANY(cutoff_weights(1:lda) > 1.0_MK)) THEN
        CALL ppmg_argfail(caller, '0 <= cutoff_weights <= 1 is violated.', info)
        RETURN
      END IF
      ALLOCATE(cw(lda))
      cw = REAL(cutoff_weights(1:lda), C_DOUBLE)
      cwp = C_LOC(cw(1))
    END IF
    IF (slave) THEN
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_wp(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, &
          lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(lda, n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(2, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1, 1))
    IF (slave) pwp = C_LOC(wp(1))
    IF (slave) THEN
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_wp(lo(1):hi(1), lo(2):hi(2), 1), slave, lda, &
          lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_dvs_2d

  SUBROUTINE ppm_rmsh_create_part_dvv_2d(xp, np, up, lda, field_up, topo_id, mesh_id, cutoff, info, resetpos, &
      cutoff_weights, field_wp, wp, lda2)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), OPTIONAL, DIMENSION(:), POINTER :: cutoff_weights
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_wp
    REAL(MK), DIMENSION(:,:), POINTER :: wp
    INTEGER, INTENT(IN) :: lda2
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = .TRUE.
    CALL ppmg_box(caller, topo_id, (/0, 0/), 2, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 2)
    lo(1) = LBOUND(field_up, 2)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 3)
    lo(2) = LBOUND(field_up, 3)
    hi(2) = hi(2) + cg(2)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 2, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .TRUE., lda2, 2, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (PRESENT(cutoff_weights)) THEN
      IF (.NOT. ASSOCIATED(cutoff_weights)) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights is not allocated', info)
        RETURN
      END IF
      IF (SIZE(cutoff_weights) < lda) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights too short', info)
        RETURN
      END IF
      IF (ANY(cutoff_weights(1:lda) < 0.0_MK) .OR. ANY(cutoff_weights(1:lda) > 1.0_MK)) THEN
        CALL ppmg_argfail(caller, '0 <= cutoff_weights <= 1 is violated.', info)
        RETURN
      END IF
      ALLOCATE(cw(lda))
      cw = REAL(cutoff_weights(1:lda), C_DOUBLE)
      cwp = C_LOC(cw(1))
    END IF
    IF (slave) THEN
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(lda, n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(lda2, n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(2, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1, 1))
    IF (slave) pwp = C_LOC(wp(1, 1))
    IF (slave) THEN
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), field_up(1:lda, lo(1):hi(1), lo(2):hi(2), 1), &
          slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_dvv_2d

  SUBROUTINE ppm_rmsh_create_part_sss_3d(xp, np, up, field_up, topo_id, mesh_id, cutoff, info, resetpos, field_wp, wp)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:), POINTER :: up
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), DIMENSION(:,:,:,:), POINTER, OPTIONAL :: field_wp
    REAL(MK), DIMENSION(:), POINTER, OPTIONAL :: wp
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda = 1
    INTEGER, PARAMETER :: lda2 = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = PRESENT(field_wp)
    CALL ppmg_box(caller, topo_id, (/0, 0, 0/), 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 1)
    lo(1) = LBOUND(field_up, 1)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 2)
    lo(2) = LBOUND(field_up, 2)
    hi(2) = hi(2) + cg(2)
    cg(3) = 1 - LBOUND(field_up, 3)
    lo(3) = LBOUND(field_up, 3)
    hi(3) = hi(3) + cg(3)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .FALSE., lda2, 3, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (slave) THEN
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(3, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1))
    IF (slave) pwp = C_LOC(wp(1))
    IF (slave) THEN
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_sss_3d

  SUBROUTINE ppm_rmsh_create_part_ssv_3d(xp, np, up, field_up, topo_id, mesh_id, cutoff, info, resetpos, field_wp, wp, &
      lda2)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:), POINTER :: up
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_wp
    REAL(MK), DIMENSION(:,:), POINTER :: wp
    INTEGER, INTENT(IN) :: lda2
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = .TRUE.
    CALL ppmg_box(caller, topo_id, (/0, 0, 0/), 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 1)
    lo(1) = LBOUND(field_up, 1)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 2)
    lo(2) = LBOUND(field_up, 2)
    hi(2) = hi(2) + cg(2)
    cg(3) = 1 - LBOUND(field_up, 3)
    lo(3) = LBOUND(field_up, 3)
    hi(3) = hi(3) + cg(3)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .TRUE., lda2, 3, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (slave) THEN
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(lda2, n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(3, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1))
    IF (slave) pwp = C_LOC(wp(1, 1))
    IF (slave) THEN
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_s(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_ssv_3d

  SUBROUTINE ppm_rmsh_create_part_svs_3d(xp, np, up, lda, field_up, topo_id, mesh_id, cutoff, info, resetpos, &
      cutoff_weights, field_wp, wp)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), OPTIONAL, DIMENSION(:), POINTER :: cutoff_weights
    REAL(MK), DIMENSION(:,:,:,:), POINTER, OPTIONAL :: field_wp
    REAL(MK), DIMENSION(:), POINTER, OPTIONAL :: wp
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda2 = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = PRESENT(field_wp)
    CALL ppmg_box(caller, topo_id, (/0, 0, 0/), 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 2)
    lo(1) = LBOUND(field_up, 2)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 3)
    lo(2) = LBOUND(field_up, 3)
    hi(2) = hi(2) + cg(2)
    cg(3) = 1 - LBOUND(field_up, 4)
    lo(3) = LBOUND(field_up, 4)
    hi(3) = hi(3) + cg(3)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .FALSE., lda2, 3, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (PRESENT(cutoff_weights)) THEN
      IF (.NOT. ASSOCIATED(cutoff_weights)) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights is not allocated', info)
        RETURN
      END IF
      IF (SIZE(cutoff_weights) < lda) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights too short', info)
        RETURN
      END IF
      IF (ANY(cutoff_weights(1:lda) < 0.0_MK) .OR. ANY(cutoff_weights(1:lda) > 1.0_MK)) THEN
        CALL ppmg_argfail(caller, '0 <= cutoff_weights <= 1 is violated.', info)
        RETURN
      END IF
      ALLOCATE(cw(lda))
      cw = REAL(cutoff_weights(1:lda), C_DOUBLE)
      cwp = C_LOC(cw(1))
    END IF
    IF (slave) THEN
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(1:lda, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(lda, n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(3, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1, 1))
    IF (slave) pwp = C_LOC(wp(1))
    IF (slave) THEN
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(1:lda, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_svs_3d

  SUBROUTINE ppm_rmsh_create_part_svv_3d(xp, np, up, lda, field_up, topo_id, mesh_id, cutoff, info, resetpos, &
      cutoff_weights, field_wp, wp, lda2)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), OPTIONAL, DIMENSION(:), POINTER :: cutoff_weights
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_wp
    REAL(MK), DIMENSION(:,:), POINTER :: wp
    INTEGER, INTENT(IN) :: lda2
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = .TRUE.
    CALL ppmg_box(caller, topo_id, (/0, 0, 0/), 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 2)
    lo(1) = LBOUND(field_up, 2)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 3)
    lo(2) = LBOUND(field_up, 3)
    hi(2) = hi(2) + cg(2)
    cg(3) = 1 - LBOUND(field_up, 4)
    lo(3) = LBOUND(field_up, 4)
    hi(3) = hi(3) + cg(3)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .TRUE., lda2, 3, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (PRESENT(cutoff_weights)) THEN
      IF (.NOT. ASSOCIATED(cutoff_weights)) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights is not allocated', info)
        RETURN
      END IF
      IF (SIZE(cutoff_weights) < lda) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights too short', info)
        RETURN
      END IF
      IF (ANY(cutoff_weights(1:lda) < 0.0_MK) .OR. ANY(cutoff_weights(1:lda) > 1.0_MK)) THEN
        CALL ppmg_argfail(caller, '0 <= cutoff_weights <= 1 is violated.', info)
        RETURN
      END IF
      ALLOCATE(cw(lda))
      cw = REAL(cutoff_weights(1:lda), C_DOUBLE)
      cwp = C_LOC(cw(1))
    END IF
    IF (slave) THEN
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(1:lda2, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(1:lda, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(lda, n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(lda2, n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(3, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1, 1))
    IF (slave) pwp = C_LOC(wp(1, 1))
    IF (slave) THEN
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(1:lda2, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_s(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(1:lda, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_svv_3d

  SUBROUTINE ppm_rmsh_create_part_dss_3d(xp, np, up, field_up, topo_id, mesh_id, cutoff, info, resetpos, field_wp, wp)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:), POINTER :: up
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), DIMENSION(:,:,:,:), POINTER, OPTIONAL :: field_wp
    REAL(MK), DIMENSION(:), POINTER, OPTIONAL :: wp
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda = 1
    INTEGER, PARAMETER :: lda2 = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = PRESENT(field_wp)
    CALL ppmg_box(caller, topo_id, (/0, 0, 0/), 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 1)
    lo(1) = LBOUND(field_up, 1)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 2)
    lo(2) = LBOUND(field_up, 2)
    hi(2) = hi(2) + cg(2)
    cg(3) = 1 - LBOUND(field_up, 3)
    lo(3) = LBOUND(field_up, 3)
    hi(3) = hi(3) + cg(3)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .FALSE., lda2, 3, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (slave) THEN
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(3, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1))
    IF (slave) pwp = C_LOC(wp(1))
    IF (slave) THEN
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_dss_3d

  SUBROUTINE ppm_rmsh_create_part_dsv_3d(xp, np, up, field_up, topo_id, mesh_id, cutoff, info, resetpos, field_wp, wp, &
      lda2)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:), POINTER :: up
    REAL(MK), DIMENSION(:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_wp
    REAL(MK), DIMENSION(:,:), POINTER :: wp
    INTEGER, INTENT(IN) :: lda2
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = .TRUE.
    CALL ppmg_box(caller, topo_id, (/0, 0, 0/), 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 1)
    lo(1) = LBOUND(field_up, 1)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 2)
    lo(2) = LBOUND(field_up, 2)
    hi(2) = hi(2) + cg(2)
    cg(3) = 1 - LBOUND(field_up, 3)
    lo(3) = LBOUND(field_up, 3)
    hi(3) = hi(3) + cg(3)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .FALSE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .TRUE., lda2, 3, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (slave) THEN
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(lda2, n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(3, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1))
    IF (slave) pwp = C_LOC(wp(1, 1))
    IF (slave) THEN
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(1:lda2, lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_d(field_up(lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_dsv_3d

  SUBROUTINE ppm_rmsh_create_part_dvs_3d(xp, np, up, lda, field_up, topo_id, mesh_id, cutoff, info, resetpos, &
      cutoff_weights, field_wp, wp)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), OPTIONAL, DIMENSION(:), POINTER :: cutoff_weights
    REAL(MK), DIMENSION(:,:,:,:), POINTER, OPTIONAL :: field_wp
    REAL(MK), DIMENSION(:), POINTER, OPTIONAL :: wp
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    INTEGER, PARAMETER :: lda2 = 1
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = PRESENT(field_wp)
    CALL ppmg_box(caller, topo_id, (/0, 0, 0/), 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 2)
    lo(1) = LBOUND(field_up, 2)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 3)
    lo(2) = LBOUND(field_up, 3)
    hi(2) = hi(2) + cg(2)
    cg(3) = 1 - LBOUND(field_up, 4)
    lo(3) = LBOUND(field_up, 4)
    hi(3) = hi(3) + cg(3)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .FALSE., lda2, 3, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (PRESENT(cutoff_weights)) THEN
      IF (.NOT. ASSOCIATED(cutoff_weights)) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights is not allocated', info)
        RETURN
      END IF
      IF (SIZE(cutoff_weights) < lda) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights too short', info)
        RETURN
      END IF
      IF (ANY(cutoff_weights(1:lda) < 0.0_MK) .OR. ANY(cutoff_weights(1:lda) > 1.0_MK)) THEN
        CALL ppmg_argfail(caller, '0 <= cutoff_weights <= 1 is violated.', info)
        RETURN
      END IF
      ALLOCATE(cw(lda))
      cw = REAL(cutoff_weights(1:lda), C_DOUBLE)
      cwp = C_LOC(cw(1))
    END IF
    IF (slave) THEN
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(1:lda, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(lda, n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(3, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1, 1))
    IF (slave) pwp = C_LOC(wp(1))
    IF (slave) THEN
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(lo(1):hi(1), lo(2):hi(2), &
          lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(1:lda, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_dvs_3d

  SUBROUTINE ppm_rmsh_create_part_dvv_3d(xp, np, up, lda, field_up, topo_id, mesh_id, cutoff, info, resetpos, &
      cutoff_weights, field_wp, wp, lda2)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(OUT) :: np
    REAL(MK), DIMENSION(:,:), POINTER :: up
    INTEGER, INTENT(IN) :: lda
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_up
    INTEGER, INTENT(IN) :: topo_id, mesh_id
    REAL(MK), DIMENSION(2), INTENT(IN) :: cutoff
    INTEGER, INTENT(OUT) :: info
    LOGICAL, OPTIONAL, INTENT(IN) :: resetpos
    REAL(MK), OPTIONAL, DIMENSION(:), POINTER :: cutoff_weights
    REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field_wp
    REAL(MK), DIMENSION(:,:), POINTER :: wp
    INTEGER, INTENT(IN) :: lda2
    INTEGER :: lo(3), hi(3), n, i
    INTEGER(C_INT) :: cg(3)
    LOGICAL :: reset, slave
    REAL(C_DOUBLE) :: cut(2)
    REAL(C_DOUBLE), DIMENSION(:), ALLOCATABLE, TARGET :: cw
    TYPE(C_PTR) :: cwp, pxp, pup, pwp
    CHARACTER(LEN=*), PARAMETER :: caller = 'ppm_rmsh_create_part'
    np = 0
    reset = .FALSE.
    IF (PRESENT(resetpos)) reset = resetpos
    slave = .TRUE.
    CALL ppmg_box(caller, topo_id, (/0, 0, 0/), 3, lo, hi, cg, info)
    IF (info /= 0) RETURN
    IF (.NOT. ASSOCIATED(field_up)) THEN
      CALL ppmg_argfail(caller, 'field_up is not associated', info)
      RETURN
    END IF
    cg(1) = 1 - LBOUND(field_up, 2)
    lo(1) = LBOUND(field_up, 2)
    hi(1) = hi(1) + cg(1)
    cg(2) = 1 - LBOUND(field_up, 3)
    lo(2) = LBOUND(field_up, 3)
    hi(2) = hi(2) + cg(2)
    cg(3) = 1 - LBOUND(field_up, 4)
    lo(3) = LBOUND(field_up, 4)
    hi(3) = hi(3) + cg(3)
    CALL ppmg_field_check(caller, LBOUND(field_up), UBOUND(field_up), .TRUE., lda, 3, lo, hi, info)
    IF (info /= 0) RETURN
    IF (slave) THEN
      IF (.NOT. ASSOCIATED(field_wp)) THEN
        CALL ppmg_argfail(caller, 'field_wp is not associated', info)
        RETURN
      END IF
      CALL ppmg_field_check(caller, LBOUND(field_wp), UBOUND(field_wp), .TRUE., lda2, 3, lo, hi, info)
      IF (info /= 0) RETURN
    END IF
    cut = REAL(cutoff, C_DOUBLE)
    cwp = C_NULL_PTR
    IF (PRESENT(cutoff_weights)) THEN
      IF (.NOT. ASSOCIATED(cutoff_weights)) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights is not allocated', info)
        RETURN
      END IF
      IF (SIZE(cutoff_weights) < lda) THEN
        CALL ppmg_argfail(caller, 'cutoff_weights too short', info)
        RETURN
      END IF
      IF (ANY(cutoff_weights(1:lda) < 0.0_MK) .OR. ANY(cutoff_weights(1:lda) > 1.0_MK)) THEN
        CALL ppmg_argfail(caller, '0 <= cutoff_weights <= 1 is violated.', info)
        RETURN
      END IF
      ALLOCATE(cw(lda))
      cw = REAL(cutoff_weights(1:lda), C_DOUBLE)
      cwp = C_LOC(cw(1))
    END IF
    IF (slave) THEN
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(1:lda2, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    ELSE
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(1:lda, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, 0, n, &
                     C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, info)
    END IF
    IF (info /= 0) RETURN
    IF (ASSOCIATED(up)) DEALLOCATE(up)
    ALLOCATE(up(lda, n))
    IF (slave) THEN
      IF (ASSOCIATED(wp)) DEALLOCATE(wp)
      ALLOCATE(wp(lda2, n))
    END IF
    IF (reset) THEN
      IF (ASSOCIATED(xp)) DEALLOCATE(xp)
      ALLOCATE(xp(3, n))
    END IF
    np = n
    IF (n == 0) RETURN
    pxp = C_NULL_PTR
    pwp = C_NULL_PTR
    IF (reset) pxp = C_LOC(xp(1, 1))
    pup = C_LOC(up(1, 1))
    IF (slave) pwp = C_LOC(wp(1, 1))
    IF (slave) THEN
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_wp(1:lda2, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    ELSE
      CALL cp_core_d(field_up(1:lda, lo(1):hi(1), lo(2):hi(2), lo(3):hi(3), 1), field_up(1:lda, lo(1):hi(1), &
          lo(2):hi(2), lo(3):hi(3), 1), slave, lda, lda2, topo_id, mesh_id, cg, cut, cwp, reset, n, i, &
                     pxp, pup, pwp, info)
    END IF
  END SUBROUTINE ppm_rmsh_create_part_dvv_3d

END MODULE ppm_module_rmsh_create_part

MODULE ppm_module_rmsh
  USE ppm_module_rmsh_comp_weights, ONLY: ppm_rmsh_comp_weights
  USE ppm_module_rmsh_remesh
  USE ppm_module_rmsh_create_part
END MODULE ppm_module_rmsh
