!-------------------------------------------------------------------------------
! ppm_module_ppmg -- ISO_C_BINDING interface of libppmg (include/ppmg.h) and
! the constants PPM clients take from ppm_param.h / ppm_module_data
! (ppm_param.h:183-407).  The drop-in modules (ppm_module_interp_p2m, ...)
! are thin wrappers over these bindings; arrays are passed as host memory and
! staged to the GPU by the library.
!-------------------------------------------------------------------------------
MODULE ppm_module_ppmg
  USE ISO_C_BINDING
  IMPLICIT NONE

  ! kinds and precision selectors (ppm_param.h)
  INTEGER, PARAMETER :: ppm_kind_single = KIND(1.0E0)
  INTEGER, PARAMETER :: ppm_kind_double = KIND(1.0D0)
  ! interpolation kernels (ppm_param.h:405-407)
  INTEGER, PARAMETER :: ppm_param_rmsh_kernel_bsp2 = 1
  INTEGER, PARAMETER :: ppm_param_rmsh_kernel_mp4  = 2
  INTEGER, PARAMETER :: ppm_param_rmsh_kernel_m3p6 = 3
  ! boundary conditions (ppm_param.h:333-338)
  INTEGER, PARAMETER :: ppm_param_bcdef_periodic  = 1
  INTEGER, PARAMETER :: ppm_param_bcdef_freespace = 2
  INTEGER, PARAMETER :: ppm_param_bcdef_symmetry  = 3
  INTEGER, PARAMETER :: ppm_param_bcdef_neumann   = 4
  INTEGER, PARAMETER :: ppm_param_bcdef_dirichlet = 5
  INTEGER, PARAMETER :: ppm_param_bcdef_robin     = 6
  ! error severities (ppm_param.h:372-375)
  INTEGER, PARAMETER :: ppm_error_notice  = -4
  INTEGER, PARAMETER :: ppm_error_warning = -3
  INTEGER, PARAMETER :: ppm_error_error   = -2
  INTEGER, PARAMETER :: ppm_error_fatal   = -1
  ! error codes (ppm_error.h:161-299)
  INTEGER, PARAMETER :: ppm_err_alloc      = 101
  INTEGER, PARAMETER :: ppm_err_part_lost  = 403
  INTEGER, PARAMETER :: ppm_err_part_unass = 405
  INTEGER, PARAMETER :: ppm_err_argument   = 503
  INTEGER, PARAMETER :: ppm_err_ppm_noinit = 903
  ! field mapping types (ppm_param.h:183-192)
  INTEGER, PARAMETER :: ppm_param_map_pop       = 1
  INTEGER, PARAMETER :: ppm_param_map_push      = 2
  INTEGER, PARAMETER :: ppm_param_map_send      = 4
  INTEGER, PARAMETER :: ppm_param_map_ghost_get = 5
  INTEGER, PARAMETER :: ppm_param_map_ghost_put = 6
  INTEGER, PARAMETER :: ppm_param_map_cancel    = 9
  INTEGER, PARAMETER :: ppm_param_map_init      = 10
  ! decomposition / assignment selectors accepted (and ignored: one sub) by ppm_mktopo
  INTEGER, PARAMETER :: ppm_param_decomp_cuboid = 5
  INTEGER, PARAMETER :: ppm_param_assign_internal = 1

  ! library-side selectors
  INTEGER(C_INT), PARAMETER :: ppmg_prec_f64 = 0, ppmg_prec_f32c = 1, ppmg_prec_f32 = 2
  INTEGER(C_INT), PARAMETER :: ppmg_mem_host = 1
  ! element kinds of ppmg_map_field_ghost_kind_reg (PPMG_KIND_* of ppmg.h)
  INTEGER(C_INT), PARAMETER :: ppmg_kind_d = 0, ppmg_kind_s = 2, ppmg_kind_i = 3, ppmg_kind_l = 4, &
                               ppmg_kind_dc = 5, ppmg_kind_sc = 6

  INTEGER :: ppm_debug = 0
  INTEGER :: ppm_dim = 0
  ! the topology the last ppm_mktopo made: the reference's current topology
  ! (ppm_topoid), which ppm_neighlist_clist works on
  INTEGER :: ppmg_current_topo = 0

  INTERFACE
    INTEGER(C_INT) FUNCTION ppmg_init(dim, prec, tolexp, debug) BIND(C, NAME='ppmg_init')
      IMPORT :: C_INT
      INTEGER(C_INT), VALUE :: dim, prec, tolexp, debug
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_finalize() BIND(C, NAME='ppmg_finalize')
      IMPORT :: C_INT
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_set_precision_mode(prec) BIND(C, NAME='ppmg_set_precision_mode')
      IMPORT :: C_INT
      INTEGER(C_INT), VALUE :: prec
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_mktopo_field(dim, min_phys, max_phys, bcdef, nm, ghostsize, topo_id, mesh_id) &
        BIND(C, NAME='ppmg_mktopo_field')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: dim
      REAL(C_DOUBLE) :: min_phys(*), max_phys(*)
      INTEGER(C_INT) :: bcdef(*), nm(*), ghostsize(*), topo_id, mesh_id
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_topo_info(topo_id, dim, nm, min_phys, max_phys, bcdef) BIND(C, NAME='ppmg_topo_info')
      IMPORT :: C_INT, C_DOUBLE
      INTEGER(C_INT), VALUE :: topo_id
      INTEGER(C_INT) :: dim, nm(*), bcdef(*)
      REAL(C_DOUBLE) :: min_phys(*), max_phys(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_interp_p2m(xp, np, up, lda, topo_id, mesh_id, kernel, ghostsize, field_up, prec, mem) &
        BIND(C, NAME='ppmg_interp_p2m')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: xp, up, field_up
      INTEGER(C_INT), VALUE :: np, lda, topo_id, mesh_id, kernel, prec, mem
      INTEGER(C_INT) :: ghostsize(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_interp_m2p(xp, np, up, lda, topo_id, mesh_id, kernel, ghostsize, field_up, prec, mem) &
        BIND(C, NAME='ppmg_interp_m2p')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: xp, up, field_up
      INTEGER(C_INT), VALUE :: np, lda, topo_id, mesh_id, kernel, prec, mem
      INTEGER(C_INT) :: ghostsize(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_rmsh_comp_weights_reg(xp, np, topo_id, mesh_id, kernel, wx1, wx2, wx3, prec, mem) &
        BIND(C, NAME='ppmg_rmsh_comp_weights_reg')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: xp, wx1, wx2, wx3
      INTEGER(C_INT), VALUE :: np, topo_id, mesh_id, kernel, prec, mem
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_rmsh_remesh_reg(xp, np, up, lda, topo_id, mesh_id, kernel, ghostsize, field_up, &
        wx1, wx2, wx3, prec, mem) BIND(C, NAME='ppmg_rmsh_remesh_reg')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: xp, up, field_up, wx1, wx2, wx3
      INTEGER(C_INT), VALUE :: np, lda, topo_id, mesh_id, kernel, prec, mem
      INTEGER(C_INT) :: ghostsize(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_rmsh_create_part_reg(field_up, lda, topo_id, mesh_id, ghostsize, cutoff, &
        cutoff_weights, resetpos, maxnp, np_out, xp, up, field_wp, lda2, wp, prec, mem) &
        BIND(C, NAME='ppmg_rmsh_create_part_reg')
      IMPORT :: C_INT, C_PTR, C_DOUBLE
      TYPE(C_PTR), VALUE :: field_up, cutoff_weights, xp, up, field_wp, wp
      INTEGER(C_INT), VALUE :: lda, topo_id, mesh_id, resetpos, maxnp, lda2, prec, mem
      INTEGER(C_INT) :: ghostsize(*), np_out
      REAL(C_DOUBLE) :: cutoff(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_map_field_ghost_reg(field_up, lda, topo_id, mesh_id, ghostsize, maptype, prec, mem) &
        BIND(C, NAME='ppmg_map_field_ghost_reg')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: field_up
      INTEGER(C_INT), VALUE :: lda, topo_id, mesh_id, maptype, prec, mem
      INTEGER(C_INT) :: ghostsize(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_map_field_ghost_kind_reg(field_up, lda, topo_id, mesh_id, ghostsize, maptype, &
                                                          kind, mask, mem) BIND(C, NAME='ppmg_map_field_ghost_kind_reg')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: field_up, mask
      INTEGER(C_INT), VALUE :: lda, topo_id, mesh_id, maptype, kind, mem
      INTEGER(C_INT) :: ghostsize(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_util_sort_reg(dim, xp, np, xmin, xmax, ncells, npbx, prec, mem) &
        BIND(C, NAME='ppmg_util_sort_reg')
      IMPORT :: C_INT, C_PTR
      INTEGER(C_INT), VALUE :: dim, np, prec, mem
      TYPE(C_PTR), VALUE :: xp, xmin, xmax, npbx
      INTEGER(C_INT) :: ncells(*)
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_util_rank_reg(dim, xp, np, xmin, xmax, ncells, ngl, lpdx, lhbx, icount, prec, mem) &
        BIND(C, NAME='ppmg_util_rank_reg')
      IMPORT :: C_INT, C_PTR
      INTEGER(C_INT), VALUE :: dim, np, prec, mem
      TYPE(C_PTR), VALUE :: xp, xmin, xmax, lpdx, lhbx
      INTEGER(C_INT) :: ncells(*), ngl(*), icount
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_impose_part_bc_reg(xp, np, topo_id, prec, mem) BIND(C, NAME='ppmg_impose_part_bc_reg')
      IMPORT :: C_INT, C_PTR
      TYPE(C_PTR), VALUE :: xp
      INTEGER(C_INT), VALUE :: np, topo_id, prec, mem
    END FUNCTION
    INTEGER(C_INT) FUNCTION ppmg_last_error(code, msg, msglen) BIND(C, NAME='ppmg_last_error')
      IMPORT :: C_INT, C_CHAR
      INTEGER(C_INT) :: code
      CHARACTER(KIND=C_CHAR) :: msg(*)
      INTEGER(C_INT), VALUE :: msglen
    END FUNCTION
  END INTERFACE

CONTAINS

  ! Report a non-zero info the way ppm_error does ("NNNN <SEVERITY> in <routine> : msg",
  ! ppm_error.f:125-169) when ppm_debug > 0; the library itself never aborts.
  SUBROUTINE ppmg_report(caller, info)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    INTEGER, INTENT(IN) :: info
    INTEGER(C_INT) :: code, sev
    CHARACTER(KIND=C_CHAR) :: buf(256)
    CHARACTER(LEN=256) :: msg
    INTEGER :: i
    IF (info == 0 .OR. ppm_debug <= 0) RETURN
    sev = ppmg_last_error(code, buf, 256_C_INT)
    msg = ' '
    DO i = 1, 256
      IF (buf(i) == C_NULL_CHAR) EXIT
      msg(i:i) = buf(i)
    END DO
    WRITE(*, '(I4.4,A,A,A,A)') code, ' ERROR in ', caller, ' : ', TRIM(msg)
  END SUBROUTINE ppmg_report

  ! Mesh nodes of a registered topology (single sub: Nm on every axis).
  SUBROUTINE ppmg_topo_nm(topo_id, dim, nm, info)
    INTEGER, INTENT(IN) :: topo_id
    INTEGER, INTENT(OUT) :: dim, info
    INTEGER, INTENT(OUT) :: nm(3)
    INTEGER(C_INT) :: cdim, cnm(3), cbc(6)
    REAL(C_DOUBLE) :: mn(3), mx(3)
    info = ppmg_topo_info(INT(topo_id, C_INT), cdim, cnm, mn, mx, cbc)
    dim = cdim
    nm = cnm
  END SUBROUTINE ppmg_topo_nm

  ! The reference indexes field_up(.., 1-g:Nm+g, ..); the library takes exactly
  ! that box.  ok = .FALSE. when the caller's bounds differ (a staging copy is
  ! then made by the wrappers).
  LOGICAL FUNCTION ppmg_field_bounds_ok(lb, ub, first, dim, nm, ghost)
    INTEGER, INTENT(IN) :: lb(:), ub(:), first, dim, nm(3), ghost(:)
    INTEGER :: d
    ppmg_field_bounds_ok = .TRUE.
    DO d = 1, dim
      IF (lb(first + d - 1) /= 1 - ghost(d)) ppmg_field_bounds_ok = .FALSE.
      IF (ub(first + d - 1) /= nm(d) + ghost(d)) ppmg_field_bounds_ok = .FALSE.
    END DO
  END FUNCTION ppmg_field_bounds_ok

END MODULE ppm_module_ppmg

!-------------------------------------------------------------------------------
! ppmg_fcore -- checks and C calls shared by the drop-in specifics.  The core
! routines take assumed-size dummies, so the specifics pass array sections
! (xp(1:dim,1:Np), field_up(1:lda,1-g:Nm+g,...,1)) and the compiler supplies a
! contiguous copy only when the caller's array is not already that box.
!-------------------------------------------------------------------------------
MODULE ppmg_fcore
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  IMPLICIT NONE
CONTAINS

  SUBROUTINE ppmg_argfail(caller, msg, info)
    CHARACTER(LEN=*), INTENT(IN) :: caller, msg
    INTEGER, INTENT(OUT) :: info
    info = ppm_error_error
    IF (ppm_debug > 0) WRITE(*, '(I4.4,A,A,A,A)') ppm_err_argument, ' ERROR in ', caller, ' : ', msg
  END SUBROUTINE ppmg_argfail

  ! Field box of topo_id: lo = 1-ghostsize, hi = Nm+ghostsize on the first dim axes.
  SUBROUTINE ppmg_box(caller, topo_id, ghostsize, dim, lo, hi, cg, info)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    INTEGER, INTENT(IN) :: topo_id, dim
    INTEGER, DIMENSION(:), INTENT(IN) :: ghostsize
    INTEGER, INTENT(OUT) :: lo(3), hi(3), info
    INTEGER(C_INT), INTENT(OUT) :: cg(3)
    INTEGER :: tdim, nm(3), d
    lo = 1
    hi = 1
    cg = 0
    CALL ppmg_topo_nm(topo_id, tdim, nm, info)
    IF (info /= 0) THEN
      CALL ppmg_report(caller, info)
      RETURN
    END IF
    IF (tdim /= dim) THEN
      CALL ppmg_argfail(caller, 'topology dimension differs from the specific called', info)
      RETURN
    END IF
    IF (SIZE(ghostsize) < dim) THEN
      CALL ppmg_argfail(caller, 'ghostsize too short', info)
      RETURN
    END IF
    DO d = 1, dim
      lo(d) = 1 - ghostsize(d)
      hi(d) = nm(d) + ghostsize(d)
      cg(d) = ghostsize(d)
    END DO
  END SUBROUTINE ppmg_box

  ! The caller's (associated) field_up, bounds lb/ub, must hold components
  ! 1..lda (when vec), the box lo..hi and sub index 1.
  SUBROUTINE ppmg_field_check(caller, lb, ub, vec, lda, dim, lo, hi, info)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    LOGICAL, INTENT(IN) :: vec
    INTEGER, DIMENSION(:), INTENT(IN) :: lb, ub
    INTEGER, INTENT(IN) :: lda, dim, lo(3), hi(3)
    INTEGER, INTENT(OUT) :: info
    INTEGER :: f, d
    info = 0
    f = 0
    IF (vec) THEN
      f = 1
      IF (lb(1) > 1 .OR. ub(1) < lda) THEN
        CALL ppmg_argfail(caller, 'leading dimension of field_up insufficient', info)
        RETURN
      END IF
    END IF
    DO d = 1, dim
      IF (lb(f + d) > lo(d) .OR. ub(f + d) < hi(d)) THEN
        CALL ppmg_argfail(caller, 'field_up does not hold the mesh and its ghost layers', info)
        RETURN
      END IF
    END DO
    IF (lb(f + dim + 1) > 1 .OR. ub(f + dim + 1) < 1) THEN
      CALL ppmg_argfail(caller, 'field_up has no sub 1', info)
      RETURN
    END IF
  END SUBROUTINE ppmg_field_check

  ! xp(dim, >=Np) and up(lda, >=Np) (lda = 0: scalar up(>=Np)).
  SUBROUTINE ppmg_part_check(caller, np, xshape, dim, ushape, lda, info)
    CHARACTER(LEN=*), INTENT(IN) :: caller
    INTEGER, INTENT(IN) :: np, dim, lda
    INTEGER, DIMENSION(:), INTENT(IN) :: xshape, ushape
    INTEGER, INTENT(OUT) :: info
    info = 0
    IF (np < 0) THEN
      CALL ppmg_argfail(caller, 'particles must be specified', info)
      RETURN
    END IF
    IF (np == 0) RETURN
    IF (xshape(1) < dim .OR. xshape(2) < np) THEN
      CALL ppmg_argfail(caller, 'xp too small', info)
      RETURN
    END IF
    IF (lda == 0) THEN
      IF (ushape(1) < np) CALL ppmg_argfail(caller, 'up too small', info)
    ELSE
      IF (ushape(1) < lda .OR. ushape(2) < np) CALL ppmg_argfail(caller, 'up too small', info)
    END IF
  END SUBROUTINE ppmg_part_check

  FUNCTION ploc_d(a, n) RESULT(p)
    REAL(C_DOUBLE), TARGET, INTENT(IN) :: a(*)
    INTEGER, INTENT(IN) :: n
    TYPE(C_PTR) :: p
    p = C_NULL_PTR
    IF (n > 0) p = C_LOC(a(1))
  END FUNCTION ploc_d

  FUNCTION ploc_s(a, n) RESULT(p)
    REAL(C_FLOAT), TARGET, INTENT(IN) :: a(*)
    INTEGER, INTENT(IN) :: n
    TYPE(C_PTR) :: p
    p = C_NULL_PTR
    IF (n > 0) p = C_LOC(a(1))
  END FUNCTION ploc_s

  SUBROUTINE p2m_core_d(xp, up, field, np, lda, topo_id, mesh_id, kernel, cg, info)
    REAL(C_DOUBLE), TARGET, INTENT(IN) :: xp(*), up(*)
    REAL(C_DOUBLE), TARGET :: field(*)
    INTEGER, INTENT(IN) :: np, lda, topo_id, mesh_id, kernel
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER, INTENT(OUT) :: info
    info = ppmg_interp_p2m(ploc_d(xp, np), INT(np, C_INT), ploc_d(up, np), INT(lda, C_INT), INT(topo_id, C_INT), &
                           INT(mesh_id, C_INT), INT(kernel, C_INT), cg, C_LOC(field(1)), ppmg_prec_f64, ppmg_mem_host)
    CALL ppmg_report('ppm_interp_p2m', info)
  END SUBROUTINE p2m_core_d

  SUBROUTINE p2m_core_s(xp, up, field, np, lda, topo_id, mesh_id, kernel, cg, info)
    REAL(C_FLOAT), TARGET, INTENT(IN) :: xp(*), up(*)
    REAL(C_FLOAT), TARGET :: field(*)
    INTEGER, INTENT(IN) :: np, lda, topo_id, mesh_id, kernel
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER, INTENT(OUT) :: info
    info = ppmg_interp_p2m(ploc_s(xp, np), INT(np, C_INT), ploc_s(up, np), INT(lda, C_INT), INT(topo_id, C_INT), &
                           INT(mesh_id, C_INT), INT(kernel, C_INT), cg, C_LOC(field(1)), ppmg_prec_f32, ppmg_mem_host)
    CALL ppmg_report('ppm_interp_p2m', info)
  END SUBROUTINE p2m_core_s

  SUBROUTINE m2p_core_d(xp, up, field, np, lda, topo_id, mesh_id, kernel, cg, info)
    REAL(C_DOUBLE), TARGET, INTENT(IN) :: xp(*), field(*)
    REAL(C_DOUBLE), TARGET :: up(*)
    INTEGER, INTENT(IN) :: np, lda, topo_id, mesh_id, kernel
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER, INTENT(OUT) :: info
    info = ppmg_interp_m2p(ploc_d(xp, np), INT(np, C_INT), ploc_d(up, np), INT(lda, C_INT), INT(topo_id, C_INT), &
                           INT(mesh_id, C_INT), INT(kernel, C_INT), cg, C_LOC(field(1)), ppmg_prec_f64, ppmg_mem_host)
    CALL ppmg_report('ppm_interp_m2p', info)
  END SUBROUTINE m2p_core_d

  SUBROUTINE m2p_core_s(xp, up, field, np, lda, topo_id, mesh_id, kernel, cg, info)
    REAL(C_FLOAT), TARGET, INTENT(IN) :: xp(*), field(*)
    REAL(C_FLOAT), TARGET :: up(*)
    INTEGER, INTENT(IN) :: np, lda, topo_id, mesh_id, kernel
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER, INTENT(OUT) :: info
    info = ppmg_interp_m2p(ploc_s(xp, np), INT(np, C_INT), ploc_s(up, np), INT(lda, C_INT), INT(topo_id, C_INT), &
                           INT(mesh_id, C_INT), INT(kernel, C_INT), cg, C_LOC(field(1)), ppmg_prec_f32, ppmg_mem_host)
    CALL ppmg_report('ppm_interp_m2p', info)
  END SUBROUTINE m2p_core_s

  SUBROUTINE remesh_core_d(xp, up, field, w1, w2, w3, np, lda, topo_id, mesh_id, kernel, cg, info)
    REAL(C_DOUBLE), TARGET, INTENT(IN) :: xp(*), up(*), w1(*), w2(*), w3(*)
    REAL(C_DOUBLE), TARGET :: field(*)
    INTEGER, INTENT(IN) :: np, lda, topo_id, mesh_id, kernel
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER, INTENT(OUT) :: info
    info = ppmg_rmsh_remesh_reg(ploc_d(xp, np), INT(np, C_INT), ploc_d(up, np), INT(lda, C_INT), INT(topo_id, C_INT), &
                                INT(mesh_id, C_INT), INT(kernel, C_INT), cg, C_LOC(field(1)), ploc_d(w1, np), &
                                ploc_d(w2, np), ploc_d(w3, np), ppmg_prec_f64, ppmg_mem_host)
    CALL ppmg_report('ppm_rmsh_remesh', info)
  END SUBROUTINE remesh_core_d

  SUBROUTINE remesh_core_s(xp, up, field, w1, w2, w3, np, lda, topo_id, mesh_id, kernel, cg, info)
    REAL(C_FLOAT), TARGET, INTENT(IN) :: xp(*), up(*), w1(*), w2(*), w3(*)
    REAL(C_FLOAT), TARGET :: field(*)
    INTEGER, INTENT(IN) :: np, lda, topo_id, mesh_id, kernel
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER, INTENT(OUT) :: info
    info = ppmg_rmsh_remesh_reg(ploc_s(xp, np), INT(np, C_INT), ploc_s(up, np), INT(lda, C_INT), INT(topo_id, C_INT), &
                                INT(mesh_id, C_INT), INT(kernel, C_INT), cg, C_LOC(field(1)), ploc_s(w1, np), &
                                ploc_s(w2, np), ploc_s(w3, np), ppmg_prec_f32, ppmg_mem_host)
    CALL ppmg_report('ppm_rmsh_remesh', info)
  END SUBROUTINE remesh_core_s

  SUBROUTINE ghost_core_d(field, lda, topo_id, mesh_id, cg, maptype, mi, has_mask, info)
    REAL(C_DOUBLE), TARGET :: field(*)
    INTEGER, INTENT(IN) :: lda, topo_id, mesh_id, maptype
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER(C_INT), TARGET, INTENT(IN) :: mi(*)
    LOGICAL, INTENT(IN) :: has_mask
    INTEGER, INTENT(OUT) :: info
    info = ppmg_map_field_ghost_kind_reg(C_LOC(field(1)), INT(lda, C_INT), INT(topo_id, C_INT), INT(mesh_id, C_INT), &
                                         cg, INT(maptype, C_INT), ppmg_kind_d, mask_ptr(mi, has_mask), ppmg_mem_host)
    CALL ppmg_report('ppm_map_field_ghost', info)
  END SUBROUTINE ghost_core_d

  SUBROUTINE ghost_core_s(field, lda, topo_id, mesh_id, cg, maptype, mi, has_mask, info)
    REAL(C_FLOAT), TARGET :: field(*)
    INTEGER, INTENT(IN) :: lda, topo_id, mesh_id, maptype
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER(C_INT), TARGET, INTENT(IN) :: mi(*)
    LOGICAL, INTENT(IN) :: has_mask
    INTEGER, INTENT(OUT) :: info
    info = ppmg_map_field_ghost_kind_reg(C_LOC(field(1)), INT(lda, C_INT), INT(topo_id, C_INT), INT(mesh_id, C_INT), &
                                         cg, INT(maptype, C_INT), ppmg_kind_s, mask_ptr(mi, has_mask), ppmg_mem_host)
    CALL ppmg_report('ppm_map_field_ghost', info)
  END SUBROUTINE ghost_core_s

  SUBROUTINE ghost_core_dc(field, lda, topo_id, mesh_id, cg, maptype, mi, has_mask, info)
    COMPLEX(C_DOUBLE_COMPLEX), TARGET :: field(*)
    INTEGER, INTENT(IN) :: lda, topo_id, mesh_id, maptype
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER(C_INT), TARGET, INTENT(IN) :: mi(*)
    LOGICAL, INTENT(IN) :: has_mask
    INTEGER, INTENT(OUT) :: info
    info = ppmg_map_field_ghost_kind_reg(C_LOC(field(1)), INT(lda, C_INT), INT(topo_id, C_INT), INT(mesh_id, C_INT), &
                                         cg, INT(maptype, C_INT), ppmg_kind_dc, mask_ptr(mi, has_mask), ppmg_mem_host)
    CALL ppmg_report('ppm_map_field_ghost', info)
  END SUBROUTINE ghost_core_dc

  SUBROUTINE ghost_core_sc(field, lda, topo_id, mesh_id, cg, maptype, mi, has_mask, info)
    COMPLEX(C_FLOAT_COMPLEX), TARGET :: field(*)
    INTEGER, INTENT(IN) :: lda, topo_id, mesh_id, maptype
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER(C_INT), TARGET, INTENT(IN) :: mi(*)
    LOGICAL, INTENT(IN) :: has_mask
    INTEGER, INTENT(OUT) :: info
    info = ppmg_map_field_ghost_kind_reg(C_LOC(field(1)), INT(lda, C_INT), INT(topo_id, C_INT), INT(mesh_id, C_INT), &
                                         cg, INT(maptype, C_INT), ppmg_kind_sc, mask_ptr(mi, has_mask), ppmg_mem_host)
    CALL ppmg_report('ppm_map_field_ghost', info)
  END SUBROUTINE ghost_core_sc

  SUBROUTINE ghost_core_i(field, lda, topo_id, mesh_id, cg, maptype, mi, has_mask, info)
    INTEGER(C_INT), TARGET :: field(*)
    INTEGER, INTENT(IN) :: lda, topo_id, mesh_id, maptype
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER(C_INT), TARGET, INTENT(IN) :: mi(*)
    LOGICAL, INTENT(IN) :: has_mask
    INTEGER, INTENT(OUT) :: info
    info = ppmg_map_field_ghost_kind_reg(C_LOC(field(1)), INT(lda, C_INT), INT(topo_id, C_INT), INT(mesh_id, C_INT), &
                                         cg, INT(maptype, C_INT), ppmg_kind_i, mask_ptr(mi, has_mask), ppmg_mem_host)
    CALL ppmg_report('ppm_map_field_ghost', info)
  END SUBROUTINE ghost_core_i

  ! Logical fields travel as C logicals (1 / 0) and come back as .TRUE. / .FALSE.
  SUBROUTINE ghost_core_l(field, lda, topo_id, mesh_id, cg, maptype, mi, has_mask, info)
    LOGICAL :: field(*)
    INTEGER, INTENT(IN) :: lda, topo_id, mesh_id, maptype
    INTEGER(C_INT), INTENT(IN) :: cg(3)
    INTEGER(C_INT), TARGET, INTENT(IN) :: mi(*)
    LOGICAL, INTENT(IN) :: has_mask
    INTEGER, INTENT(OUT) :: info
    INTEGER(C_INT), ALLOCATABLE, TARGET :: fi(:)
    INTEGER :: n, d, nm(3), tdim
    CALL ppmg_topo_nm(topo_id, tdim, nm, info)
    IF (info /= 0) THEN
      CALL ppmg_report('ppm_map_field_ghost', info)
      RETURN
    END IF
    n = lda
    DO d = 1, tdim
      n = n * (nm(d) + 2 * cg(d))
    END DO
    ALLOCATE(fi(n))
    fi = MERGE(1_C_INT, 0_C_INT, field(1:n))
    info = ppmg_map_field_ghost_kind_reg(C_LOC(fi(1)), INT(lda, C_INT), INT(topo_id, C_INT), INT(mesh_id, C_INT), &
                                         cg, INT(maptype, C_INT), ppmg_kind_l, mask_ptr(mi, has_mask), ppmg_mem_host)
    IF (info == 0) field(1:n) = fi /= 0
    CALL ppmg_report('ppm_map_field_ghost', info)
  END SUBROUTINE ghost_core_l

  FUNCTION mask_ptr(mi, has_mask) RESULT(p)
    INTEGER(C_INT), TARGET, INTENT(IN) :: mi(*)
    LOGICAL, INTENT(IN) :: has_mask
    TYPE(C_PTR) :: p
    p = C_NULL_PTR
    IF (has_mask) p = C_LOC(mi(1))
  END FUNCTION mask_ptr

END MODULE ppmg_fcore
