!-------------------------------------------------------------------------------
! Setup modules of the drop-in: ppm_module_data (constants), ppm_module_init,
! ppm_module_finalize, ppm_module_mktopo (field topology of one sub covering
! the domain, what ppm_topo_mkfield builds on one process) and
! ppm_module_impose_part_bc.  Same module, generic and argument names as the
! reference (ppm_module_init.f:29-45, ppm_init.f:161; ppm_module_mktopo.f:43-52,
! ppm_topo_mkfield.f:316-319; ppm_module_impose_part_bc.f).
!-------------------------------------------------------------------------------
MODULE ppm_module_data
  USE ppm_module_ppmg, ONLY: ppm_kind_single, ppm_kind_double, ppm_param_rmsh_kernel_bsp2, &
       ppm_param_rmsh_kernel_mp4, ppm_param_rmsh_kernel_m3p6, ppm_param_bcdef_periodic, &
       ppm_param_bcdef_freespace, ppm_param_bcdef_symmetry, ppm_param_bcdef_neumann, &
       ppm_param_bcdef_dirichlet, ppm_param_bcdef_robin, ppm_error_notice, ppm_error_warning, &
       ppm_error_error, ppm_error_fatal, ppm_err_alloc, ppm_err_part_unass, ppm_err_argument, &
       ppm_err_ppm_noinit, ppm_err_part_lost, ppm_param_decomp_cuboid, ppm_param_assign_internal, ppm_debug, ppm_dim, &
       ppm_param_map_pop, ppm_param_map_push, ppm_param_map_send, ppm_param_map_ghost_get, &
       ppm_param_map_ghost_put, ppm_param_map_cancel, ppm_param_map_init
  IMPLICIT NONE
END MODULE ppm_module_data

MODULE ppm_module_init
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  IMPLICIT NONE
  INTERFACE ppm_init
    MODULE PROCEDURE ppm_init
  END INTERFACE
CONTAINS
  ! ppm_init(dim,prec,tolexp,comm,debug,info,logfile,stderr,stdout): comm is the
  ! MPI communicator of the reference; one process drives one GPU here.
  SUBROUTINE ppm_init(dim, prec, tolexp, comm, debug, info, logfile, stderr, stdout)
    INTEGER, INTENT(IN) :: dim, prec, comm, debug, tolexp
    INTEGER, OPTIONAL :: logfile, stderr, stdout
    INTEGER, INTENT(OUT) :: info
    INTEGER(C_INT) :: p
    p = ppmg_prec_f64
    IF (prec == ppm_kind_single) p = ppmg_prec_f32
    ppm_debug = debug
    ppm_dim = dim
    info = ppmg_init(INT(dim, C_INT), p, INT(tolexp, C_INT), INT(debug, C_INT))
    CALL ppmg_report('ppm_init', info)
  END SUBROUTINE ppm_init
END MODULE ppm_module_init

MODULE ppm_module_finalize
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  IMPLICIT NONE
  INTERFACE ppm_finalize
    MODULE PROCEDURE ppm_finalize
  END INTERFACE
CONTAINS
  SUBROUTINE ppm_finalize(info)
    INTEGER, INTENT(OUT) :: info
    info = ppmg_finalize()
  END SUBROUTINE ppm_finalize
END MODULE ppm_module_finalize

MODULE ppm_module_mktopo
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  IMPLICIT NONE
  INTERFACE ppm_mktopo
    MODULE PROCEDURE ppm_topo_mkfield_s
    MODULE PROCEDURE ppm_topo_mkfield_d
  END INTERFACE
CONTAINS
  SUBROUTINE mkfield_common(dim, mn, mx, Nm, bcdef, ghostsize, topo_id, mesh_id, nsubs, isublist, nsublist, &
                            istart, ndata, sub2proc, info)
    INTEGER, INTENT(IN) :: dim
    REAL(C_DOUBLE), INTENT(IN) :: mn(3), mx(3)
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, bcdef, ghostsize
    INTEGER, INTENT(INOUT) :: topo_id, mesh_id, nsubs
    INTEGER, DIMENSION(:), POINTER :: isublist, sub2proc
    INTEGER, DIMENSION(:,:), POINTER :: istart, ndata
    INTEGER, INTENT(OUT) :: nsublist, info
    INTEGER(C_INT) :: cnm(3), cbc(6), cg(3), tid, mid
    INTEGER :: d
    cnm = 2
    cbc = 1
    cg = 0
    DO d = 1, dim
      cnm(d) = Nm(d)
      cg(d) = ghostsize(d)
    END DO
    DO d = 1, 2 * dim
      cbc(d) = bcdef(d)
    END DO
    info = ppmg_mktopo_field(INT(dim, C_INT), mn, mx, cbc, cnm, cg, tid, mid)
    CALL ppmg_report('ppm_mktopo', info)
    IF (info /= 0) RETURN
    topo_id = tid
    mesh_id = mid
    ppmg_current_topo = tid
    nsubs = 1
    nsublist = 1
    ALLOCATE(isublist(1), sub2proc(1), istart(dim, 1), ndata(dim, 1))
    isublist(1) = 1
    sub2proc(1) = 0
    istart(:, 1) = 1
    ndata(1:dim, 1) = Nm(1:dim)
  END SUBROUTINE mkfield_common

  SUBROUTINE ppm_topo_mkfield_s(xp, Npart, Nm, decomp, assig, min_phys, max_phys, bcdef, ghostsize, topo_id, &
                                mesh_id, min_sub, max_sub, cost, sub2proc, nsubs, isublist, nsublist, istart, &
                                ndata, info, ndom, pcost)
    INTEGER, PARAMETER :: MK = ppm_kind_single
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(IN) :: Npart, assig
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, bcdef, ghostsize
    REAL(MK), DIMENSION(:), INTENT(IN) :: min_phys, max_phys
    REAL(MK), DIMENSION(:), OPTIONAL, INTENT(IN) :: pcost
    INTEGER, OPTIONAL, INTENT(IN) :: ndom
    INTEGER, INTENT(INOUT) :: nsubs, topo_id, mesh_id, decomp
    REAL(MK), DIMENSION(:,:), POINTER :: min_sub, max_sub
    REAL(MK), DIMENSION(:), POINTER :: cost
    INTEGER, DIMENSION(:,:), POINTER :: istart, ndata
    INTEGER, DIMENSION(:), POINTER :: sub2proc, isublist
    INTEGER, INTENT(OUT) :: nsublist, info
    REAL(C_DOUBLE) :: mn(3), mx(3)
    INTEGER :: dim
    dim = SIZE(Nm)
    mn = 0.0_C_DOUBLE
    mx = 1.0_C_DOUBLE
    mn(1:dim) = REAL(min_phys(1:dim), C_DOUBLE)
    mx(1:dim) = REAL(max_phys(1:dim), C_DOUBLE)
    CALL mkfield_common(dim, mn, mx, Nm, bcdef, ghostsize, topo_id, mesh_id, nsubs, isublist, nsublist, &
                        istart, ndata, sub2proc, info)
    IF (info /= 0) RETURN
    ALLOCATE(min_sub(dim, 1), max_sub(dim, 1), cost(1))
    min_sub(:, 1) = min_phys(1:dim)
    max_sub(:, 1) = max_phys(1:dim)
    cost(1) = REAL(Npart, MK)
  END SUBROUTINE ppm_topo_mkfield_s

  SUBROUTINE ppm_topo_mkfield_d(xp, Npart, Nm, decomp, assig, min_phys, max_phys, bcdef, ghostsize, topo_id, &
                                mesh_id, min_sub, max_sub, cost, sub2proc, nsubs, isublist, nsublist, istart, &
                                ndata, info, ndom, pcost)
    INTEGER, PARAMETER :: MK = ppm_kind_double
    REAL(MK), DIMENSION(:,:), POINTER :: xp
    INTEGER, INTENT(IN) :: Npart, assig
    INTEGER, DIMENSION(:), INTENT(IN) :: Nm, bcdef, ghostsize
    REAL(MK), DIMENSION(:), INTENT(IN) :: min_phys, max_phys
    REAL(MK), DIMENSION(:), OPTIONAL, INTENT(IN) :: pcost
    INTEGER, OPTIONAL, INTENT(IN) :: ndom
    INTEGER, INTENT(INOUT) :: nsubs, topo_id, mesh_id, decomp
    REAL(MK), DIMENSION(:,:), POINTER :: min_sub, max_sub
    REAL(MK), DIMENSION(:), POINTER :: cost
    INTEGER, DIMENSION(:,:), POINTER :: istart, ndata
    INTEGER, DIMENSION(:), POINTER :: sub2proc, isublist
    INTEGER, INTENT(OUT) :: nsublist, info
    REAL(C_DOUBLE) :: mn(3), mx(3)
    INTEGER :: dim
    dim = SIZE(Nm)
    mn = 0.0_C_DOUBLE
    mx = 1.0_C_DOUBLE
    mn(1:dim) = min_phys(1:dim)
    mx(1:dim) = max_phys(1:dim)
    CALL mkfield_common(dim, mn, mx, Nm, bcdef, ghostsize, topo_id, mesh_id, nsubs, isublist, nsublist, &
                        istart, ndata, sub2proc, info)
    IF (info /= 0) RETURN
    ALLOCATE(min_sub(dim, 1), max_sub(dim, 1), cost(1))
    min_sub(:, 1) = min_phys(1:dim)
    max_sub(:, 1) = max_phys(1:dim)
    cost(1) = REAL(Npart, MK)
  END SUBROUTINE ppm_topo_mkfield_d
END MODULE ppm_module_mktopo

MODULE ppm_module_impose_part_bc
  USE ISO_C_BINDING
  USE ppm_module_ppmg
  IMPLICIT NONE
  INTERFACE ppm_impose_part_bc
    MODULE PROCEDURE ppm_impose_part_bc_s
    MODULE PROCEDURE ppm_impose_part_bc_d
  END INTERFACE
CONTAINS
  SUBROUTINE ppm_impose_part_bc_s(xp, Npart, topo_id, info)
    REAL(ppm_kind_single), DIMENSION(:,:), TARGET, CONTIGUOUS, INTENT(INOUT) :: xp
    INTEGER, INTENT(IN) :: Npart, topo_id
    INTEGER, INTENT(OUT) :: info
    info = 0
    IF (Npart <= 0) RETURN
    info = ppmg_impose_part_bc_reg(C_LOC(xp), INT(Npart, C_INT), INT(topo_id, C_INT), ppmg_prec_f32, ppmg_mem_host)
    CALL ppmg_report('ppm_impose_part_bc', info)
  END SUBROUTINE ppm_impose_part_bc_s
  SUBROUTINE ppm_impose_part_bc_d(xp, Npart, topo_id, info)
    REAL(ppm_kind_double), DIMENSION(:,:), TARGET, CONTIGUOUS, INTENT(INOUT) :: xp
    INTEGER, INTENT(IN) :: Npart, topo_id
    INTEGER, INTENT(OUT) :: info
    info = 0
    IF (Npart <= 0) RETURN
    info = ppmg_impose_part_bc_reg(C_LOC(xp), INT(Npart, C_INT), INT(topo_id, C_INT), ppmg_prec_f64, ppmg_mem_host)
    CALL ppmg_report('ppm_impose_part_bc', info)
  END SUBROUTINE ppm_impose_part_bc_d
END MODULE ppm_module_impose_part_bc
