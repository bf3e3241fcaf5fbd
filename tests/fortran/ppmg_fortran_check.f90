!-------------------------------------------------------------------------------
! ppmg_fortran_check -- a PPM client program written against the reference's
! module and routine names, linked against libppm_fortran.a + libppmg.so.
! Reads particles from an input stream file, runs the interpolation and
! remeshing routines through the drop-in modules and writes every result to
! an output stream file, which the test suite checks.
!
!   ppmg_fortran_check <input> <output>
!
! Input (stream, native endian): nm, np, lda, kernel, ghost (int32); then
! xp(3,np), up(lda,np), xq(3,np) (float64).
!-------------------------------------------------------------------------------
PROGRAM ppmg_fortran_check
  USE ppm_module_data
  USE ppm_module_init
  USE ppm_module_finalize
  USE ppm_module_mktopo
  USE ppm_module_interp_p2m
  USE ppm_module_interp_m2p
  USE ppm_module_rmsh
  USE ppm_module_util_rank
  USE ppm_module_impose_part_bc
  USE ppm_module_map_field_ghost
  USE ppm_module_util_sort
  USE ppm_module_neighlist
  IMPLICIT NONE
  INTEGER, PARAMETER :: MK = ppm_kind_double
  CHARACTER(LEN=1024) :: fin, fout
  INTEGER :: nm, np, lda, kernel, g, info, topo, mesh, topo2, mesh2, decomp, nsubs, nsublist, npn, i
  INTEGER, DIMENSION(3) :: nmv, gs
  INTEGER, DIMENSION(6) :: bc, ngl
  REAL(MK), DIMENSION(3) :: minp, maxp
  REAL(MK), DIMENSION(:,:), POINTER :: xp => NULL(), up => NULL(), xq => NULL(), uq => NULL()
  REAL(MK), DIMENSION(:,:), POINTER :: xn => NULL(), un => NULL(), wn => NULL(), x2 => NULL()
  REAL(MK), DIMENSION(:,:,:,:,:), POINTER :: field => NULL(), field2 => NULL(), fbig => NULL(), fg => NULL()
  REAL(MK), DIMENSION(:,:,:), POINTER :: f2d => NULL()
  REAL(MK), DIMENSION(:), POINTER :: u1 => NULL()
  REAL(ppm_kind_single), DIMENSION(:,:), POINTER :: xs => NULL()
  REAL(ppm_kind_single), DIMENSION(:), POINTER :: us => NULL()
  REAL(ppm_kind_single), DIMENSION(:,:,:,:), POINTER :: fs => NULL()
  REAL(MK), DIMENSION(:,:), POINTER :: min_sub => NULL(), max_sub => NULL()
  REAL(MK), DIMENSION(:), POINTER :: cost => NULL()
  INTEGER, DIMENSION(:,:), POINTER :: istart => NULL(), ndata => NULL()
  INTEGER, DIMENSION(:), POINTER :: sub2proc => NULL(), isublist => NULL(), lpdx => NULL(), lhbx => NULL()
  REAL(MK), DIMENSION(2) :: cutoff
  INTEGER, DIMENSION(:), POINTER :: npbx => NULL()
  INTEGER, DIMENSION(:,:), POINTER :: nmc => NULL()
  TYPE(ppm_type_ptr_to_clist), DIMENSION(:), POINTER :: clist => NULL()
  REAL(MK), DIMENSION(3) :: rc
  INTEGER, DIMENSION(:,:,:,:,:), POINTER :: fi => NULL()
  LOGICAL, DIMENSION(:,:,:,:), POINTER :: fl => NULL(), msk => NULL()
  LOGICAL, DIMENSION(:,:,:), POINTER :: msk2 => NULL()
  COMPLEX(MK), DIMENSION(:,:,:,:), POINTER :: fc => NULL()
  INTEGER :: topo3, mesh3, j, k, c

  CALL GET_COMMAND_ARGUMENT(1, fin)
  CALL GET_COMMAND_ARGUMENT(2, fout)
  OPEN(10, FILE=TRIM(fin), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='OLD')
  READ(10) nm, np, lda, kernel, g
  ALLOCATE(xp(3, np), up(lda, np), xq(3, np), uq(lda, np))
  READ(10) xp, up, xq
  CLOSE(10)
  OPEN(20, FILE=TRIM(fout), ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')

  CALL ppm_init(3, ppm_kind_double, -15, 0, 1, info)
  CALL check(info, 'ppm_init')
  nmv = nm
  gs = g
  bc = ppm_param_bcdef_periodic
  minp = 0.0_MK
  maxp = 1.0_MK
  decomp = ppm_param_decomp_cuboid
  CALL ppm_mktopo(xp, np, nmv, decomp, ppm_param_assign_internal, minp, maxp, bc, gs, topo, mesh, &
                  min_sub, max_sub, cost, sub2proc, nsubs, isublist, nsublist, istart, ndata, info)
  CALL check(info, 'ppm_mktopo')

  ! 1. P2M, vector double (dv_3d)
  ALLOCATE(field(lda, 1-g:nm+g, 1-g:nm+g, 1-g:nm+g, 1))
  CALL ppm_interp_p2m(xp, np, up, lda, topo, mesh, kernel, gs, field, info)
  CALL check(info, 'ppm_interp_p2m dv_3d')
  WRITE(20) field

  ! 2. M2P from that field at xq
  CALL ppm_interp_m2p(xq, np, uq, lda, topo, mesh, kernel, gs, field, info)
  CALL check(info, 'ppm_interp_m2p dv_3d')
  WRITE(20) uq

  ! 3. the same P2M into a field with wider bounds than the box (staged copy)
  ALLOCATE(fbig(lda+1, -g:nm+g+1, -g:nm+g+1, -g:nm+g+1, 1))
  fbig = -7.0_MK
  CALL ppm_interp_p2m(xp, np, up, lda, topo, mesh, kernel, gs, fbig, info)
  CALL check(info, 'ppm_interp_p2m dv_3d (wide field)')
  WRITE(20) fbig

  ! 4. comp_weights + remesh (field allocated by remesh)
  CALL ppm_rmsh_comp_weights(xp, np, topo, mesh, kernel, info)
  CALL check(info, 'ppm_rmsh_comp_weights')
  CALL ppm_rmsh_remesh(xp, np, up, lda, topo, mesh, kernel, gs, info, field2)
  CALL check(info, 'ppm_rmsh_remesh')
  WRITE(20) LBOUND(field2), UBOUND(field2)
  WRITE(20) field2

  ! 5. create_part with a slave field (svv form): master = P2M field, slave = remeshed field
  cutoff(1) = 0.05_MK * MAXVAL(ABS(field))
  cutoff(2) = 0.6_MK * MAXVAL(ABS(field)) * SQRT(REAL(lda, MK))
  CALL ppm_rmsh_create_part(xn, npn, un, lda, field, topo, mesh, cutoff, info, .TRUE., field_wp=field2, wp=wn, &
                            lda2=lda)
  CALL check(info, 'ppm_rmsh_create_part dvv_3d')
  WRITE(20) npn
  WRITE(20) cutoff
  WRITE(20) xn, un, wn

  ! 6. util_rank3d on the mesh cells with one ghost layer each side
  ngl = 1
  CALL ppm_util_rank3d(xp, np, minp, maxp, nmv - 1, ngl, lpdx, lhbx, info)
  CALL check(info, 'ppm_util_rank3d')
  WRITE(20) SIZE(lpdx), SIZE(lhbx)
  WRITE(20) lpdx, lhbx

  ! 7. P2M single precision scalar (ss_3d) of component 1
  ALLOCATE(xs(3, np), us(np), fs(1-g:nm+g, 1-g:nm+g, 1-g:nm+g, 1))
  xs = REAL(xp, ppm_kind_single)
  us = REAL(up(1, :), ppm_kind_single)
  CALL ppm_interp_p2m(xs, np, us, topo, mesh, kernel, gs, fs, info)
  CALL check(info, 'ppm_interp_p2m ss_3d')
  WRITE(20) fs

  ! 8. impose_part_bc on shifted positions
  ALLOCATE(x2(3, np))
  x2 = xp + 0.75_MK
  CALL ppm_impose_part_bc(x2, np, topo, info)
  CALL check(info, 'ppm_impose_part_bc')
  WRITE(20) x2

  ! 9. 2D topology, P2M double scalar (ds_2d) of the xy projection
  CALL ppm_mktopo(xp, np, nmv(1:2), decomp, ppm_param_assign_internal, minp(1:2), maxp(1:2), bc(1:4), gs(1:2), &
                  topo2, mesh2, min_sub, max_sub, cost, sub2proc, nsubs, isublist, nsublist, istart, ndata, info)
  CALL check(info, 'ppm_mktopo 2d')
  ALLOCATE(f2d(1-g:nm+g, 1-g:nm+g, 1), u1(np))
  u1 = up(1, :)
  CALL ppm_interp_p2m(xp(1:2, :), np, u1, topo2, mesh2, kernel, gs(1:2), f2d, info)
  CALL check(info, 'ppm_interp_p2m ds_2d')
  WRITE(20) f2d

  ! 10. error path: a particle outside the domain is the reference's fatal 405
  xp(2, 3) = 1.5_MK
  CALL ppm_interp_p2m(xp, np, up, lda, topo, mesh, kernel, gs, field, info)
  WRITE(20) info

  ! 11. ghost mapping on a copy of the remeshed field: ghost_put, then ghost_get,
  !     each as init / map / push / send / pop; then a pop without push
  ALLOCATE(fg(lda, 1-g:nm+g, 1-g:nm+g, 1-g:nm+g, 1))
  fg = field2
  CALL ppm_map_field_ghost(fg, lda, topo, mesh, gs, ppm_param_map_init, info)
  CALL check(info, 'ppm_map_field_ghost init')
  CALL ppm_map_field_ghost(fg, lda, topo, mesh, gs, ppm_param_map_ghost_put, info)
  CALL check(info, 'ppm_map_field_ghost ghost_put')
  CALL ppm_map_field_ghost(fg, lda, topo, mesh, gs, ppm_param_map_push, info)
  CALL check(info, 'ppm_map_field_ghost push')
  CALL ppm_map_field_ghost(fg, lda, topo, mesh, gs, ppm_param_map_send, info)
  CALL check(info, 'ppm_map_field_ghost send')
  CALL ppm_map_field_ghost(fg, lda, topo, mesh, gs, ppm_param_map_pop, info)
  CALL check(info, 'ppm_map_field_ghost pop (put)')
  WRITE(20) fg
  CALL ppm_map_field_ghost(fg, lda, topo, mesh, gs, ppm_param_map_ghost_get, info)
  CALL check(info, 'ppm_map_field_ghost ghost_get')
  CALL ppm_map_field_ghost(fg, lda, topo, mesh, gs, ppm_param_map_push, info)
  CALL check(info, 'ppm_map_field_ghost push')
  CALL ppm_map_field_ghost(fg, lda, topo, mesh, gs, ppm_param_map_send, info)
  CALL check(info, 'ppm_map_field_ghost send')
  CALL ppm_map_field_ghost(fg, lda, topo, mesh, gs, ppm_param_map_pop, info)
  CALL check(info, 'ppm_map_field_ghost pop (get)')
  WRITE(20) fg
  CALL ppm_map_field_ghost(fg, lda, topo, mesh, gs, ppm_param_map_pop, info)
  WRITE(20) info

  ! 12. util_sort3d on the (wrapped) shifted positions x2: 5 x 4 x 3 boxes
  nmv = (/ 5, 4, 3 /)
  CALL ppm_util_sort3d(x2, np, minp, maxp, nmv, npbx, info)
  CALL check(info, 'ppm_util_sort3d')
  WRITE(20) x2
  WRITE(20) npbx

  ! 13. cell list of the P2M particles (3D topology made last is the 2D one:
  !     make the 3D one current again), cutoff 0.3 / 0.25 / 0.2, no symmetry
  CALL ppm_mktopo(xq, np, (/ nm, nm, nm /), decomp, ppm_param_assign_internal, minp, maxp, bc, gs, topo2, mesh2, &
                  min_sub, max_sub, cost, sub2proc, nsubs, isublist, nsublist, istart, ndata, info)
  CALL check(info, 'ppm_mktopo (current topology for the cell list)')
  rc = (/ 0.3_MK, 0.25_MK, 0.2_MK /)
  CALL ppm_neighlist_clist(xq, np, rc, .FALSE., clist, nmc, info)
  CALL check(info, 'ppm_neighlist_clist')
  WRITE(20) SHAPE(nmc), nmc
  WRITE(20) SIZE(clist(1)%lpdx), SIZE(clist(1)%lhbx)
  WRITE(20) clist(1)%lpdx, clist(1)%lhbx
  CALL ppm_clist_destroy(clist, info)
  CALL check(info, 'ppm_clist_destroy')
  WRITE(20) ASSOCIATED(clist)

  ! 14. ghost mapping of the other element kinds, with the mask: an integer
  !     vector field (ghost_put, masked), a logical scalar field (ghost_put,
  !     then ghost_get, masked), a double complex vector field on a 2D
  !     topology (ghost_get, masked); values and masks are closed forms of the
  !     node indices
  ALLOCATE(fi(2, 1-g:nm+g, 1-g:nm+g, 1-g:nm+g, 1), fl(1-g:nm+g, 1-g:nm+g, 1-g:nm+g, 1))
  ALLOCATE(msk(1-g:nm+g, 1-g:nm+g, 1-g:nm+g, 1))
  DO k = 1-g, nm+g
    DO j = 1-g, nm+g
      DO i = 1-g, nm+g
        DO c = 1, 2
          fi(c, i, j, k, 1) = MOD(7*i + 13*j + 29*k + 3*c + 1000, 101) - 50
        END DO
        fl(i, j, k, 1) = MOD(i + 2*j + 3*k + 100, 5) /= 0
        msk(i, j, k, 1) = MOD(i*j + k + 100, 3) /= 0
      END DO
    END DO
  END DO
  CALL ppm_map_field_ghost(fi, 2, topo, mesh, gs, ppm_param_map_init, info)
  CALL ppm_map_field_ghost(fi, 2, topo, mesh, gs, ppm_param_map_ghost_put, info)
  CALL ppm_map_field_ghost(fi, 2, topo, mesh, gs, ppm_param_map_push, info)
  CALL ppm_map_field_ghost(fi, 2, topo, mesh, gs, ppm_param_map_send, info)
  CALL ppm_map_field_ghost(fi, 2, topo, mesh, gs, ppm_param_map_pop, info, msk)
  CALL check(info, 'ppm_map_field_ghost 3d_vec_i pop (put, masked)')
  WRITE(20) fi
  CALL ppm_map_field_ghost(fl, topo, mesh, gs, ppm_param_map_ghost_put, info)
  CALL ppm_map_field_ghost(fl, topo, mesh, gs, ppm_param_map_push, info)
  CALL ppm_map_field_ghost(fl, topo, mesh, gs, ppm_param_map_send, info)
  CALL ppm_map_field_ghost(fl, topo, mesh, gs, ppm_param_map_pop, info)
  CALL check(info, 'ppm_map_field_ghost 3d_sca_l pop (put)')
  CALL ppm_map_field_ghost(fl, topo, mesh, gs, ppm_param_map_ghost_get, info)
  CALL ppm_map_field_ghost(fl, topo, mesh, gs, ppm_param_map_push, info)
  CALL ppm_map_field_ghost(fl, topo, mesh, gs, ppm_param_map_send, info)
  CALL ppm_map_field_ghost(fl, topo, mesh, gs, ppm_param_map_pop, info, msk)
  CALL check(info, 'ppm_map_field_ghost 3d_sca_l pop (get, masked)')
  WRITE(20) MERGE(1, 0, fl)
  CALL ppm_mktopo(xp, np, (/ nm, nm /), decomp, ppm_param_assign_internal, minp(1:2), maxp(1:2), bc(1:4), &
                  gs(1:2), topo3, mesh3, min_sub, max_sub, cost, sub2proc, nsubs, isublist, nsublist, istart, ndata, info)
  CALL check(info, 'ppm_mktopo 2d (complex ghost mapping)')
  ALLOCATE(fc(2, 1-g:nm+g, 1-g:nm+g, 1), msk2(1-g:nm+g, 1-g:nm+g, 1))
  DO j = 1-g, nm+g
    DO i = 1-g, nm+g
      DO c = 1, 2
        fc(c, i, j, 1) = CMPLX(REAL(i, MK) + 0.5_MK * c, REAL(j, MK) - 0.25_MK * c, MK)
      END DO
      msk2(i, j, 1) = MOD(i + 3*j + 100, 4) /= 0
    END DO
  END DO
  CALL ppm_map_field_ghost(fc, 2, topo3, mesh3, gs(1:2), ppm_param_map_ghost_get, info)
  CALL ppm_map_field_ghost(fc, 2, topo3, mesh3, gs(1:2), ppm_param_map_push, info)
  CALL ppm_map_field_ghost(fc, 2, topo3, mesh3, gs(1:2), ppm_param_map_send, info)
  CALL ppm_map_field_ghost(fc, 2, topo3, mesh3, gs(1:2), ppm_param_map_pop, info, msk2)
  CALL check(info, 'ppm_map_field_ghost 2d_vec_dc pop (get, masked)')
  WRITE(20) fc

  CLOSE(20)
  CALL ppm_finalize(info)
  PRINT '(A)', 'ppmg_fortran_check: done'

CONTAINS

  SUBROUTINE check(info, what)
    INTEGER, INTENT(IN) :: info
    CHARACTER(LEN=*), INTENT(IN) :: what
    IF (info /= 0) THEN
      PRINT '(A,A,A,I0)', 'ppmg_fortran_check: ', what, ' failed, info = ', info
      STOP 2
    END IF
  END SUBROUTINE check

END PROGRAM ppmg_fortran_check
