!------------------------------------------------------------------------------
! fc_poisson -- ISO_C_BINDING layer over the gfx950 pressure-Poisson library
! (include/fc_poisson.h).  It gives Fortran host code the reference's
! subroutine surface for the hot path:
!
!   fcp_solve_poisson(pPrime, div, p)   <- solve_poisson()            CORE_POISSON.F90:83
!   fcp_mg_solve(var, rr)               <- mg_solve(var, rr)          CORE_MG.F90:1572
!   fcp_ghostcell_pressure(pres, r, p)  <- GCM_ghostcell_pressure     GCM_GHOSTCELL_UPDATE.F90:361
!   fcp_calc_residual(var, r, resm, loc) <- calc_residual             CORE_POISSON.F90:460
!   fcp_calc_residual2(var, p_temp, resm, loc) <- calc_residual2      CORE_POISSON.F90:679
!   fcp_itsolv(var, r)                  <- itsolv(var, r)             CORE_POISSON.F90:205
!   fcp_rhs_poisson(face_u, face_v, face_w, dtinv, div, sumd, resDiv)
!                                       <- rhs_poisson(sumd)          CORE_POISSON.F90:9
!   fcp_correct_vel(pPrime, dt, u, v, w)  <- correct_vel()            CORE_FRAC_STEP_ROUTINES.F90:134
!   fcp_correct_ghostcell_velocity(pVar, dt_fac, u, v, w)
!                                       <- correct_ghostcell_velocity GCM_GHOSTCELL_UPDATE.F90:789
!   fcp_face_vel(u, v, w, bcxu, bcyv, bczw, face_u, face_v, face_w)
!                                       <- face_vel(u_vel,v_vel,w_vel) CORE_FRAC_STEP_ROUTINES.F90:6
!   fcp_update_pressure(pPrime, p)      <- update_pressure()          CORE_FRAC_STEP_ROUTINES.F90:212
!   fcp_set_solve_ad(reinv, dt, cn, omega_ad) <- set_solve_ad()      CORE_SET_SOLVERS.F90:591
!   fcp_itsolv_ad(var, r, face_u, face_v, face_w)
!                                       <- itsolv_ad(var, r)          CORE_SOLVE_AD.F90:131
!   fcp_calc_residual_ad(var, r, face_u, face_v, face_w, resm, loc)
!                                       <- calc_residual_ad           CORE_SOLVE_AD.F90:780
!   fcp_init / fcp_set_masks / fcp_set_gcm / fcp_set_bc               <- MG_initial, mg_prepare,
!                                          GCM_set_internal_boundary outputs
!
! MODULE fc_poisson_ref (fc_poisson_ref.f90) gives the same routines under
! the reference's own names and argument lists.
!
! Arrays are the reference's module arrays (0:nx+1, 0:ny+1, 0:nz+1) passed
! as they are; masks are INTEGER(KIND=1).  The reference's non-convergence
! messages are printed here (rc = 1); errors (rc < 0) stop the run.
!------------------------------------------------------------------------------
module fc_poisson
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: FCP_PREC_FP64 = 0, FCP_PREC_MIXED = 1
  integer, parameter, public :: FCP_OUTER_REFERENCE = 0, FCP_OUTER_RELRES = 1

  type, bind(c), public :: fc_grid_t
     integer(c_int) :: nx, ny, nz
     type(c_ptr)    :: x, y, z
  end type fc_grid_t

  type, bind(c), public :: fc_params_t
     integer(c_int) :: it_solver_type
     real(c_double) :: omega_pson
     integer(c_int) :: itermax_pson
     real(c_double) :: restol_pson
     integer(c_int) :: iterInter, mgrids_x, mgrids_y, mgrids_z
     real(c_double) :: restol_gcp
     integer(c_int) :: itermax_gc, mix_GC_form, is_PBC_homogeneous
     integer(c_int) :: pbc(6)
     real(c_double) :: ppp(6)
     integer(c_int) :: bc_vel(6)
     integer(c_int) :: infoconv, precision, outer_rule
     real(c_double) :: rel_tol
     integer(c_int) :: fp32_cycles, aa_depth, launch_mode, cycle
     integer(c_int) :: vc_walls, coarse_solve
     integer(c_int) :: reserved(3)
  end type fc_params_t

  type, bind(c), public :: fc_stats_t
     integer(c_int) :: outer_iters, vcycles_fp64, vcycles_fp32, refine_steps, gcm_sweeps, converged
     real(c_double) :: maxres, maxres2, rel_resid, gcm_change, time_ms
     real(c_double) :: reserved(6)
  end type fc_stats_t

  interface
     subroutine fc_pson_default_params(p) bind(c, name='fc_pson_default_params')
       import :: fc_params_t
       type(fc_params_t), intent(out) :: p
     end subroutine
     integer(c_int) function fc_pson_create(g, p, h) bind(c, name='fc_pson_create')
       import :: c_int, c_ptr, fc_grid_t, fc_params_t
       type(fc_grid_t), intent(in) :: g
       type(fc_params_t), intent(in) :: p
       type(c_ptr), intent(out) :: h
     end function
     integer(c_int) function fc_pson_destroy(h) bind(c, name='fc_pson_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function
     integer(c_int) function fc_pson_set_params(h, p) bind(c, name='fc_pson_set_params')
       import :: c_int, c_ptr, fc_params_t
       type(c_ptr), value :: h
       type(fc_params_t), intent(in) :: p
     end function
     integer(c_int) function fc_pson_set_masks(h, ib, gh, de) bind(c, name='fc_pson_set_masks')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, ib, gh, de
     end function
     integer(c_int) function fc_pson_set_gcm(h, nG, iG, jG, kG, iC, jC, kC, cN, al, dpG, &
          nD, iD, jD, kD, iDC, jDC, kDC, cD, plD, dpD, ipD) bind(c, name='fc_pson_set_gcm')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
       integer(c_int), value :: nG, nD
       type(c_ptr), value :: iG, jG, kG, iC, jC, kC, cN, al, dpG, iD, jD, kD, iDC, jDC, kDC, cD, plD, dpD, ipD
     end function
     integer(c_int) function fc_pson_set_bc(h, px1, px2, py1, py2, pz1, pz2, area) bind(c, name='fc_pson_set_bc')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h, px1, px2, py1, py2, pz1, pz2
       real(c_double), value :: area
     end function
     integer(c_int) function fc_pson_get_bc(h, px1, px2, py1, py2, pz1, pz2) bind(c, name='fc_pson_get_bc')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, px1, px2, py1, py2, pz1, pz2
     end function
     integer(c_int) function fc_pson_solve(h, pp, dv, pv, rl, st) bind(c, name='fc_pson_solve')
       import :: c_int, c_ptr, fc_stats_t
       type(c_ptr), value :: h, pp, dv, pv, rl
       type(fc_stats_t), intent(out) :: st
     end function
     integer(c_int) function fc_pson_mg_solve(h, var, rr, st) bind(c, name='fc_pson_mg_solve')
       import :: c_int, c_ptr, fc_stats_t
       type(c_ptr), value :: h, var, rr
       type(fc_stats_t), intent(out) :: st
     end function
     integer(c_int) function fc_pson_ghostcell_pressure(h, pres, r, pv, rl, st) bind(c, name='fc_pson_ghostcell_pressure')
       import :: c_int, c_ptr, fc_stats_t
       type(c_ptr), value :: h, pres, r, pv, rl
       type(fc_stats_t), intent(out) :: st
     end function
     integer(c_int) function fc_pson_calc_residual(h, var, r, resm, loc) bind(c, name='fc_pson_calc_residual')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h, var, r
       real(c_double), intent(out) :: resm
       integer(c_int), intent(out) :: loc(3)
     end function
     integer(c_int) function fc_pson_calc_residual2(h, var, pt, resm, loc) bind(c, name='fc_pson_calc_residual2')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h, var, pt
       real(c_double), intent(out) :: resm
       integer(c_int), intent(out) :: loc(3)
     end function
     integer(c_int) function fc_pson_itsolv(h, var, r) bind(c, name='fc_pson_itsolv')
       import :: c_int, c_ptr
       type(c_ptr), value :: h, var, r
     end function
     integer(c_int) function fc_pson_rhs(h, fu, fv, fw, dtinv, div, sumd, resdiv) bind(c, name='fc_pson_rhs')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h, fu, fv, fw, div
       real(c_double), value :: dtinv
       real(c_double), intent(out) :: sumd, resdiv
     end function
     integer(c_int) function fc_vel_set_bc(h, bu, bv, bw) bind(c, name='fc_vel_set_bc')
       import :: c_ptr, c_int
       type(c_ptr), value :: h, bu, bv, bw
     end function
     integer(c_int) function fc_vel_get_bc(h, bu, bv, bw) bind(c, name='fc_vel_get_bc')
       import :: c_ptr, c_int
       type(c_ptr), value :: h, bu, bv, bw
     end function
     integer(c_int) function fc_pson_get_outflow(h, area, corr) bind(c, name='fc_pson_get_outflow')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: area, corr
     end function
     integer(c_int) function fc_correct_vel(h, pp, dt, u, v, w) bind(c, name='fc_correct_vel')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h, pp, u, v, w
       real(c_double), value :: dt
     end function
     integer(c_int) function fc_correct_ghostcell_velocity(h, pv, dt, u, v, w) &
          bind(c, name='fc_correct_ghostcell_velocity')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h, pv, u, v, w
       real(c_double), value :: dt
     end function
     integer(c_int) function fc_face_vel(h, u, v, w, fu, fv, fw) bind(c, name='fc_face_vel')
       import :: c_ptr, c_int
       type(c_ptr), value :: h, u, v, w, fu, fv, fw
     end function
     integer(c_int) function fc_update_pressure(h, pp, p) bind(c, name='fc_update_pressure')
       import :: c_ptr, c_int
       type(c_ptr), value :: h, pp, p
     end function
     integer(c_int) function fc_set_solve_ad(h, reinv, dt, cn, om) bind(c, name='fc_set_solve_ad')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h
       real(c_double), value :: reinv, dt, om
       integer(c_int), value :: cn
     end function
     integer(c_int) function fc_itsolv_ad(h, var, r, fu, fv, fw, ns) bind(c, name='fc_itsolv_ad')
       import :: c_ptr, c_int
       type(c_ptr), value :: h, var, r, fu, fv, fw
       integer(c_int), value :: ns
     end function
     integer(c_int) function fc_calc_residual_ad(h, var, r, fu, fv, fw, resm, loc, r2) &
          bind(c, name='fc_calc_residual_ad')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h, var, r, fu, fv, fw, r2
       real(c_double) :: resm
       integer(c_int) :: loc(3)
     end function
     integer(c_int) function fc_vel_set_outer(h, vals, freq, y0, y1) bind(c, name='fc_vel_set_outer')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: vals(18), freq(18)
       real(c_double), value :: y0, y1
     end function
     integer(c_int) function fc_set_gcm_vel(h, cg, cd, ubi, ubid, pl) bind(c, name='fc_set_gcm_vel')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h, cg, cd, ubi, ubid
       real(c_double), value :: pl
     end function
     integer(c_int) function fc_solve_ad(h, u, v, w, nlu, nlv, nlw, fu, fv, fw, time, itmax, rtol, gtol, &
          iters, res, loc) bind(c, name='fc_solve_ad')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h, u, v, w, nlu, nlv, nlw, fu, fv, fw
       real(c_double), value :: time, rtol, gtol
       integer(c_int), value :: itmax
       integer(c_int), intent(out) :: iters, loc(3)
       real(c_double), intent(out) :: res
     end function
     integer(c_int) function fc_set_visc_tot(h, vt) bind(c, name='fc_set_visc_tot')
       import :: c_ptr, c_int
       type(c_ptr), value :: h, vt
     end function
     integer(c_int) function fc_rhs_advec_diff(h, uo, vo, wo, fu, fv, fw, alfa, dt, nlu, nlv, nlw) &
          bind(c, name='fc_rhs_advec_diff')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: h, uo, vo, wo, fu, fv, fw, nlu, nlv, nlw
       real(c_double), value :: alfa, dt
     end function
     type(c_ptr) function fc_pson_last_error() bind(c, name='fc_pson_last_error')
       import :: c_ptr
     end function
  end interface

  type(c_ptr), save :: handle = c_null_ptr
  type(fc_params_t), save, public :: fcp_params
  type(fc_stats_t), save, public :: fcp_last_stats

  public :: fcp_default_params, fcp_init, fcp_finalize, fcp_set_params, fcp_set_masks, fcp_set_gcm, fcp_set_bc
  public :: fcp_get_bc
  public :: fcp_solve_poisson, fcp_mg_solve, fcp_ghostcell_pressure, fcp_calc_residual
  public :: fcp_calc_residual2, fcp_itsolv, fcp_rhs_poisson, fcp_get_outflow, fcp_vel_get_bc
  public :: fcp_correct_vel, fcp_correct_ghostcell_velocity, fcp_face_vel, fcp_update_pressure
  public :: fcp_set_solve_ad, fcp_itsolv_ad, fcp_calc_residual_ad
  public :: fcp_vel_set_outer, fcp_set_gcm_vel, fcp_solve_ad, fcp_rhs_advec_diff

contains

  subroutine fcp_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc >= 0) return
    call c_f_pointer(fc_pson_last_error(), msg, [512])
    n = 0
    do while (n < 512)
       if (msg(n + 1) == c_null_char) exit
       n = n + 1
    end do
    write(*, '(3a,i0,a,512a)') 'fc_poisson: ', what, ' failed (', rc, '): ', msg(1:n)
    error stop 'fc_poisson error'
  end subroutine fcp_check

  subroutine fcp_default_params(p)
    type(fc_params_t), intent(out) :: p
    call fc_pson_default_params(p)
  end subroutine fcp_default_params

  ! MG_initial: builds the hierarchy for the grid x(0:nx+1), y, z
  subroutine fcp_init(nx, ny, nz, x, y, z, p)
    integer, intent(in) :: nx, ny, nz
    real(c_double), intent(in), target :: x(0:nx+1), y(0:ny+1), z(0:nz+1)
    type(fc_params_t), intent(in) :: p
    type(fc_grid_t) :: g
    g%nx = nx; g%ny = ny; g%nz = nz
    g%x = c_loc(x); g%y = c_loc(y); g%z = c_loc(z)
    fcp_params = p
    call fcp_check(fc_pson_create(g, p, handle), 'fc_pson_create')
  end subroutine fcp_init

  subroutine fcp_finalize()
    integer(c_int) :: rc
    if (c_associated(handle)) rc = fc_pson_destroy(handle)
    handle = c_null_ptr
  end subroutine fcp_finalize

  subroutine fcp_set_params(p)
    type(fc_params_t), intent(in) :: p
    fcp_params = p
    call fcp_check(fc_pson_set_params(handle, p), 'fc_pson_set_params')
  end subroutine fcp_set_params

  ! iblank, ghostCellMark, dead_cell (INTEGER(KIND=1)) -> mg_prepare
  subroutine fcp_set_masks(iblank, ghostCellMark, dead_cell)
    integer(c_int8_t), intent(in), target, contiguous :: iblank(:,:,:), ghostCellMark(:,:,:), dead_cell(:,:,:)
    call fcp_check(fc_pson_set_masks(handle, c_loc(iblank), c_loc(ghostCellMark), c_loc(dead_cell)), &
         'fc_pson_set_masks')
  end subroutine fcp_set_masks

  subroutine fcp_set_gcm(nGhost, iGhost, jGhost, kGhost, iCellIndex, jCellIndex, kCellIndex, coeffGCMN, &
       alphaGhost, dpdnBodyIntercept, nDead, iDead, jDead, kDead, iDeadCellIndex, jDeadCellIndex, &
       kDeadCellIndex, coeffGCMDeadN, probeLengthDead, dpdnInterceptDead, xyzDCIP)
    integer, intent(in) :: nGhost, nDead
    integer(c_int), intent(in), target :: iGhost(*), jGhost(*), kGhost(*), iCellIndex(*), jCellIndex(*), kCellIndex(*)
    integer(c_int), intent(in), target :: iDead(*), jDead(*), kDead(*), iDeadCellIndex(*), jDeadCellIndex(*), &
         kDeadCellIndex(*)
    real(c_double), intent(in), target :: coeffGCMN(8,*), alphaGhost(*), dpdnBodyIntercept(*)
    real(c_double), intent(in), target :: coeffGCMDeadN(8,*), probeLengthDead(*), dpdnInterceptDead(*), xyzDCIP(3,*)
    call fcp_check(fc_pson_set_gcm(handle, int(nGhost, c_int), c_loc(iGhost), c_loc(jGhost), c_loc(kGhost), &
         c_loc(iCellIndex), c_loc(jCellIndex), c_loc(kCellIndex), c_loc(coeffGCMN), c_loc(alphaGhost), &
         c_loc(dpdnBodyIntercept), int(nDead, c_int), c_loc(iDead), c_loc(jDead), c_loc(kDead), &
         c_loc(iDeadCellIndex), c_loc(jDeadCellIndex), c_loc(kDeadCellIndex), c_loc(coeffGCMDeadN), &
         c_loc(probeLengthDead), c_loc(dpdnInterceptDead), c_loc(xyzDCIP)), 'fc_pson_set_gcm')
  end subroutine fcp_set_gcm

  subroutine fcp_set_bc(pgradx1, pgradx2, pgrady1, pgrady2, pgradz1, pgradz2, outflow_area)
    real(c_double), intent(in), target, contiguous :: pgradx1(:,:), pgradx2(:,:), pgrady1(:,:), pgrady2(:,:), &
         pgradz1(:,:), pgradz2(:,:)
    real(c_double), intent(in) :: outflow_area
    call fcp_check(fc_pson_set_bc(handle, c_loc(pgradx1), c_loc(pgradx2), c_loc(pgrady1), c_loc(pgrady2), &
         c_loc(pgradz1), c_loc(pgradz2), outflow_area), 'fc_pson_set_bc')
  end subroutine fcp_set_bc

  ! the pgrad planes as the solve left them (GCM_enforce_p_compatibility
  ! corrects those of zero-gradient faces)
  subroutine fcp_get_bc(pgradx1, pgradx2, pgrady1, pgrady2, pgradz1, pgradz2)
    real(c_double), intent(out), target, contiguous :: pgradx1(:,:), pgradx2(:,:), pgrady1(:,:), pgrady2(:,:), &
         pgradz1(:,:), pgradz2(:,:)
    call fcp_check(fc_pson_get_bc(handle, c_loc(pgradx1), c_loc(pgradx2), c_loc(pgrady1), c_loc(pgrady2), &
         c_loc(pgradz1), c_loc(pgradz2)), 'fc_pson_get_bc')
  end subroutine fcp_get_bc

  ! solve_poisson(): pPrime in/out, div in, p = previous-step pressure
  subroutine fcp_solve_poisson(pPrime, div, p)
    real(c_double), intent(inout), target, contiguous :: pPrime(:,:,:)
    real(c_double), intent(in), target, contiguous :: div(:,:,:), p(:,:,:)
    integer(c_int) :: rc
    rc = fc_pson_solve(handle, c_loc(pPrime), c_loc(div), c_loc(p), c_null_ptr, fcp_last_stats)
    call fcp_check(rc, 'solve_poisson')
    if (rc == 1) then
       print*, 'Pressure did not converge in ', fcp_params%itermax_pson, ' iterations'
       print*, 'Final residual = ', fcp_last_stats%maxres
    end if
  end subroutine fcp_solve_poisson

  ! mg_solve(var, rr)
  subroutine fcp_mg_solve(var, rr)
    real(c_double), intent(inout), target, contiguous :: var(:,:,:)
    real(c_double), intent(in), target, contiguous :: rr(:,:,:)
    call fcp_check(fc_pson_mg_solve(handle, c_loc(var), c_loc(rr), fcp_last_stats), 'mg_solve')
  end subroutine fcp_mg_solve

  ! GCM_ghostcell_pressure(pres, r) with the previous-step pressure p
  subroutine fcp_ghostcell_pressure(pres, r, p)
    real(c_double), intent(inout), target, contiguous :: pres(:,:,:)
    real(c_double), intent(in), target, contiguous :: r(:,:,:), p(:,:,:)
    integer(c_int) :: rc
    rc = fc_pson_ghostcell_pressure(handle, c_loc(pres), c_loc(r), c_loc(p), c_null_ptr, fcp_last_stats)
    call fcp_check(rc, 'GCM_ghostcell_pressure')
    if (rc == 1) then
       print*, 'GhostCell Pressure did not converge in ', fcp_params%itermax_gc, ' iterations'
       print*, 'Final residual = ', fcp_last_stats%gcm_change
    end if
  end subroutine fcp_ghostcell_pressure

  ! calc_residual(var, r, resm, loc)
  subroutine fcp_calc_residual(var, r, resm, loc)
    real(c_double), intent(in), target, contiguous :: var(:,:,:), r(:,:,:)
    real(c_double), intent(out) :: resm
    integer, intent(out), optional :: loc(3)
    integer(c_int) :: l(3)
    call fcp_check(fc_pson_calc_residual(handle, c_loc(var), c_loc(r), resm, l), 'calc_residual')
    if (present(loc)) loc = int(l)
  end subroutine fcp_calc_residual

  ! calc_residual2(var, r, resm, loc): max |var - p_temp| over iblank = 0
  subroutine fcp_calc_residual2(var, p_temp, resm, loc)
    real(c_double), intent(in), target, contiguous :: var(:,:,:), p_temp(:,:,:)
    real(c_double), intent(out) :: resm
    integer, intent(out), optional :: loc(3)
    integer(c_int) :: l(3)
    call fcp_check(fc_pson_calc_residual2(handle, c_loc(var), c_loc(p_temp), resm, l), 'calc_residual2')
    if (present(loc)) loc = int(l)
  end subroutine fcp_calc_residual2

  ! itsolv(var, r): one plain line-SOR pass
  subroutine fcp_itsolv(var, r)
    real(c_double), intent(inout), target, contiguous :: var(:,:,:)
    real(c_double), intent(in), target, contiguous :: r(:,:,:)
    call fcp_check(fc_pson_itsolv(handle, c_loc(var), c_loc(r)), 'itsolv')
  end subroutine fcp_itsolv

  ! rhs_poisson(sumd): div from the face velocities, dtinv = 1/dt; with
  ! zero-gradient faces the face velocities are corrected in place
  ! (enforce_global_mass_consv)
  subroutine fcp_rhs_poisson(face_u, face_v, face_w, dtinv, div, sumd, resDiv)
    real(c_double), intent(inout), target, contiguous :: face_u(:,:,:), face_v(:,:,:), face_w(:,:,:)
    real(c_double), intent(in) :: dtinv
    real(c_double), intent(inout), target, contiguous :: div(:,:,:)
    real(c_double), intent(out) :: sumd, resDiv
    call fcp_check(fc_pson_rhs(handle, c_loc(face_u), c_loc(face_v), c_loc(face_w), dtinv, c_loc(div), sumd, &
         resDiv), 'rhs_poisson')
  end subroutine fcp_rhs_poisson

  ! outflow_area and correction_vel of the last rhs_poisson
  subroutine fcp_get_outflow(outflow_area, correction_vel)
    real(c_double), intent(out) :: outflow_area, correction_vel
    call fcp_check(fc_pson_get_outflow(handle, outflow_area, correction_vel), 'get_outflow')
  end subroutine fcp_get_outflow

  ! bcxu(0:1,:,:), bcyv(:,0:1,:), bczw(:,:,0:1) as rhs_poisson left them
  subroutine fcp_vel_get_bc(bcxu, bcyv, bczw)
    real(c_double), intent(out), target, contiguous :: bcxu(:,:,:), bcyv(:,:,:), bczw(:,:,:)
    call fcp_check(fc_vel_get_bc(handle, c_loc(bcxu), c_loc(bcyv), c_loc(bczw)), 'fc_vel_get_bc')
  end subroutine fcp_vel_get_bc

  ! correct_vel(): u, v, w -= dt * grad pPrime on the non-solid interior cells
  subroutine fcp_correct_vel(pPrime, dt, u, v, w)
    real(c_double), intent(in), target, contiguous :: pPrime(:,:,:)
    real(c_double), intent(in) :: dt
    real(c_double), intent(inout), target, contiguous :: u(:,:,:), v(:,:,:), w(:,:,:)
    call fcp_check(fc_correct_vel(handle, c_loc(pPrime), dt, c_loc(u), c_loc(v), c_loc(w)), 'correct_vel')
  end subroutine fcp_correct_vel

  ! correct_ghostcell_velocity(pVar); dt_fac = dt (2*rk_bet(rk_step)*dt under RK3)
  subroutine fcp_correct_ghostcell_velocity(pVar, dt_fac, u, v, w)
    real(c_double), intent(in), target, contiguous :: pVar(:,:,:)
    real(c_double), intent(in) :: dt_fac
    real(c_double), intent(inout), target, contiguous :: u(:,:,:), v(:,:,:), w(:,:,:)
    call fcp_check(fc_correct_ghostcell_velocity(handle, c_loc(pVar), dt_fac, c_loc(u), c_loc(v), c_loc(w)), &
         'correct_ghostcell_velocity')
  end subroutine fcp_correct_ghostcell_velocity

  ! face_vel(u_vel, v_vel, w_vel) with the outer-face arrays bcxu(0:1,:,:),
  ! bcyv(:,0:1,:), bczw(:,:,0:1); face_u, face_v, face_w in/out
  subroutine fcp_face_vel(u, v, w, bcxu, bcyv, bczw, face_u, face_v, face_w)
    real(c_double), intent(in), target, contiguous :: u(:,:,:), v(:,:,:), w(:,:,:)
    real(c_double), intent(in), target, contiguous :: bcxu(:,:,:), bcyv(:,:,:), bczw(:,:,:)
    real(c_double), intent(inout), target, contiguous :: face_u(:,:,:), face_v(:,:,:), face_w(:,:,:)
    call fcp_check(fc_vel_set_bc(handle, c_loc(bcxu), c_loc(bcyv), c_loc(bczw)), 'fc_vel_set_bc')
    call fcp_check(fc_face_vel(handle, c_loc(u), c_loc(v), c_loc(w), c_loc(face_u), c_loc(face_v), &
         c_loc(face_w)), 'face_vel')
  end subroutine fcp_face_vel

  ! update_pressure(): pPrime in/out, p out
  subroutine fcp_update_pressure(pPrime, p)
    real(c_double), intent(inout), target, contiguous :: pPrime(:,:,:), p(:,:,:)
    call fcp_check(fc_update_pressure(handle, c_loc(pPrime), c_loc(p)), 'update_pressure')
  end subroutine fcp_update_pressure

  ! set_solve_ad() with nu = reinv; cn = 1 when advec_scheme == CRANK_NICOLSON2
  subroutine fcp_set_solve_ad(reinv, dt, cn, omega_ad)
    real(c_double), intent(in) :: reinv, dt, omega_ad
    integer, intent(in) :: cn
    call fcp_check(fc_set_solve_ad(handle, reinv, dt, int(cn, c_int), omega_ad), 'set_solve_ad')
  end subroutine fcp_set_solve_ad

  ! itsolv_ad(var, r) with the current face velocities
  subroutine fcp_itsolv_ad(var, r, face_u, face_v, face_w)
    real(c_double), intent(inout), target, contiguous :: var(:,:,:)
    real(c_double), intent(in), target, contiguous :: r(:,:,:), face_u(:,:,:), face_v(:,:,:), face_w(:,:,:)
    call fcp_check(fc_itsolv_ad(handle, c_loc(var), c_loc(r), c_loc(face_u), c_loc(face_v), c_loc(face_w), &
         1_c_int), 'itsolv_ad')
  end subroutine fcp_itsolv_ad

  ! calc_residual_ad(var, r, resm, loc)
  subroutine fcp_calc_residual_ad(var, r, face_u, face_v, face_w, resm, loc)
    real(c_double), intent(in), target, contiguous :: var(:,:,:), r(:,:,:), face_u(:,:,:), face_v(:,:,:), &
         face_w(:,:,:)
    real(c_double), intent(out) :: resm
    integer, intent(out) :: loc(3)
    integer(c_int) :: l(3)
    call fcp_check(fc_calc_residual_ad(handle, c_loc(var), c_loc(r), c_loc(face_u), c_loc(face_v), c_loc(face_w), &
         resm, l, c_null_ptr), 'calc_residual_ad')
    loc = int(l)
  end subroutine fcp_calc_residual_ad

  ! The inputs of set_outer_velocity_bc: ux1, ux2, vx1, vx2, wx1, wx2, uy1,
  ! ..., wz2, the frequencies of the pulsatile faces (same order), yOrigin
  ! and yout of the user-specified channel profile
  subroutine fcp_vel_set_outer(vals, freq, yOrigin, yout)
    real(c_double), intent(in) :: vals(18), freq(18), yOrigin, yout
    call fcp_check(fc_vel_set_outer(handle, vals, freq, yOrigin, yout), 'fc_vel_set_outer')
  end subroutine fcp_vel_set_outer

  ! GCM_ghostcell_velocity's relations: coeffGCMD(nGhost,8),
  ! coeffGCMDeadD(nDead,8), body-intercept velocities uvwBI(nGhost,3),
  ! uvwBIDead(nDead,3), probeLengthNormalizedD
  subroutine fcp_set_gcm_vel(coeffGCMD, coeffGCMDeadD, uvwBI, uvwBIDead, probeLengthNormalizedD)
    real(c_double), intent(in), target, contiguous :: coeffGCMD(:,:), coeffGCMDeadD(:,:), uvwBI(:,:), uvwBIDead(:,:)
    real(c_double), intent(in) :: probeLengthNormalizedD
    call fcp_check(fc_set_gcm_vel(handle, c_loc(coeffGCMD), c_loc(coeffGCMDeadD), c_loc(uvwBI), c_loc(uvwBIDead), &
         probeLengthNormalizedD), 'set_gcm_vel')
  end subroutine fcp_set_gcm_vel

  ! solve_ad(): u, v, w and face_u, face_v, face_w in/out, nlu, nlv, nlw in
  subroutine fcp_solve_ad(u, v, w, nlu, nlv, nlw, face_u, face_v, face_w, time, itermax_ad, restol_ad, &
       restol_gcv)
    real(c_double), intent(inout), target, contiguous :: u(:,:,:), v(:,:,:), w(:,:,:)
    real(c_double), intent(in), target, contiguous :: nlu(:,:,:), nlv(:,:,:), nlw(:,:,:)
    real(c_double), intent(inout), target, contiguous :: face_u(:,:,:), face_v(:,:,:), face_w(:,:,:)
    real(c_double), intent(in) :: time, restol_ad, restol_gcv
    integer, intent(in) :: itermax_ad
    integer(c_int) :: rc, iters, l(3)
    real(c_double) :: res
    rc = fc_solve_ad(handle, c_loc(u), c_loc(v), c_loc(w), c_loc(nlu), c_loc(nlv), c_loc(nlw), c_loc(face_u), &
         c_loc(face_v), c_loc(face_w), time, int(itermax_ad, c_int), restol_ad, restol_gcv, iters, res, l)
    call fcp_check(rc, 'solve_ad')
    if (rc == 1) then
       print*, 'Velocity did not converge in ', itermax_ad, ' iterations'
       print*, 'Final residual = ', res
    end if
  end subroutine fcp_solve_ad

  ! rhs_advec_diff(): nlu, nlv, nlw (interior cells) from uOld, vOld, wOld
  ! and the face velocities; viscTot defaults to the reinv of
  ! fcp_set_solve_ad
  subroutine fcp_rhs_advec_diff(uOld, vOld, wOld, face_u, face_v, face_w, alfa_upwind, dt, nlu, nlv, nlw, viscTot)
    real(c_double), intent(in), target, contiguous :: uOld(:,:,:), vOld(:,:,:), wOld(:,:,:)
    real(c_double), intent(in), target, contiguous :: face_u(:,:,:), face_v(:,:,:), face_w(:,:,:)
    real(c_double), intent(in) :: alfa_upwind, dt
    real(c_double), intent(inout), target, contiguous :: nlu(:,:,:), nlv(:,:,:), nlw(:,:,:)
    real(c_double), intent(in), target, contiguous, optional :: viscTot(:,:,:)
    if (present(viscTot)) then
       call fcp_check(fc_set_visc_tot(handle, c_loc(viscTot)), 'set_visc_tot')
    else
       call fcp_check(fc_set_visc_tot(handle, c_null_ptr), 'set_visc_tot')
    end if
    call fcp_check(fc_rhs_advec_diff(handle, c_loc(uOld), c_loc(vOld), c_loc(wOld), c_loc(face_u), c_loc(face_v), &
         c_loc(face_w), alfa_upwind, dt, c_loc(nlu), c_loc(nlv), c_loc(nlw)), 'rhs_advec_diff')
  end subroutine fcp_rhs_advec_diff

end module fc_poisson
