!------------------------------------------------------------------------------
! fc_poisson_ref -- the hot path's subroutines under the reference's own
! names and argument lists, for a flow solver that replaces them by
! "USE fc_poisson_ref" (the reference's definitions are then left out of the
! build):
!
!   solve_poisson()                      CORE_POISSON.F90:83-200
!   rhs_poisson(sumd)                    CORE_POISSON.F90:9-79
!   itsolv(var, r)                       CORE_POISSON.F90:205-456
!   calc_residual(var, r, resm, loc)     CORE_POISSON.F90:460-551
!   calc_residual2(var, r, resm, loc)    CORE_POISSON.F90:679-724
!   mg_solve(var, rr)                    CORE_MG.F90:1572-1848
!   GCM_ghostcell_pressure(pres, r)      GCM_GHOSTCELL_UPDATE.F90:361-739
!
! The reference's versions read module arrays (pPrime, div, p, p_temp,
! face_u/v/w, dt).  Here those arrays are bound once with fcp_bind_flow (the
! arrays must stay allocated, and at the same address, while bound: the
! reference's module allocatables are allocated once, CORE_MEMORY_ALLOCATE.
! F90:6-120); the grid, masks and GCM lists go in through MODULE fc_poisson
! (fcp_init, fcp_set_masks, fcp_set_gcm), as MG_initial, mg_prepare and
! GCM_set_internal_boundary leave them.
!------------------------------------------------------------------------------
module fc_poisson_ref
  use iso_c_binding
  use fc_poisson
  implicit none
  private

  real(c_double), pointer, contiguous :: b_pPrime(:,:,:) => null(), b_div(:,:,:) => null(), b_p(:,:,:) => null()
  real(c_double), pointer, contiguous :: b_ptemp(:,:,:) => null()
  real(c_double), pointer, contiguous :: b_fu(:,:,:) => null(), b_fv(:,:,:) => null(), b_fw(:,:,:) => null()
  real(c_double), save :: b_dt = 1.0_c_double
  real(c_double), save, public :: resDiv = 0.0_c_double

  public :: fcp_bind_flow, solve_poisson, rhs_poisson, itsolv, calc_residual, calc_residual2, mg_solve
  public :: GCM_ghostcell_pressure

contains

  subroutine need(ok, what)
    logical, intent(in) :: ok
    character(len=*), intent(in) :: what
    if (.not. ok) then
       write(*, '(3a)') 'fc_poisson_ref: ', what, ' is not bound (fcp_bind_flow)'
       error stop 'fc_poisson_ref'
    end if
  end subroutine need

  ! the reference's module arrays the routines read: pPrime, div, p
  ! (previous-step pressure); optionally p_temp, the face velocities and dt
  subroutine fcp_bind_flow(pPrime, div, p, p_temp, face_u, face_v, face_w, dt)
    real(c_double), intent(inout), target, contiguous :: pPrime(:,:,:), div(:,:,:), p(:,:,:)
    real(c_double), intent(inout), target, contiguous, optional :: p_temp(:,:,:), face_u(:,:,:), face_v(:,:,:), &
         face_w(:,:,:)
    real(c_double), intent(in), optional :: dt
    b_pPrime => pPrime
    b_div => div
    b_p => p
    if (present(p_temp)) b_ptemp => p_temp
    if (present(face_u)) b_fu => face_u
    if (present(face_v)) b_fv => face_v
    if (present(face_w)) b_fw => face_w
    if (present(dt)) b_dt = dt
  end subroutine fcp_bind_flow

  subroutine solve_poisson()
    call need(associated(b_pPrime), 'pPrime')
    call fcp_solve_poisson(b_pPrime, b_div, b_p)
  end subroutine solve_poisson

  subroutine rhs_poisson(sumd)
    real(c_double), intent(out) :: sumd
    call need(associated(b_fu) .and. associated(b_fv) .and. associated(b_fw), 'face_u/v/w')
    call fcp_rhs_poisson(b_fu, b_fv, b_fw, 1.0_c_double / b_dt, b_div, sumd, resDiv)
  end subroutine rhs_poisson

  subroutine itsolv(var, r)
    real(c_double), intent(inout), contiguous :: var(:,:,:)
    real(c_double), intent(in), contiguous :: r(:,:,:)
    call fcp_itsolv(var, r)
  end subroutine itsolv

  subroutine calc_residual(var, r, resm, loc)
    real(c_double), intent(in), contiguous :: var(:,:,:), r(:,:,:)
    real(c_double), intent(out) :: resm
    integer, intent(out) :: loc(3)
    call fcp_calc_residual(var, r, resm, loc)
  end subroutine calc_residual

  ! as the reference, reads the bound p_temp and not r
  subroutine calc_residual2(var, r, resm, loc)
    real(c_double), intent(in), contiguous :: var(:,:,:), r(:,:,:)
    real(c_double), intent(out) :: resm
    integer, intent(out) :: loc(3)
    call need(associated(b_ptemp), 'p_temp')
    call fcp_calc_residual2(var, b_ptemp, resm, loc)
  end subroutine calc_residual2

  subroutine mg_solve(var, rr)
    real(c_double), intent(inout), contiguous :: var(:,:,:)
    real(c_double), intent(in), contiguous :: rr(:,:,:)
    call fcp_mg_solve(var, rr)
  end subroutine mg_solve

  subroutine GCM_ghostcell_pressure(pres, r)
    real(c_double), intent(inout), contiguous :: pres(:,:,:)
    real(c_double), intent(in), contiguous :: r(:,:,:)
    call need(associated(b_p), 'p')
    call fcp_ghostcell_pressure(pres, r, b_p)
  end subroutine GCM_ghostcell_pressure

end module fc_poisson_ref
