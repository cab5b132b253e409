!------------------------------------------------------------------------------
! fc_step_driver -- a Fortran host program that runs the calls around the
! pressure solve of a time step through MODULE fc_poisson: the projection
! (correct_vel, correct_ghostcell_velocity, face_vel, update_pressure;
! CORE_TIME_STEP.F90:211-222), one itsolv_ad / calc_residual_ad pair of the
! momentum step (CORE_SOLVE_AD.F90) and rhs_advec_diff (CORE_SET_SOLVERS.F90)
! of the projected velocities.  Inputs come from integer formulas, so a
! checker can rebuild them exactly:
!   grid  x(i) = (i-1)/(n-1), ghost nodes mirrored (make_grid)
!   field s at (i,j,k) = mod(7i + 13j + 29k + 31s, 17)/17 - 0.5
!   iblank = 1 where mod(i + 2j + 3k, 11) = 0; dead_cell = -1 on those with
!   i + j + k even
! Usage: fc_step_driver <n> <result file>; the result is a stream file of
! u, v, w, face_u, face_v, face_w, pPrime, p (after the projection), var
! (after itsolv_ad), resm and loc(3) (calc_residual_ad), then nlu, nlv, nlw
! (rhs_advec_diff of u, v, w with alfa_upwind 1/2, starting from the fields
! 8, 9, 10 outside the interior).
!------------------------------------------------------------------------------
program fc_step_driver
  use iso_c_binding
  use fc_poisson
  implicit none
  character(len=1024) :: arg, fout
  integer :: n, i, j, k, iu, loc(3)
  real(c_double), allocatable, target :: x(:), y(:), z(:)
  real(c_double), allocatable, target :: P(:,:,:), pn(:,:,:), u(:,:,:), v(:,:,:), w(:,:,:)
  real(c_double), allocatable, target :: fu(:,:,:), fv(:,:,:), fw(:,:,:), nlv(:,:,:), var(:,:,:)
  real(c_double), allocatable, target :: n1(:,:,:), n2(:,:,:), n3(:,:,:)
  real(c_double), allocatable, target :: bcxu(:,:,:), bcyv(:,:,:), bczw(:,:,:)
  integer(c_int8_t), allocatable, target :: ib(:,:,:), gh(:,:,:), de(:,:,:)
  integer(c_int), target :: idum(1)
  real(c_double), target :: d8(8, 1), d1(1), d3(3, 1)
  type(fc_params_t) :: prm
  real(c_double) :: resm

  call get_command_argument(1, arg)
  read(arg, *) n
  call get_command_argument(2, fout)
  allocate(x(0:n+1), y(0:n+1), z(0:n+1))
  do i = 1, n
    x(i) = real(i - 1, c_double) / real(n - 1, c_double)
  end do
  x(0) = 2.0_c_double * x(1) - x(2)
  x(n+1) = 2.0_c_double * x(n) - x(n-1)
  y = x
  z = x
  allocate(ib(0:n+1,0:n+1,0:n+1), gh(0:n+1,0:n+1,0:n+1), de(0:n+1,0:n+1,0:n+1))
  ib = 0; gh = 0; de = 0
  do k = 0, n+1
    do j = 0, n+1
      do i = 0, n+1
        if (mod(i + 2*j + 3*k, 11) == 0) then
          ib(i,j,k) = 1
          if (mod(i + j + k, 2) == 0) de(i,j,k) = -1
        end if
      end do
    end do
  end do
  call field(P, 0); call field(u, 1); call field(v, 2); call field(w, 3)
  call field(fu, 4); call field(fv, 5); call field(fw, 6); call field(nlv, 7)
  allocate(pn(0:n+1,0:n+1,0:n+1))
  pn = 0.0_c_double
  allocate(bcxu(0:1,0:n+1,0:n+1), bcyv(0:n+1,0:1,0:n+1), bczw(0:n+1,0:n+1,0:1))
  do k = 0, n+1
    do j = 0, n+1
      do i = 0, 1
        bcxu(i,j,k) = real(mod(3*i + 5*j + 11*k, 13), c_double) / 13.0_c_double
        bcyv(j,i,k) = real(mod(5*i + 3*j + 11*k, 13), c_double) / 13.0_c_double
        bczw(j,k,i) = real(mod(11*i + 5*j + 3*k, 13), c_double) / 13.0_c_double
      end do
    end do
  end do

  call fcp_default_params(prm)
  call fcp_init(n, n, n, x, y, z, prm)
  call fcp_set_masks(ib, gh, de)
  idum = 0; d8 = 0.0_c_double; d1 = 0.0_c_double; d3 = 0.0_c_double
  call fcp_set_gcm(0, idum, idum, idum, idum, idum, idum, d8, d1, d1, 0, idum, idum, idum, idum, idum, idum, &
       d8, d1, d1, d3)

  ! CORE_TIME_STEP.F90:211-222
  call fcp_correct_vel(P, 0.01_c_double, u, v, w)
  call fcp_correct_ghostcell_velocity(P, 0.01_c_double, u, v, w)
  call fcp_face_vel(u, v, w, bcxu, bcyv, bczw, fu, fv, fw)
  call fcp_update_pressure(P, pn)

  ! one line-solve call of solve_ad on v with the new face velocities
  call fcp_set_solve_ad(1.0_c_double / 200.0_c_double, 0.002_c_double, 1, 1.0_c_double)
  allocate(var(0:n+1,0:n+1,0:n+1))
  var = v
  call fcp_itsolv_ad(var, nlv, fu, fv, fw)
  call fcp_calc_residual_ad(var, nlv, fu, fv, fw, resm, loc)
  ! the explicit right-hand side of the next momentum step
  call field(n1, 8)
  call field(n2, 9)
  call field(n3, 10)
  call fcp_rhs_advec_diff(u, v, w, fu, fv, fw, 0.5_c_double, 0.002_c_double, n1, n2, n3)
  call fcp_finalize()

  open(newunit=iu, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(iu) u, v, w, fu, fv, fw, P, pn, var, resm, loc
  write(iu) n1, n2, n3
  close(iu)
  print '(a, es12.4, 3i5)', 'fc_step_driver: calc_residual_ad', resm, loc

contains

  subroutine field(a, s)
    real(c_double), allocatable, intent(out) :: a(:,:,:)
    integer, intent(in) :: s
    integer :: ii, jj, kk
    allocate(a(0:n+1,0:n+1,0:n+1))
    do kk = 0, n+1
      do jj = 0, n+1
        do ii = 0, n+1
          a(ii,jj,kk) = real(mod(7*ii + 13*jj + 29*kk + 31*s, 17), c_double) / 17.0_c_double - 0.5_c_double
        end do
      end do
    end do
  end subroutine field

end program fc_step_driver
