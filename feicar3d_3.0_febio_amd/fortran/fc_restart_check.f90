! fc_restart_check -- writes and reads a restart file with Fortran's own
! unformatted sequential I/O, in the record layout of the reference's
! write_restart / read_restart (MAIN.F90:983-1122): a check that the C-ABI's
! fc_restart_write produces the bytes Fortran writes and that
! fc_restart_read reads what Fortran writes.
!
!   fc_restart_check write <path> nx ny nz   (arrays filled with the pattern)
!   fc_restart_check read  <path> nx ny nz   (exit 0 when the file holds it)
!
! The pattern: the a-th real*8 array of record 2 (u = 1 ... pgradz2 = 26)
! holds 1000 a + 0.5 i at 0-based position i (Fortran order); iblank holds
! mod(i,3) - 1, ghostCellMark mod(i,2); time = 1.25, ntime = 77.  Two bodies
! (7 markers and 4 elements, 5 and 3): the q-th marker array (x = 1 ...
! zMarkerForce = 9) of body b holds 100 b + 10 q + 0.25 m at 0-based m,
! triElemNeig(r, e) = 3 (e-1) + r + b, and the twelve scalars 0.5 (b + s)
! for s = 1..12.
program fc_restart_check
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  character(len=256) :: mode, path, arg
  integer :: nx, ny, nz, X, Y, Z, ios
  real(dp), allocatable :: f3(:,:,:,:), bcx(:,:,:,:), bcy(:,:,:,:), bcz(:,:,:,:)
  real(dp), allocatable :: pgx(:,:,:), pgy(:,:,:), pgz(:,:,:)
  integer(1), allocatable :: ibl(:,:,:), gcm(:,:,:)
  real(dp), allocatable :: mk(:,:), sc(:)
  integer, allocatable :: tri(:,:)
  real(dp) :: time
  integer :: ntime, b, bad

  call get_command_argument(1, mode)
  call get_command_argument(2, path)
  call get_command_argument(3, arg); read(arg, *) nx
  call get_command_argument(4, arg); read(arg, *) ny
  call get_command_argument(5, arg); read(arg, *) nz
  X = nx + 2; Y = ny + 2; Z = nz + 2
  allocate(f3(0:nx+1, 0:ny+1, 0:nz+1, 8), bcx(0:1, 0:ny+1, 0:nz+1, 4), bcy(0:nx+1, 0:1, 0:nz+1, 4))
  allocate(bcz(0:nx+1, 0:ny+1, 0:1, 4), pgx(0:ny+1, 0:nz+1, 2), pgy(0:nx+1, 0:nz+1, 2), pgz(0:nx+1, 0:ny+1, 2))
  allocate(ibl(0:nx+1, 0:ny+1, 0:nz+1), gcm(0:nx+1, 0:ny+1, 0:nz+1))

  if (trim(mode) == 'write') then
    call fill_flow()
    open(unit=21, file=trim(path), form='unformatted', status='replace')
    time = 1.25_dp; ntime = 77
    write(21) time, ntime
    write(21) f3(:,:,:,1), f3(:,:,:,2), f3(:,:,:,3), f3(:,:,:,4), &
              f3(:,:,:,5), f3(:,:,:,6), f3(:,:,:,7), f3(:,:,:,8), &
              bcx(:,:,:,1), bcx(:,:,:,2), bcx(:,:,:,3), bcx(:,:,:,4), &
              bcy(:,:,:,1), bcy(:,:,:,2), bcy(:,:,:,3), bcy(:,:,:,4), &
              bcz(:,:,:,1), bcz(:,:,:,2), bcz(:,:,:,3), bcz(:,:,:,4), &
              pgx(:,:,1), pgx(:,:,2), pgy(:,:,1), pgy(:,:,2), pgz(:,:,1), pgz(:,:,2), &
              ibl(:,:,:), gcm(:,:,:)
    do b = 1, 2
      call fill_body(b)
      write(21) mk(:,1), mk(:,2), mk(:,3), mk(:,4), mk(:,5), mk(:,6), mk(:,7), mk(:,8), mk(:,9), &
                tri(1:3,:), sc(1), sc(2), sc(3), sc(4), sc(5), sc(6), sc(7), sc(8), sc(9), sc(10), sc(11), sc(12)
    end do
    close(21)
  else
    open(unit=21, file=trim(path), form='unformatted', status='old', iostat=ios)
    if (ios /= 0) stop 2
    read(21) time, ntime
    read(21) f3(:,:,:,1), f3(:,:,:,2), f3(:,:,:,3), f3(:,:,:,4), &
             f3(:,:,:,5), f3(:,:,:,6), f3(:,:,:,7), f3(:,:,:,8), &
             bcx(:,:,:,1), bcx(:,:,:,2), bcx(:,:,:,3), bcx(:,:,:,4), &
             bcy(:,:,:,1), bcy(:,:,:,2), bcy(:,:,:,3), bcy(:,:,:,4), &
             bcz(:,:,:,1), bcz(:,:,:,2), bcz(:,:,:,3), bcz(:,:,:,4), &
             pgx(:,:,1), pgx(:,:,2), pgy(:,:,1), pgy(:,:,2), pgz(:,:,1), pgz(:,:,2), &
             ibl(:,:,:), gcm(:,:,:)
    bad = 0
    if (time /= 1.25_dp .or. ntime /= 77) bad = bad + 1
    bad = bad + check_flow()
    do b = 1, 2
      call fill_body(b)
      mk = -1.0_dp; tri = -1; sc = -1.0_dp
      read(21) mk(:,1), mk(:,2), mk(:,3), mk(:,4), mk(:,5), mk(:,6), mk(:,7), mk(:,8), mk(:,9), &
               tri(1:3,:), sc(1), sc(2), sc(3), sc(4), sc(5), sc(6), sc(7), sc(8), sc(9), sc(10), sc(11), sc(12)
      bad = bad + check_body(b)
    end do
    close(21)
    if (bad /= 0) then
      print *, 'mismatches:', bad
      stop 1
    end if
    print *, 'restart file holds the pattern'
  end if

contains

  real(dp) function pat(a, i)
    integer, intent(in) :: a, i
    pat = 1000.0_dp * a + 0.5_dp * i
  end function

  subroutine fill_flow()
    integer :: i, j, k, a, q
    do a = 1, 8
      q = 0
      do k = 0, nz + 1; do j = 0, ny + 1; do i = 0, nx + 1
        f3(i, j, k, a) = pat(a, q); q = q + 1
      end do; end do; end do
    end do
    do a = 1, 4
      q = 0
      do k = 0, nz + 1; do j = 0, ny + 1; do i = 0, 1
        bcx(i, j, k, a) = pat(8 + a, q); q = q + 1
      end do; end do; end do
      q = 0
      do k = 0, nz + 1; do j = 0, 1; do i = 0, nx + 1
        bcy(i, j, k, a) = pat(12 + a, q); q = q + 1
      end do; end do; end do
      q = 0
      do k = 0, 1; do j = 0, ny + 1; do i = 0, nx + 1
        bcz(i, j, k, a) = pat(16 + a, q); q = q + 1
      end do; end do; end do
    end do
    do a = 1, 2
      q = 0
      do k = 0, nz + 1; do j = 0, ny + 1
        pgx(j, k, a) = pat(20 + a, q); q = q + 1
      end do; end do
      q = 0
      do k = 0, nz + 1; do i = 0, nx + 1
        pgy(i, k, a) = pat(22 + a, q); q = q + 1
      end do; end do
      q = 0
      do j = 0, ny + 1; do i = 0, nx + 1
        pgz(i, j, a) = pat(24 + a, q); q = q + 1
      end do; end do
    end do
    q = 0
    do k = 0, nz + 1; do j = 0, ny + 1; do i = 0, nx + 1
      ibl(i, j, k) = int(mod(q, 3) - 1, 1); gcm(i, j, k) = int(mod(q, 2), 1); q = q + 1
    end do; end do; end do
  end subroutine

  integer function check_flow()
    real(dp), allocatable :: g3(:,:,:,:), gx(:,:,:,:), gy(:,:,:,:), gz(:,:,:,:), hx(:,:,:), hy(:,:,:), hz(:,:,:)
    integer(1), allocatable :: gi(:,:,:), gg(:,:,:)
    g3 = f3; gx = bcx; gy = bcy; gz = bcz; hx = pgx; hy = pgy; hz = pgz; gi = ibl; gg = gcm
    call fill_flow()
    check_flow = count(g3 /= f3) + count(gx /= bcx) + count(gy /= bcy) + count(gz /= bcz) &
                 + count(hx /= pgx) + count(hy /= pgy) + count(hz /= pgz) + count(gi /= ibl) + count(gg /= gcm)
  end function

  subroutine fill_body(b)
    integer, intent(in) :: b
    integer :: npts, nelm, q, m, e, r, s
    npts = merge(7, 5, b == 1); nelm = merge(4, 3, b == 1)
    if (allocated(mk)) deallocate(mk, tri, sc)
    allocate(mk(npts, 9), tri(3, nelm), sc(12))
    do q = 1, 9
      do m = 1, npts
        mk(m, q) = 100.0_dp * b + 10.0_dp * q + 0.25_dp * (m - 1)
      end do
    end do
    do e = 1, nelm
      do r = 1, 3
        tri(r, e) = 3 * (e - 1) + r + b
      end do
    end do
    do s = 1, 12
      sc(s) = 0.5_dp * (b + s)
    end do
  end subroutine

  integer function check_body(b)
    integer, intent(in) :: b
    real(dp), allocatable :: m2(:,:), s2(:)
    integer, allocatable :: t2(:,:)
    m2 = mk; t2 = tri; s2 = sc
    call fill_body(b)
    check_body = count(m2 /= mk) + count(t2 /= tri) + count(s2 /= sc)
  end function

end program fc_restart_check
