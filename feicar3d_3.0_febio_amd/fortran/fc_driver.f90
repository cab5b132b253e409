!------------------------------------------------------------------------------
! fc_driver -- a Fortran host program that drives the pressure-Poisson path
! through MODULE fc_poisson exactly the way the flow solver's time step does
! (rhs_poisson output -> solve_poisson, CORE_TIME_STEP.F90:171-178).
!
!   fc_driver case.bin out.bin
!
! case.bin (stream, little-endian): nx,ny,nz; x,y,z; fc_params_t; iblank,
! ghostCellMark, dead_cell; nGhost and the ghost lists; nDead and the dead
! lists; div, p, pPrime.  out.bin: pPrime after the solve, then fc_stats_t.
!------------------------------------------------------------------------------
program fc_driver
  use iso_c_binding
  use fc_poisson
  implicit none
  character(len=512) :: fin, fout
  integer(c_int) :: nx, ny, nz, nG, nD
  real(c_double), allocatable, target :: x(:), y(:), z(:)
  integer(c_int8_t), allocatable, target :: iblank(:,:,:), ghost(:,:,:), dead(:,:,:)
  integer(c_int), allocatable, target :: iG(:), jG(:), kG(:), iC(:), jC(:), kC(:)
  integer(c_int), allocatable, target :: iD(:), jD(:), kD(:), iDC(:), jDC(:), kDC(:)
  real(c_double), allocatable, target :: cG(:,:), alpha(:), dpG(:), cD(:,:), plD(:), dpD(:), ipD(:,:)
  real(c_double), allocatable, target :: div(:,:,:), p(:,:,:), pPrime(:,:,:)
  type(fc_params_t) :: prm
  type(fc_stats_t) :: stbuf
  character(len=:), allocatable :: raw
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) nx, ny, nz
  allocate(x(0:nx+1), y(0:ny+1), z(0:nz+1))
  read(u) x, y, z
  ! the C struct is read as raw bytes (unformatted derived-type transfer
  ! would skip the alignment padding between components)
  allocate(character(len=c_sizeof(prm)) :: raw)
  read(u) raw
  prm = transfer(raw, prm)
  deallocate(raw)
  allocate(iblank(0:nx+1,0:ny+1,0:nz+1), ghost(0:nx+1,0:ny+1,0:nz+1), dead(0:nx+1,0:ny+1,0:nz+1))
  read(u) iblank, ghost, dead
  read(u) nG
  allocate(iG(max(nG,1)), jG(max(nG,1)), kG(max(nG,1)), iC(max(nG,1)), jC(max(nG,1)), kC(max(nG,1)))
  allocate(cG(8,max(nG,1)), alpha(max(nG,1)), dpG(max(nG,1)))
  if (nG > 0) read(u) iG(1:nG), jG(1:nG), kG(1:nG), iC(1:nG), jC(1:nG), kC(1:nG), cG(:,1:nG), alpha(1:nG), dpG(1:nG)
  read(u) nD
  allocate(iD(max(nD,1)), jD(max(nD,1)), kD(max(nD,1)), iDC(max(nD,1)), jDC(max(nD,1)), kDC(max(nD,1)))
  allocate(cD(8,max(nD,1)), plD(max(nD,1)), dpD(max(nD,1)), ipD(3,max(nD,1)))
  if (nD > 0) read(u) iD(1:nD), jD(1:nD), kD(1:nD), iDC(1:nD), jDC(1:nD), kDC(1:nD), cD(:,1:nD), plD(1:nD), &
       dpD(1:nD), ipD(:,1:nD)
  allocate(div(0:nx+1,0:ny+1,0:nz+1), p(0:nx+1,0:ny+1,0:nz+1), pPrime(0:nx+1,0:ny+1,0:nz+1))
  read(u) div, p, pPrime
  close(u)

  write(*,'(a,3i6,a,i6,a,es10.3,a,i6)') 'fc_driver: grid', nx, ny, nz, '  itermax_pson', prm%itermax_pson, &
       '  restol_pson', prm%restol_pson, '  nGhost', nG
  flush(6)
  call fcp_init(nx, ny, nz, x, y, z, prm)          ! MG_initial
  call fcp_set_masks(iblank, ghost, dead)          ! mg_prepare
  call fcp_set_gcm(nG, iG, jG, kG, iC, jC, kC, cG, alpha, dpG, nD, iD, jD, kD, iDC, jDC, kDC, cD, plD, dpD, ipD)
  write(*,'(a)') 'fc_driver: hierarchy, masks and GCM lists on the device'
  flush(6)
  call fcp_solve_poisson(pPrime, div, p)           ! solve_poisson()
  write(*,'(a,i6,a,i6,a,es12.5,a,es12.5)') 'fc_driver: outer ', fcp_last_stats%outer_iters, '  V-cycles ', &
       fcp_last_stats%vcycles_fp64 + fcp_last_stats%vcycles_fp32, '  maxres ', fcp_last_stats%maxres, &
       '  rel ', fcp_last_stats%rel_resid
  flush(6)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) pPrime
  stbuf = fcp_last_stats
  write(u) transfer(stbuf, repeat(' ', int(c_sizeof(stbuf))))
  close(u)
  call fcp_finalize()
  write(*,'(a)') 'fc_driver: done'
  flush(6)
end program fc_driver
