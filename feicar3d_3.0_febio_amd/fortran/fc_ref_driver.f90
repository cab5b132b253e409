!------------------------------------------------------------------------------
! fc_ref_driver -- a Fortran host program that calls the hot path under the
! reference's own names (MODULE fc_poisson_ref): solve_poisson(), then
! calc_residual(pPrime, div, resm, loc), calc_residual2(pPrime, div, resm2,
! loc2) against p_temp = p, and itsolv(var, div) on a copy of the input
! pPrime.
!
!   fc_ref_driver case.bin out.bin
!
! case.bin: as fc_driver.  out.bin (stream): pPrime, fc_stats_t, resm, loc(3)
! (int32), resm2, loc2(3), var after itsolv.
!------------------------------------------------------------------------------
program fc_ref_driver
  use iso_c_binding
  use fc_poisson
  use fc_poisson_ref
  implicit none
  character(len=512) :: fin, fout
  integer(c_int) :: nx, ny, nz, nG, nD
  real(c_double), allocatable, target :: x(:), y(:), z(:)
  integer(c_int8_t), allocatable, target :: iblank(:,:,:), ghost(:,:,:), dead(:,:,:)
  integer(c_int), allocatable, target :: iG(:), jG(:), kG(:), iC(:), jC(:), kC(:)
  integer(c_int), allocatable, target :: iD(:), jD(:), kD(:), iDC(:), jDC(:), kDC(:)
  real(c_double), allocatable, target :: cG(:,:), alpha(:), dpG(:), cD(:,:), plD(:), dpD(:), ipD(:,:)
  real(c_double), allocatable, target :: div(:,:,:), p(:,:,:), pPrime(:,:,:), var(:,:,:)
  type(fc_params_t) :: prm
  type(fc_stats_t) :: stbuf
  character(len=:), allocatable :: raw
  real(c_double) :: resm, resm2
  integer :: loc(3), loc2(3)
  integer :: u

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) nx, ny, nz
  allocate(x(0:nx+1), y(0:ny+1), z(0:nz+1))
  read(u) x, y, z
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
  allocate(var(0:nx+1,0:ny+1,0:nz+1))
  var = pPrime

  ! MG_initial, mg_prepare and the GCM lists, then the module arrays
  call fcp_init(nx, ny, nz, x, y, z, prm)
  call fcp_set_masks(iblank, ghost, dead)
  call fcp_set_gcm(nG, iG, jG, kG, iC, jC, kC, cG, alpha, dpG, nD, iD, jD, kD, iDC, jDC, kDC, cD, plD, dpD, ipD)
  call fcp_bind_flow(pPrime, div, p, p_temp=p)
  write(*,'(a,3i6)') 'fc_ref_driver: grid', nx, ny, nz
  flush(6)

  call solve_poisson()
  call calc_residual(pPrime, div, resm, loc)
  call calc_residual2(pPrime, div, resm2, loc2)
  call itsolv(var, div)
  write(*,'(a,i6,a,es12.5,a,3i5)') 'fc_ref_driver: outer ', fcp_last_stats%outer_iters, '  calc_residual ', &
       resm, '  at', loc
  flush(6)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) pPrime
  stbuf = fcp_last_stats
  write(u) transfer(stbuf, repeat(' ', int(c_sizeof(stbuf))))
  write(u) resm, int(loc, c_int), resm2, int(loc2, c_int)
  write(u) var
  close(u)
  call fcp_finalize()
  write(*,'(a)') 'fc_ref_driver: done'
  flush(6)
end program fc_ref_driver
