! TEST INFRASTRUCTURE ONLY.  A driver program around the reference's own
! multigrid pressure solver, compiled unmodified from the reference tree by
! oracle/ref.mk into oracle/_ref/.  The tests run it beside the C oracle on
! the same case file and compare the two (the oracle restates the same
! subroutines).  Nothing in the product path loads it.
!
! What it calls, all from the reference sources:
!   slice_assignment   CORE_PARALLEL.F90:91-211   (one rank, uniform slabs)
!   allocate_memory    CORE_MEMORY_ALLOCATE.F90:6-203
!   make_grid          CORE_GRID.F90:29-172       (nonuniform branch, reads xgrid.dat ...)
!   MG_initial         CORE_MG.F90:22-599
!   mg_prepare         CORE_MG.F90:1365-1555
!   mg_solve           CORE_MG.F90:1572-1848      (mode 0)
!   itsolv             CORE_POISSON.F90:205-456   (mode 1)
!   calc_residual      CORE_POISSON.F90:460-550
!   metrics            CORE_GRID.F90:175-207          (mode 2)
!   correct_vel        CORE_FRAC_STEP_ROUTINES.F90:134-210   (mode 2)
!   face_vel           CORE_FRAC_STEP_ROUTINES.F90:6-130     (mode 2)
!   update_pressure    CORE_FRAC_STEP_ROUTINES.F90:212-299   (mode 2)
!   set_solve_ad       CORE_SET_SOLVERS.F90:591-676          (mode 3)
!   rhs_advec_diff     CORE_SET_SOLVERS.F90:10-61            (mode 4)
!   BOUNDARY_allocate_memory     GCM_MEMORY_ALLOCATE.F90:2-112     (mode 5)
!   calculate_arclength_norm_ds  BOUNDARY_SET_ARCLENGTH_NORM.F90:2-189 (mode 5)
!   set_iblank_body_fast         BOUNDARY_IBLANK_FAST.F90:2-854     (mode 5)
!   GCM_set_internal_boundary    GCM_SET_BOUNDARY.F90:2-604         (mode 5;
!     smooth_iblank, GCM_define_ghostcells, calc_bodyIntercept_Unstruc,
!     GCM_Calc_VanMatrixDN with LAPACK's DGETRF/DGETRI, GCM_set_deadcells)
!   GCM_SetBodyInterceptValues   GCM_BODYINTERCEPT_VALUES.F90:22-178 (mode 5,
!                                hdr(23) = 1)
!   compute_marker_stress        UTIL_DRAG_LIFT.F90:29-342   (mode 5,
!                                hdr(23) = 1)
! What it sets itself: the MPI state MPI_Initialize (CORE_PARALLEL.F90:2-88)
! would set for one flow rank (that routine also reads mpi_setup.dat and the
! body input, which the harness has no use for), the solver parameters
! (normally read from input.dat) and the masks iblank / ghostCellMark /
! dead_cell, which GCM_MEMORY_ALLOCATE.F90:16-20 allocates with these bounds.
!
! Usage: ref_harness <case file> <result file>, run in a scratch directory
! (make_grid and MG_initial write fort.* listings into the current one).
!
! Case file (stream, little endian):
!   int32 hdr(32): nx ny nz mode iterInter is_PBC_homogeneous mgrids_x mgrids_y
!                  mgrids_z pbc(6) bc_vel(6) nsweep  (the rest zero)
!   float64 rhdr(8): omega_pson restol_pson ppp(6)
!   float64 x(1:nx) y(1:ny) z(1:nz)
!   int8 iblank, ghostCellMark, dead_cell   each (0:nx+1,0:ny+1,0:nz+1)
!   float64 pgradx1 pgradx2 (0:ny+1,0:nz+1), pgrady1 pgrady2 (0:nx+1,0:nz+1),
!           pgradz1 pgradz2 (0:nx+1,0:ny+1)
!   float64 var, rr   each (0:nx+1,0:ny+1,0:nz+1)
!   mode 2 only, after rr: float64 dt; u, v, w, face_u, face_v, face_w (each
!   (0:nx+1,0:ny+1,0:nz+1)); bcxu (0:1,0:ny+1,0:nz+1), bcyv (0:nx+1,0:1,
!   0:nz+1), bczw (0:nx+1,0:ny+1,0:1).  var is pPrime.
! Result file, mode 2: float64 u, v, w after correct_vel; face_u, face_v,
!   face_w after face_vel(u, v, w); pPrime, p after update_pressure.
!   mode 3 only, after rr: float64 reinv, dt.
! Result file, mode 3: float64 amx_ad, apx_ad (0:nx+1), amy_ad, apy_ad
!   (0:ny+1), amz_ad, apz_ad (0:nz+1) after set_solve_ad.
!   mode 4 only, after rr: float64 dt, alfa_upwind; uOld, vOld, wOld, face_u,
!   face_v, face_w, viscTot, nlu, nlv, nlw (each (0:nx+1,0:ny+1,0:nz+1)).
! Result file, mode 4: float64 nlu, nlv, nlw after rhs_advec_diff; int32
!   nGhost (the cells with |ghostCellMark| = 1, numbered in array order);
!   float64 ghostNScoeff(8:10, 1:nGhost).
!   mode 5 only, after rr: int32 nBody; float64 membrane_tkns,
!   probeLengthNormalized, probeLengthNormalizedD; per body int32 nPts, nElem,
!   unstruc_surface_type, body_dim; per body float64 xBodyMarker(nPts),
!   yBodyMarker, zBodyMarker, then int32 triElemNeig(3,nElem).
!   With hdr(23) = 1 there follow float64 dt, reinv; per body
!   u/v/wBodyMarker and u/v/wBodyMarkerOld (nPts each); p, u, v, w (each
!   (0:nx+1,0:ny+1,0:nz+1)).
! Result file, mode 5: int8 iblank, ghostCellMark, dead_cell, bodyNum (each
!   (0:nx+1,0:ny+1,0:nz+1)); per body float64 triElemNormX/Y/Z, triElemArea,
!   triElemCentX/Y/Z (nElem each), xNormBodyMarker/y/z (nPts each); int32
!   nGhost; int32 iGhost jGhost kGhost iCellIndex jCellIndex kCellIndex
!   closestElementGC; float64 xBodyIntercept y z, xBodyInterceptNorm y z,
!   probeLength, xImagePoint y z, coeffGCMD(8,:), coeffGCMN(8,:), alphaGhost;
!   int32 nDead; int32 iDead jDead kDead iDeadCellIndex jDeadCellIndex
!   kDeadCellIndex closestElementDead; float64 xBodyInterceptDead y z,
!   xNormInterceptDead y z, probeLengthDead, coeffGCMDeadD(8,:),
!   coeffGCMDeadN(8,:); float64 xyzDCIP(0:nx+1,0:ny+1,0:nz+1,0:2).
!   With hdr(23) = 1 then: float64 u/v/wBodyIntercept, dpdnBodyIntercept
!   (nGhost each), u/v/wBodyInterceptDead, dpdnInterceptDead (nDead each);
!   per body cxt cyt czt cmxt cmyt cmzt cxs cys czs, then
!   x/y/zMarkerForce, x/y/zMarkerShear, MarkerArea, x/y/zMarkerStress
!   (nPts each) after compute_marker_stress (fsi_filter_f = 1: no blending
!   with the previous step's forces).
! Result file, modes 0 and 1:
!   int32 mLevel; float64 resm (calc_residual after the solve)
!   float64 var (0:nx+1,0:ny+1,0:nz+1)
!   per level n = 1..mLevel: int32 I J K; int8 igk ibk; float64 rhs phi res,
!   each (0:I+1,0:J+1,0:K+1)
program ref_harness
  use global_parameters
  use flow_parameters
  use flow_arrays
  use grid_arrays
  use boundary_arrays
  use mg_module
  use solver_arrays
  use MPI_module
  use multiuse_arrays, only: nlu, nlv, nlw
  use GCM_arrays, only: ghostNScoeff
  implicit none
  character(len=1024) :: fin, fout
  integer :: hdr(32), iu, i, n, mode, nsweep, loc(3)
  real(kind=CGREAL) :: rhdr(8), resm
  real(kind=CGREAL), allocatable :: xs(:), ys(:), zs(:), var(:,:,:), rr(:,:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=iu, file=trim(fin), access='stream', form='unformatted', status='old')
  read(iu) hdr, rhdr
  nx = hdr(1); ny = hdr(2); nz = hdr(3)
  mode = hdr(4); nsweep = hdr(22)

  ! one flow rank: the state MPI_Initialize leaves for nProcY = nProcZ = 1
  call MPI_INIT(ierr)
  call MPI_COMM_SIZE(MPI_COMM_WORLD, nProcs_g, ierr)
  call MPI_COMM_RANK(MPI_COMM_WORLD, lProc_g, ierr)
  if (nProcs_g /= 1) stop 'ref_harness: run it as one process'
  flow_comm = MPI_COMM_WORLD
  nProcs = 1; lProc = 0; nElastic = 0; PROC_M = 0
  nProcY = 1; nProcZ = 1; jProc = 0; kProc = 0
  lProc_leftmost = 0; lProc_rightmost = 0; lProc_btommost = 0; lProc_topmost = 0
  lProc_left = -1; lProc_right = 1; lProc_top = 1; lProc_bottom = -1
  y_slab_unif = UNIFORM_SLAB; z_slab_unif = UNIFORM_SLAB
  call slice_assignment()
  call allocate_memory()

  ! grid: the nonuniform branch of make_grid reads the node files
  allocate(xs(nx), ys(ny), zs(nz))
  read(iu) xs, ys, zs
  call write_axis('xgrid.dat', xs)
  call write_axis('ygrid.dat', ys)
  call write_axis('zgrid.dat', zs)
  xgrid_unif = NONUNIFORM_GRID; ygrid_unif = NONUNIFORM_GRID; zgrid_unif = NONUNIFORM_GRID
  call make_grid()
  do i = 1, nx
    if (x(i) /= xs(i)) stop 'ref_harness: x node did not round-trip'
  end do
  do i = 1, ny
    if (y(i) /= ys(i)) stop 'ref_harness: y node did not round-trip'
  end do
  do i = 1, nz
    if (z(i) /= zs(i)) stop 'ref_harness: z node did not round-trip'
  end do

  ! solver parameters (input.dat in the reference)
  it_solver_type = IT_SOLVER_TYPE_MG
  iterInter = hdr(5)
  is_PBC_homogeneous = hdr(6)
  mgrids_x = hdr(7); mgrids_y = hdr(8); mgrids_z = hdr(9)
  pbcx1 = hdr(10); pbcx2 = hdr(11); pbcy1 = hdr(12); pbcy2 = hdr(13); pbcz1 = hdr(14); pbcz2 = hdr(15)
  bcx1 = hdr(16); bcx2 = hdr(17); bcy1 = hdr(18); bcy2 = hdr(19); bcz1 = hdr(20); bcz2 = hdr(21)
  omega_pson = rhdr(1); restol_pson = rhdr(2)
  pppx1 = rhdr(3); pppx2 = rhdr(4); pppy1 = rhdr(5); pppy2 = rhdr(6); pppz1 = rhdr(7); pppz2 = rhdr(8)
  infoconv = 0; ntime = 1; nmonitor = huge(1); iRedBlack = 0

  allocate(iblank(0:nx+1,yb1:yb2,zb1:zb2), dead_cell(0:nx+1,yb1:yb2,zb1:zb2))
  allocate(ghostCellMark(0:nx+1,yb1:yb2,zb1:zb2), igmark(0:nx+1,yb1:yb2,zb1:zb2))
  read(iu) iblank, ghostCellMark, dead_cell
  read(iu) pgradx1, pgradx2, pgrady1, pgrady2, pgradz1, pgradz2
  allocate(var(0:nx+1,yb1:yb2,zb1:zb2), rr(0:nx+1,yb1:yb2,zb1:zb2))
  read(iu) var, rr
  close(iu)

  if (mode == 2) then
    call projection(trim(fin), trim(fout))
    call MPI_FINALIZE(ierr)
    stop
  end if
  if (mode == 3) then
    call ad_coefficients(trim(fin), trim(fout))
    call MPI_FINALIZE(ierr)
    stop
  end if
  if (mode == 4) then
    call advec_diff_rhs(trim(fin), trim(fout))
    call MPI_FINALIZE(ierr)
    stop
  end if
  if (mode == 5) then
    call geometry_refresh(trim(fin), trim(fout))
    call MPI_FINALIZE(ierr)
    stop
  end if

  call MG_initial()
  call mg_prepare()

  select case (mode)
  case (0)
    call mg_solve(var, rr)
  case (1)
    do n = 1, nsweep
      call itsolv(var, rr)
    end do
  case default
    stop 'ref_harness: unknown mode'
  end select
  call calc_residual(var, rr, resm, loc)

  open(newunit=iu, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(iu) mLevel, resm
  write(iu) var
  do n = 1, mLevel
    write(iu) mgrid_I(n), mgrid_J(n), mgrid_K(n)
    write(iu) mg_array(n)%igk, mg_array(n)%ibk
    write(iu) mg_array(n)%rhs, mg_array(n)%phi, mg_array(n)%res
  end do
  close(iu)
  call MPI_FINALIZE(ierr)

contains

  ! the velocity projection after a solve (CORE_TIME_STEP.F90:211-222; the
  ! ghost-cell correction between correct_vel and face_vel lives in
  ! GCM_GHOSTCELL_UPDATE.F90, which flang does not compile)
  subroutine projection(fin, fout)
    character(len=*), intent(in) :: fin, fout
    integer :: v1
    integer(kind=8) :: ex, ey, ez, pos
    call metrics()
    ! the mode-2 data follow the common part of the case file
    open(newunit=v1, file=fin, access='stream', form='unformatted', status='old')
    ex = nx + 2; ey = ny + 2; ez = nz + 2
    pos = 32 * 4 + 8 * 8 + 8 * (ex + ey + ez - 6) + 3 * ex * ey * ez &
        + 8 * 2 * (ey * ez + ex * ez + ex * ey) + 8 * 2 * ex * ey * ez + 1
    read(v1, pos=pos) dt
    read(v1) u, v, w, face_u, face_v, face_w
    read(v1) bcxu, bcyv, bczw
    close(v1)
    pPrime = var
    open(newunit=v1, file=fout, access='stream', form='unformatted', status='replace')
    call correct_vel()
    write(v1) u, v, w
    call face_vel(u, v, w)
    write(v1) face_u, face_v, face_w
    ! At one rank update_pressure's kProc == nProcZ-1 branch (CORE_FRAC_STEP_
    ! ROUTINES.F90:278-284) also assigns the plane k = zc_start-2 = -1, which
    ! lies before pPrime(0:nx+1, 0:ny+1, 0:nz+1) in memory; the rest of that
    ! branch, and the two corner blocks that test kProc == nProcZ-1, repeat
    ! assignments the kProc == 0 branches make.  Declaring a second z rank
    ! for this call skips those branches and nothing else.
    nProcZ = 2
    call update_pressure()
    nProcZ = 1
    write(v1) pPrime, p
    close(v1)
  end subroutine projection

  ! the line coefficients of the advection-diffusion solve
  subroutine ad_coefficients(fin, fout)
    character(len=*), intent(in) :: fin, fout
    integer :: v1
    integer(kind=8) :: ex, ey, ez, pos
    open(newunit=v1, file=fin, access='stream', form='unformatted', status='old')
    ex = nx + 2; ey = ny + 2; ez = nz + 2
    pos = 32 * 4 + 8 * 8 + 8 * (ex + ey + ez - 6) + 3 * ex * ey * ez &
        + 8 * 2 * (ey * ez + ex * ez + ex * ey) + 8 * 2 * ex * ey * ez + 1
    read(v1, pos=pos) reinv, dt
    close(v1)
    call set_solve_ad()
    open(newunit=v1, file=fout, access='stream', form='unformatted', status='replace')
    write(v1) amx_ad, apx_ad, amy_ad, apy_ad, amz_ad, apz_ad
    close(v1)
  end subroutine ad_coefficients

  ! the explicit right-hand side of the advection-diffusion step
  subroutine advec_diff_rhs(fin, fout)
    character(len=*), intent(in) :: fin, fout
    integer :: v1, ng, ii, jj, kk
    integer(kind=8) :: ex, ey, ez, pos
    call metrics()
    open(newunit=v1, file=fin, access='stream', form='unformatted', status='old')
    ex = nx + 2; ey = ny + 2; ez = nz + 2
    pos = 32 * 4 + 8 * 8 + 8 * (ex + ey + ez - 6) + 3 * ex * ey * ez &
        + 8 * 2 * (ey * ez + ex * ez + ex * ey) + 8 * 2 * ex * ey * ez + 1
    read(v1, pos=pos) dt, alfa_upwind
    read(v1) uOld, vOld, wOld, face_u, face_v, face_w, viscTot, nlu, nlv, nlw
    close(v1)
    ! GCM_MEMORY_ALLOCATE.F90:21, 172 allocate these; the ghost cells are
    ! numbered in array order
    allocate(ghostCellIndex(0:nx+1,yb1:yb2,zb1:zb2))
    ghostCellIndex = 0
    ng = 0
    do kk = zb1, zb2
      do jj = yb1, yb2
        do ii = 0, nx + 1
          if (abs(ghostCellMark(ii,jj,kk)) == 1) then
            ng = ng + 1
            ghostCellIndex(ii,jj,kk) = ng
          end if
        end do
      end do
    end do
    allocate(ghostNScoeff(10, max(ng, 1)))
    ghostNScoeff = 0.0_CGREAL
    call rhs_advec_diff()
    open(newunit=v1, file=fout, access='stream', form='unformatted', status='replace')
    write(v1) nlu, nlv, nlw
    write(v1) ng
    if (ng > 0) write(v1) ghostNScoeff(8:10, 1:ng)
    close(v1)
  end subroutine advec_diff_rhs

  ! the immersed-boundary geometry of one time step (CORE_TIME_STEP.F90:
  ! 100-118 for moving bodies): surface normals, iblank, ghost and dead cells
  ! with their relations.  The surfaces are given directly (the reference
  ! reads them from its body input, BOUNDARY_MARKER_INIT.F90).
  subroutine geometry_refresh(fin, fout)
    use GCM_arrays
    use fsi_module, only: fsi_filter_f
    character(len=*), intent(in) :: fin, fout
    integer :: v1, v2, ib, nb, m
    integer(kind=8) :: ex, ey, ez, pos, posd
    integer, allocatable :: npt(:), nel(:), stype(:), bdim(:), tri(:,:)
    real(kind=CGREAL) :: rp(3)
    real(kind=CGREAL), allocatable :: a(:)
    call metrics()
    open(newunit=v1, file=fin, access='stream', form='unformatted', status='old')
    ex = nx + 2; ey = ny + 2; ez = nz + 2
    pos = 32 * 4 + 8 * 8 + 8 * (ex + ey + ez - 6) + 3 * ex * ey * ez &
        + 8 * 2 * (ey * ez + ex * ez + ex * ey) + 8 * 2 * ex * ey * ez + 1
    read(v1, pos=pos) nb
    read(v1) rp
    allocate(npt(nb), nel(nb), stype(nb), bdim(nb))
    do ib = 1, nb
      read(v1) npt(ib), nel(ib), stype(ib), bdim(ib)
    end do
    ! the body parameters MAIN.F90:440-447 allocates and reads
    nBody = nb
    allocate(canonical_body_type(nb), body_dim(nb), boundary_motion_type(nb), unstruc_surface_type(nb))
    allocate(wall_type(nb), nPtsBodyMarkerOrig(nb), nPtsBodyMarker(nb), totNumTriElem(nb))
    canonical_body_type = UNSTRUCTURED_SURFACE
    boundary_motion_type = 0; wall_type = 0
    body_dim = bdim; unstruc_surface_type = stype
    nPtsBodyMarker = npt; totNumTriElem = nel
    membrane_tkns = rp(1); probeLengthNormalized = rp(2); probeLengthNormalizedD = rp(3)
    ! the masks the main program allocated are allocated again, with the
    ! rest of the boundary arrays, by BOUNDARY_allocate_memory
    deallocate(iblank, dead_cell, ghostCellMark, igmark)
    call BOUNDARY_allocate_memory()
    do ib = 1, nb
      allocate(a(npt(ib)))
      read(v1) a
      xBodyMarker(1:npt(ib), ib) = a
      read(v1) a
      yBodyMarker(1:npt(ib), ib) = a
      read(v1) a
      zBodyMarker(1:npt(ib), ib) = a
      deallocate(a)
      allocate(tri(3, nel(ib)))
      read(v1) tri
      triElemNeig(1:3, 1:nel(ib), ib) = tri
      deallocate(tri)
    end do
    inquire(unit=v1, pos=posd)
    close(v1)
    ! a three-dimensional run, a time step after the first: no orientation
    ! pass, no array release
    nDim = DIM_3D
    ntime = 1; ntime_start = 1
    call calculate_arclength_norm_ds()
    call set_iblank_body_fast()
    call GCM_set_internal_boundary()
    open(newunit=v1, file=fout, access='stream', form='unformatted', status='replace')
    write(v1) iblank, ghostCellMark, dead_cell, bodyNum
    do ib = 1, nb
      m = nel(ib)
      write(v1) triElemNormX(1:m, ib), triElemNormY(1:m, ib), triElemNormZ(1:m, ib), triElemArea(1:m, ib)
      write(v1) triElemCentX(1:m, ib), triElemCentY(1:m, ib), triElemCentZ(1:m, ib)
      m = npt(ib)
      write(v1) xNormBodyMarker(1:m, ib), yNormBodyMarker(1:m, ib), zNormBodyMarker(1:m, ib)
    end do
    write(v1) nGhost
    if (nGhost > 0) then
      write(v1) iGhost(1:nGhost), jGhost(1:nGhost), kGhost(1:nGhost)
      write(v1) iCellIndex(1:nGhost), jCellIndex(1:nGhost), kCellIndex(1:nGhost), closestElementGC(1:nGhost)
      write(v1) xBodyIntercept(1:nGhost), yBodyIntercept(1:nGhost), zBodyIntercept(1:nGhost)
      write(v1) xBodyInterceptNorm(1:nGhost), yBodyInterceptNorm(1:nGhost), zBodyInterceptNorm(1:nGhost)
      write(v1) probeLength(1:nGhost), xImagePoint(1:nGhost), yImagePoint(1:nGhost), zImagePoint(1:nGhost)
      write(v1) coeffGCMD(1:8, 1:nGhost), coeffGCMN(1:8, 1:nGhost), alphaGhost(1:nGhost)
    end if
    write(v1) nDead
    if (nDead > 0) then
      write(v1) iDead(1:nDead), jDead(1:nDead), kDead(1:nDead)
      write(v1) iDeadCellIndex(1:nDead), jDeadCellIndex(1:nDead), kDeadCellIndex(1:nDead), closestElementDead(1:nDead)
      write(v1) xBodyInterceptDead(1:nDead), yBodyInterceptDead(1:nDead), zBodyInterceptDead(1:nDead)
      write(v1) xNormInterceptDead(1:nDead), yNormInterceptDead(1:nDead), zNormInterceptDead(1:nDead)
      write(v1) probeLengthDead(1:nDead), coeffGCMDeadD(1:8, 1:nDead), coeffGCMDeadN(1:8, 1:nDead)
    end if
    write(v1) xyzDCIP
    if (hdr(23) == 1) then
      ! the body-intercept values and the marker stresses of the step
      open(newunit=v2, file=fin, access='stream', form='unformatted', status='old')
      read(v2, pos=posd) dt, reinv
      do ib = 1, nb
        m = npt(ib)
        read(v2) uBodyMarker(1:m, ib), vBodyMarker(1:m, ib), wBodyMarker(1:m, ib)
        read(v2) uBodyMarkerOld(1:m, ib), vBodyMarkerOld(1:m, ib), wBodyMarkerOld(1:m, ib)
      end do
      read(v2) p, u, v, w
      close(v2)
      ntime = 2
      call GCM_SetBodyInterceptValues()
      fsi_filter_f = 1.0_CGREAL
      xMarkerForceOld = 0.0_CGREAL; yMarkerForceOld = 0.0_CGREAL; zMarkerForceOld = 0.0_CGREAL
      call compute_marker_stress()
      if (nGhost > 0) write(v1) uBodyIntercept(1:nGhost), vBodyIntercept(1:nGhost), &
          wBodyIntercept(1:nGhost), dpdnBodyIntercept(1:nGhost)
      if (nDead > 0) write(v1) uBodyInterceptDead(1:nDead), vBodyInterceptDead(1:nDead), &
          wBodyInterceptDead(1:nDead), dpdnInterceptDead(1:nDead)
      do ib = 1, nb
        m = npt(ib)
        write(v1) cxt(ib), cyt(ib), czt(ib), cmxt(ib), cmyt(ib), cmzt(ib), cxs(ib), cys(ib), czs(ib)
        write(v1) xMarkerForce(1:m, ib), yMarkerForce(1:m, ib), zMarkerForce(1:m, ib)
        write(v1) xMarkerShear(1:m, ib), yMarkerShear(1:m, ib), zMarkerShear(1:m, ib), MarkerArea(1:m, ib)
        write(v1) xMarkerStress(1:m, ib), yMarkerStress(1:m, ib), zMarkerStress(1:m, ib)
      end do
    end if
    close(v1)
  end subroutine geometry_refresh

  ! one node per line as make_grid reads it (index, coordinate); 17
  ! significant digits make the decimal text round-trip to the same double
  subroutine write_axis(name, a)
    character(len=*), intent(in) :: name
    real(kind=CGREAL), intent(in) :: a(:)
    integer :: v, m
    open(newunit=v, file=name, status='replace', form='formatted')
    do m = 1, size(a)
      write(v, '(I8,1X,ES26.17E3)') m, a(m)
    end do
    close(v)
  end subroutine write_axis

end program ref_harness
