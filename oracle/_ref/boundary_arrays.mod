﻿!mod$ v1 sum:5bbbe129bd8f8e08
!need$ aa821a1d41b92a25 n global_parameters
module boundary_arrays
use global_parameters,only:int_k
use global_parameters,only:cgreal
use global_parameters,only:viscous_flow
use global_parameters,only:potential_flow
use global_parameters,only:uniform_grid
use global_parameters,only:nonuniform_grid
use global_parameters,only:bc_type_dirichlet
use global_parameters,only:bc_type_zero_gradient
use global_parameters,only:bc_type_pulsatile_inflow
use global_parameters,only:bc_type_symmetry
use global_parameters,only:bc_type_periodic
use global_parameters,only:bc_type_user_specified
use global_parameters,only:bc_type_shear
use global_parameters,only:it_solver_type_lsor
use global_parameters,only:it_solver_type_petsc
use global_parameters,only:it_solver_type_mg
use global_parameters,only:it_solver_type_az
use global_parameters,only:fixed_boundary
use global_parameters,only:moving_boundary
use global_parameters,only:stationary
use global_parameters,only:prescribed_rigid
use global_parameters,only:induced_rigid
use global_parameters,only:prescribed_arbitrary
use global_parameters,only:elastic_motion
use global_parameters,only:specify_velocity
use global_parameters,only:specify_position
use global_parameters,only:pbc_dirichlet
use global_parameters,only:pbc_neumann
use global_parameters,only:adams_bashforth2
use global_parameters,only:runge_kutta3
use global_parameters,only:crank_nicolson2
use global_parameters,only:nonporous_and_nonslip
use global_parameters,only:porous_or_slip
use global_parameters,only:none
use global_parameters,only:general
use global_parameters,only:canonical
use global_parameters,only:elliptic_cylinder
use global_parameters,only:general_cylinder
use global_parameters,only:ellipsoid
use global_parameters,only:unstructured_surface
use global_parameters,only:solid_body
use global_parameters,only:membrane
use global_parameters,only:body_dim2
use global_parameters,only:body_dim3
use global_parameters,only:intr_bound_none
use global_parameters,only:intr_bound_present
use global_parameters,only:no_van_kan
use global_parameters,only:van_kan
use global_parameters,only:tecplot
use global_parameters,only:fieldview
use global_parameters,only:matlab
use global_parameters,only:iblank_read
use global_parameters,only:dim_2d
use global_parameters,only:dim_3d
use global_parameters,only:iblank_used
use global_parameters,only:no_internal_boundary
use global_parameters,only:ssm_method
use global_parameters,only:gcm_method
use global_parameters,only:inviscid
use global_parameters,only:icoord
use global_parameters,only:jcoord
use global_parameters,only:kcoord
use global_parameters,only:stats_none
use global_parameters,only:stdout
use global_parameters,only:inactive
use global_parameters,only:active
use global_parameters,only:err_none
use global_parameters,only:yes
use global_parameters,only:no
integer(4),parameter::nsize=1000_4
integer(4),parameter::msize=20_4
integer(4)::nfresh
integer(4)::ndead
integer(4)::iblank_reset
integer(4),allocatable::marker2did(:,:)
integer(4)::nprobebody
integer(4),allocatable::markerprobe(:,:)
integer(1),allocatable::igmark(:,:,:)
integer(1),allocatable::iblank(:,:,:)
integer(1),allocatable::dead_cell(:,:,:)
integer(1),allocatable::bodynum(:,:,:)
integer(1),allocatable::ghostcellmark(:,:,:)
integer(1),allocatable::iblanktemp(:,:,:)
integer(1),allocatable::iblankundecided(:,:,:)
real(8),allocatable::iblank_real(:,:,:)
integer(1),allocatable::ps1(:,:,:)
integer(1),allocatable::ps2(:,:,:)
integer(1),allocatable::goa(:,:,:)
real(8),allocatable::xyzdcip(:,:,:,:)
integer(4),allocatable::ghostcellindex(:,:,:)
real(8),allocatable::xbitable(:,:,:)
real(8),allocatable::ybitable(:,:,:)
real(8),allocatable::zbitable(:,:,:)
integer(4),allocatable::celtable(:,:,:)
integer(1),allocatable::is_biset(:,:,:)
real(8),allocatable::xbodymarker(:,:)
real(8),allocatable::ybodymarker(:,:)
real(8),allocatable::zbodymarker(:,:)
real(8),allocatable::ubodymarker(:,:)
real(8),allocatable::vbodymarker(:,:)
real(8),allocatable::wbodymarker(:,:)
real(8),allocatable::axbodymarker(:,:)
real(8),allocatable::aybodymarker(:,:)
real(8),allocatable::azbodymarker(:,:)
real(8),allocatable::sbodymarker(:,:)
real(8),allocatable::dsbodymarker(:,:)
real(8),allocatable::xnormbodymarker(:,:)
real(8),allocatable::ynormbodymarker(:,:)
real(8),allocatable::znormbodymarker(:,:)
real(8),allocatable::dpdnbodymarker(:,:)
real(8),allocatable::u1bodymarker(:,:)
real(8),allocatable::v1bodymarker(:,:)
real(8),allocatable::w1bodymarker(:,:)
real(8),allocatable::pgradx1(:,:)
real(8),allocatable::pgradx2(:,:)
real(8),allocatable::pgrady1(:,:)
real(8),allocatable::pgrady2(:,:)
real(8),allocatable::pgradz1(:,:)
real(8),allocatable::pgradz2(:,:)
real(8),allocatable::xbodymarkerold(:,:)
real(8),allocatable::ybodymarkerold(:,:)
real(8),allocatable::zbodymarkerold(:,:)
real(8),allocatable::ubodymarkerold(:,:)
real(8),allocatable::vbodymarkerold(:,:)
real(8),allocatable::wbodymarkerold(:,:)
real(8),allocatable::ubodymarkertmp(:,:)
real(8),allocatable::vbodymarkertmp(:,:)
real(8),allocatable::wbodymarkertmp(:,:)
real(8),allocatable::trielemnormx(:,:)
real(8),allocatable::trielemnormy(:,:)
real(8),allocatable::trielemnormz(:,:)
real(8),allocatable::trielemarea(:,:)
real(8),allocatable::trielemcentx(:,:)
real(8),allocatable::trielemcenty(:,:)
real(8),allocatable::trielemcentz(:,:)
real(8),allocatable::trielemtang1x(:,:)
real(8),allocatable::trielemtang1y(:,:)
real(8),allocatable::trielemtang1z(:,:)
real(8),allocatable::trielemtang2x(:,:)
real(8),allocatable::trielemtang2y(:,:)
real(8),allocatable::trielemtang2z(:,:)
real(8),allocatable::pointoutsidebodyx(:)
real(8),allocatable::pointoutsidebodyy(:)
real(8),allocatable::pointoutsidebodyz(:)
real(8),allocatable::surfarea(:)
integer(4),allocatable::trielemneig(:,:,:)
integer(4),allocatable::totnumtrielem(:)
integer(4),allocatable::flag_outside_marker(:,:)
real(8)::normdirflag
real(8),allocatable::markerarea(:,:)
real(8),allocatable::xmarkerforce(:,:)
real(8),allocatable::ymarkerforce(:,:)
real(8),allocatable::zmarkerforce(:,:)
real(8),allocatable::xmarkershear(:,:)
real(8),allocatable::ymarkershear(:,:)
real(8),allocatable::zmarkershear(:,:)
real(8),allocatable::xmarkerstress(:,:)
real(8),allocatable::ymarkerstress(:,:)
real(8),allocatable::zmarkerstress(:,:)
real(8),allocatable::xyzmarkerforcetmp(:)
real(8),allocatable::xyzmarkerforcetmp_i(:)
real(8),allocatable::xmarkerforceold(:,:)
real(8),allocatable::ymarkerforceold(:,:)
real(8),allocatable::zmarkerforceold(:,:)
real(8),allocatable::xmarkerforce0(:,:)
real(8),allocatable::ymarkerforce0(:,:)
real(8),allocatable::zmarkerforce0(:,:)
end
