﻿!mod$ v1 sum:78fafbc9f1e6a907
!need$ aa821a1d41b92a25 n global_parameters
module gcm_arrays
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
integer(4)::nghost
integer(4),allocatable::inci(:)
integer(4),allocatable::incj(:)
integer(4),allocatable::inck(:)
integer(4),allocatable::ipvt(:)
integer(4),allocatable::closestmarker(:)
integer(4),allocatable::closestelementgc(:)
integer(4),allocatable::ighost(:)
integer(4),allocatable::jghost(:)
integer(4),allocatable::kghost(:)
integer(4),allocatable::icellindex(:)
integer(4),allocatable::jcellindex(:)
integer(4),allocatable::kcellindex(:)
integer(4),allocatable::idead(:)
integer(4),allocatable::jdead(:)
integer(4),allocatable::kdead(:)
integer(4),allocatable::ideadcellindex(:)
integer(4),allocatable::jdeadcellindex(:)
integer(4),allocatable::kdeadcellindex(:)
integer(4),allocatable::closestmarkerdead(:)
integer(4),allocatable::closestelementdead(:)
real(8)::det(1_8:2_8)
real(8),allocatable::work(:)
real(8),allocatable::closestmarkerratio(:)
real(8),allocatable::probelength(:)
real(8),allocatable::xbodyintercepttang(:)
real(8),allocatable::ybodyintercepttang(:)
real(8),allocatable::zbodyintercepttang(:)
real(8),allocatable::xbodyinterceptnorm(:)
real(8),allocatable::ybodyinterceptnorm(:)
real(8),allocatable::zbodyinterceptnorm(:)
real(8),allocatable::xbodyintercept(:)
real(8),allocatable::ybodyintercept(:)
real(8),allocatable::zbodyintercept(:)
real(8),allocatable::ubodyintercept(:)
real(8),allocatable::vbodyintercept(:)
real(8),allocatable::wbodyintercept(:)
real(8),allocatable::pbodyintercept(:)
real(8),allocatable::dpdnbodyintercept(:)
real(8),allocatable::dpdtbodyintercept(:)
real(8),allocatable::ximagepoint(:)
real(8),allocatable::yimagepoint(:)
real(8),allocatable::zimagepoint(:)
real(8),allocatable::res_laplacian(:)
real(8),allocatable::alphaghost(:)
real(8),allocatable::closestmarkerratiodead(:)
real(8),allocatable::xbodyinterceptdead(:)
real(8),allocatable::ybodyinterceptdead(:)
real(8),allocatable::zbodyinterceptdead(:)
real(8),allocatable::ubodyinterceptdead(:)
real(8),allocatable::vbodyinterceptdead(:)
real(8),allocatable::wbodyinterceptdead(:)
real(8),allocatable::xnorminterceptdead(:)
real(8),allocatable::ynorminterceptdead(:)
real(8),allocatable::znorminterceptdead(:)
real(8),allocatable::dpdninterceptdead(:)
real(8),allocatable::probelengthdead(:)
real(8),allocatable::alphadead(:)
real(8),allocatable::ghostnscoeff(:,:)
real(8),allocatable::coeffgcmd(:,:)
real(8),allocatable::coeffgcmn(:,:)
real(8),allocatable::coeffgcmdeadd(:,:)
real(8),allocatable::coeffgcmdeadn(:,:)
real(8),allocatable::coeffgcmdirc(:)
real(8),allocatable::coeffgcmneum(:)
real(8),allocatable::vanmatrixd(:,:)
real(8),allocatable::vanmatrixn(:,:)
real(8),allocatable::xbodycentroid(:,:)
real(8),allocatable::ybodycentroid(:,:)
real(8),allocatable::sbodycentroid(:,:)
real(8),allocatable::xcentroidtang(:,:)
real(8),allocatable::ycentroidtang(:,:)
real(8),allocatable::xcentroidnorm(:,:)
real(8),allocatable::ycentroidnorm(:,:)
integer(4),allocatable::inti(:)
integer(4),allocatable::intj(:)
integer(4),allocatable::intk(:)
end
