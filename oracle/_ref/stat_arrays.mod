﻿!mod$ v1 sum:79c2f3e154d73868
!need$ aa821a1d41b92a25 n global_parameters
module stat_arrays
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
integer(4)::statctr
real(8),allocatable::uav(:,:,:)
real(8),allocatable::vav(:,:,:)
real(8),allocatable::wav(:,:,:)
real(8),allocatable::pav(:,:,:)
real(8),allocatable::uvav(:,:,:)
real(8),allocatable::vwav(:,:,:)
real(8),allocatable::uwav(:,:,:)
real(8),allocatable::uuav(:,:,:)
real(8),allocatable::vvav(:,:,:)
real(8),allocatable::wwav(:,:,:)
end
