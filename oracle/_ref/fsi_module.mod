﻿!mod$ v1 sum:1507f1e47cafc7c7
!need$ aa821a1d41b92a25 n global_parameters
module fsi_module
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
integer(4)::itermax_fsi
integer(4)::iter_fsi
integer(4)::is_fsi_cvg
integer(4)::nsub
integer(4)::fsi_on
real(8)::restol_fsip
real(8)::restol_fsiv
real(8)::restol_fsidsp
real(8)::resmax_fsip
real(8)::resmax_fsiv
real(8)::resmax_fsidsp
real(8)::resmax_fsiv_old
real(8)::relax_fsip0
real(8)::relax_fsiv0
real(8)::resmax_fsip_old
real(8)::relax_fsip
real(8)::relax_fsiv
real(8)::relax_fsidsp
real(8)::fsi_delay
real(8)::fsi_filter_f
real(8)::fsi_filter_f0
real(8)::fsi_filter_v
integer(4)::ifuelasid
integer(4),allocatable::bodyinflowid(:)
integer(4),allocatable::elasticmarkerid(:,:)
integer(4)::nptsmax_s
integer(4)::nelasticbody
integer(4)::ndata_s
integer(4)::nread_s
integer(4)::nrestart_s
integer(4)::nstep_s
real(8)::dt_s
real(8),allocatable::xdata_s(:,:)
real(8),allocatable::ydata_s(:,:)
real(8),allocatable::zdata_s(:,:)
real(8),allocatable::cnt_fx(:,:)
real(8),allocatable::cnt_fy(:,:)
real(8),allocatable::cnt_fz(:,:)
real(8),allocatable::cnt_force(:,:)
real(8),allocatable::cfx(:,:)
real(8),allocatable::cfy(:,:)
real(8),allocatable::cfz(:,:)
real(8),allocatable::cnt_d12(:,:)
real(8),allocatable::cnt_d21(:,:)
real(8),allocatable::cnt_d13(:,:)
real(8),allocatable::cnt_d31(:,:)
real(8),allocatable::cnt_d23(:,:)
real(8),allocatable::cnt_d32(:,:)
real(8),allocatable::cntd12(:)
real(8),allocatable::cntd21(:)
real(8),allocatable::cntd13(:)
real(8),allocatable::cntd31(:)
real(8),allocatable::cntd23(:)
real(8),allocatable::cntd32(:)
integer(4),allocatable::cnt_em12(:,:)
integer(4),allocatable::cnt_em21(:,:)
integer(4),allocatable::cnt_em13(:,:)
integer(4),allocatable::cnt_em31(:,:)
integer(4),allocatable::cnt_em23(:,:)
integer(4),allocatable::cnt_em32(:,:)
integer(4),allocatable::elemnum12(:)
integer(4),allocatable::elemnum21(:)
integer(4),allocatable::elemnum13(:)
integer(4),allocatable::elemnum31(:)
integer(4),allocatable::elemnum23(:)
integer(4),allocatable::elemnum32(:)
integer(4),allocatable::dotp(:,:)
integer(4),allocatable::dotp2(:,:)
integer(4),allocatable::set_old(:,:)
integer(4),allocatable::cnt_fg12(:,:)
integer(4),allocatable::cntflg12(:,:)
integer(4),allocatable::cnt_fg21(:,:)
integer(4),allocatable::cntflg21(:,:)
integer(4),allocatable::cnt_fg13(:,:)
integer(4),allocatable::cntflg13(:,:)
integer(4),allocatable::cnt_fg31(:,:)
integer(4),allocatable::cntflg31(:,:)
integer(4),allocatable::cnt_fg23(:,:)
integer(4),allocatable::cntflg23(:,:)
integer(4),allocatable::cnt_fg32(:,:)
integer(4),allocatable::cntflg32(:,:)
end
