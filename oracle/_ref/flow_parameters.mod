﻿!mod$ v1 sum:0a9719ac12704417
!need$ aa821a1d41b92a25 n global_parameters
module flow_parameters
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
real(8),parameter::sidw=2._8
real(8)::pi
integer(4)::nread
integer(4)::ndim
integer(4)::flow_type
integer(4)::nx
integer(4)::ny
integer(4)::nz
integer(4)::xgrid_unif
integer(4)::ygrid_unif
integer(4)::zgrid_unif
integer(4)::bcx1
integer(4)::bcx2
integer(4)::bcy1
integer(4)::bcy2
integer(4)::bcz1
integer(4)::bcz2
integer(4)::no_tsteps
integer(4)::nmonitor
integer(4)::ndump
integer(4)::nrestart
integer(4)::nstat
integer(4)::nmonitor_probe_liftdrag
integer(4)::format_dump
integer(4)::ntime
integer(4)::ntime_start
integer(4)::it_solver_type
integer(4)::advec_scheme
integer(4)::boundary_motion
integer(4)::nbody
integer(4)::body_type
integer(4)::nptsmax
integer(4)::nghostmax
integer(4)::ndeadmax
integer(4)::boundary_formulation
integer(4)::mix_gc_form
integer(4)::elastic_present
integer(4)::itermax_pson
integer(4)::itermax_ad
integer(4)::itermax_gc
integer(4)::ifubodyin
integer(4)::ifuinput
integer(4)::ifuiblankin
integer(4)::ifumarkerin
integer(4)::ifuprobein
integer(4)::ifurstrtflowin
integer(4)::ifurstrtbodyin
integer(4)::ifuunstrucsurfin
integer(4)::ifuunstrucsurfout
integer(4)::ifubodyout
integer(4)::ifudragout
integer(4)::ifufreshcellout
integer(4)::ifumarkerout
integer(4)::ifuprobeout
integer(4)::ifurstrtflowout
integer(4)::ifurstrtbodyout
integer(4)::ifustatout
integer(4)::ifustatplot
integer(4)::internal_boundary_present
integer(4)::frac_step_type
integer(4),allocatable::nptsbodymarkerorig(:)
integer(4),allocatable::canonical_body_type(:)
integer(4),allocatable::body_dim(:)
integer(4),allocatable::boundary_motion_type(:)
integer(4),allocatable::wall_type(:)
integer(4),allocatable::unstruc_surface_type(:)
integer(4),allocatable::nptsbodymarker(:)
real(8)::xorigin
real(8)::yorigin
real(8)::zorigin
real(8)::xout
real(8)::yout
real(8)::zout
real(8)::uinit
real(8)::vinit
real(8)::winit
real(8)::ux1
real(8)::ux2
real(8)::vx1
real(8)::vx2
real(8)::wx1
real(8)::wx2
real(8)::uy1
real(8)::uy2
real(8)::vy1
real(8)::vy2
real(8)::wy1
real(8)::wy2
real(8)::uz1
real(8)::uz2
real(8)::vz1
real(8)::vz2
real(8)::wz1
real(8)::wz2
real(8)::freq_ux1
real(8)::freq_vx1
real(8)::freq_wx1
real(8)::freq_ux2
real(8)::freq_vx2
real(8)::freq_wx2
real(8)::freq_uy1
real(8)::freq_vy1
real(8)::freq_wy1
real(8)::freq_uy2
real(8)::freq_vy2
real(8)::freq_wy2
real(8)::freq_uz1
real(8)::freq_vz1
real(8)::freq_wz1
real(8)::freq_uz2
real(8)::freq_vz2
real(8)::freq_wz2
real(8)::re
real(8)::dt
real(8)::reinv
real(8)::dtinv
real(8)::restol_ad
real(8)::restol_gcp
real(8)::restol_gcv
real(8)::omega_pson
real(8)::omega_ad
real(8)::restol_pson
real(8)::restol_phiprime
real(8)::restol_div
real(8)::resdiv
real(8)::time
real(8)::time_start
real(8)::bodyinterceptweight
real(8)::imagepointweight
real(8)::probelengthnormalized
real(8)::probelengthnormalizedd
real(8)::gcmflag
real(8)::membrane_tkns_factor
real(8)::membrane_tkns
real(8)::area_left
real(8)::area_right
real(8)::area_bot
real(8)::area_top
real(8)::area_back
real(8)::area_front
real(8)::outflow_area
integer(4),allocatable::n_theta(:)
integer(4),allocatable::n_phi(:)
real(8),allocatable::radiusx(:)
real(8),allocatable::radiusy(:)
real(8),allocatable::radiusz(:)
real(8),allocatable::vxcent(:)
real(8),allocatable::vycent(:)
real(8),allocatable::vzcent(:)
real(8),allocatable::angvx(:)
real(8),allocatable::angvy(:)
real(8),allocatable::angvz(:)
real(8),allocatable::xcent(:)
real(8),allocatable::ycent(:)
real(8),allocatable::zcent(:)
real(8),allocatable::angvx_old(:)
real(8),allocatable::angvy_old(:)
real(8),allocatable::angvz_old(:)
real(8),allocatable::xcentinit(:)
real(8),allocatable::ycentinit(:)
real(8),allocatable::zcentinit(:)
real(8),allocatable::vxcenttrans(:)
real(8),allocatable::vycenttrans(:)
real(8),allocatable::vzcenttrans(:)
real(8),allocatable::ampx(:)
real(8),allocatable::ampy(:)
real(8),allocatable::ampz(:)
real(8),allocatable::freqx(:)
real(8),allocatable::freqy(:)
real(8),allocatable::freqz(:)
real(8),allocatable::angvxinit(:)
real(8),allocatable::angvyinit(:)
real(8),allocatable::angvzinit(:)
real(8),allocatable::ampangx(:)
real(8),allocatable::ampangy(:)
real(8),allocatable::ampangz(:)
real(8),allocatable::freqangx(:)
real(8),allocatable::freqangy(:)
real(8),allocatable::freqangz(:)
real(8),allocatable::angx(:)
real(8),allocatable::angy(:)
real(8),allocatable::angz(:)
real(8),allocatable::angxinit(:)
real(8),allocatable::angyinit(:)
real(8),allocatable::angzinit(:)
real(8),allocatable::xpivot(:)
real(8),allocatable::ypivot(:)
real(8),allocatable::zpivot(:)
real(8),allocatable::vxphase(:)
real(8),allocatable::vyphase(:)
real(8),allocatable::vzphase(:)
real(8),allocatable::angxphase(:)
real(8),allocatable::angyphase(:)
real(8),allocatable::angzphase(:)
integer(4),allocatable::forced_motion_spec(:)
real(8),allocatable::xcentconstr(:)
real(8),allocatable::ycentconstr(:)
real(8),allocatable::zcentconstr(:)
real(8)::density_fluid
real(8),allocatable::density_solid(:)
real(8),allocatable::zoom_factor(:)
integer(4)::is_pbc_homogeneous
integer(4)::pbcx1
integer(4)::pbcx2
integer(4)::pbcy1
integer(4)::pbcy2
integer(4)::pbcz1
integer(4)::pbcz2
real(8)::pppx1
real(8)::pppx2
real(8)::pppy1
real(8)::pppy2
real(8)::pppz1
real(8)::pppz2
logical(4)::smooth_divergence
real(8)::alfa_upwind
real(8)::x1_upwind
real(8)::alfa
real(8)::vper
integer(4)::imonitor_probe
integer(4)::nprobeflow
integer(4),allocatable::iprobe(:)
integer(4),allocatable::jprobe(:)
integer(4),allocatable::kprobe(:)
integer(4)::imonitor_drag
integer(4),parameter::nintpdatamax=20_4
integer(4),parameter::irowmax=8_4
integer(4)::rk_step
real(8)::rk_bet(1_8:3_8)
real(8)::rk_gam(1_8:3_8)
real(8)::rk_zet(1_8:3_8)
real(8)::i_xx
real(8)::i_yy
real(8)::i_zz
real(8)::i_xy
real(8)::i_yz
real(8)::i_xz
real(8)::moment_x
real(8)::moment_y
real(8)::moment_z
real(8)::force_x
real(8)::force_y
real(8)::force_z
real(8),allocatable::cxt(:)
real(8),allocatable::cyt(:)
real(8),allocatable::czt(:)
real(8),allocatable::cxs(:)
real(8),allocatable::cys(:)
real(8),allocatable::czs(:)
real(8),allocatable::cmxt(:)
real(8),allocatable::cmyt(:)
real(8),allocatable::cmzt(:)
real(8),allocatable::cmxs(:)
real(8),allocatable::cmys(:)
real(8),allocatable::cmzs(:)
real(8)::vq1_i
real(8)::vq1
real(8)::vq2_i
real(8)::vq2
real(8)::vq3_i
real(8)::vq3
real(8)::vq4_i
real(8)::vq4
real(8)::goa_area_i
real(8)::goa_area
end
