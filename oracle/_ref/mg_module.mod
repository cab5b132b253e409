﻿!mod$ v1 sum:281c3826fd9e5dbf
!need$ 0a9719ac12704417 n flow_parameters
!need$ aa821a1d41b92a25 n global_parameters
module mg_module
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
use flow_parameters,only:sidw
use flow_parameters,only:pi
use flow_parameters,only:nread
use flow_parameters,only:ndim
use flow_parameters,only:flow_type
use flow_parameters,only:nx
use flow_parameters,only:ny
use flow_parameters,only:nz
use flow_parameters,only:xgrid_unif
use flow_parameters,only:ygrid_unif
use flow_parameters,only:zgrid_unif
use flow_parameters,only:bcx1
use flow_parameters,only:bcx2
use flow_parameters,only:bcy1
use flow_parameters,only:bcy2
use flow_parameters,only:bcz1
use flow_parameters,only:bcz2
use flow_parameters,only:no_tsteps
use flow_parameters,only:nmonitor
use flow_parameters,only:ndump
use flow_parameters,only:nrestart
use flow_parameters,only:nstat
use flow_parameters,only:nmonitor_probe_liftdrag
use flow_parameters,only:format_dump
use flow_parameters,only:ntime
use flow_parameters,only:ntime_start
use flow_parameters,only:it_solver_type
use flow_parameters,only:advec_scheme
use flow_parameters,only:boundary_motion
use flow_parameters,only:nbody
use flow_parameters,only:body_type
use flow_parameters,only:nptsmax
use flow_parameters,only:nghostmax
use flow_parameters,only:ndeadmax
use flow_parameters,only:boundary_formulation
use flow_parameters,only:mix_gc_form
use flow_parameters,only:elastic_present
use flow_parameters,only:itermax_pson
use flow_parameters,only:itermax_ad
use flow_parameters,only:itermax_gc
use flow_parameters,only:ifubodyin
use flow_parameters,only:ifuinput
use flow_parameters,only:ifuiblankin
use flow_parameters,only:ifumarkerin
use flow_parameters,only:ifuprobein
use flow_parameters,only:ifurstrtflowin
use flow_parameters,only:ifurstrtbodyin
use flow_parameters,only:ifuunstrucsurfin
use flow_parameters,only:ifuunstrucsurfout
use flow_parameters,only:ifubodyout
use flow_parameters,only:ifudragout
use flow_parameters,only:ifufreshcellout
use flow_parameters,only:ifumarkerout
use flow_parameters,only:ifuprobeout
use flow_parameters,only:ifurstrtflowout
use flow_parameters,only:ifurstrtbodyout
use flow_parameters,only:ifustatout
use flow_parameters,only:ifustatplot
use flow_parameters,only:internal_boundary_present
use flow_parameters,only:frac_step_type
use flow_parameters,only:nptsbodymarkerorig
use flow_parameters,only:canonical_body_type
use flow_parameters,only:body_dim
use flow_parameters,only:boundary_motion_type
use flow_parameters,only:wall_type
use flow_parameters,only:unstruc_surface_type
use flow_parameters,only:nptsbodymarker
use flow_parameters,only:xorigin
use flow_parameters,only:yorigin
use flow_parameters,only:zorigin
use flow_parameters,only:xout
use flow_parameters,only:yout
use flow_parameters,only:zout
use flow_parameters,only:uinit
use flow_parameters,only:vinit
use flow_parameters,only:winit
use flow_parameters,only:ux1
use flow_parameters,only:ux2
use flow_parameters,only:vx1
use flow_parameters,only:vx2
use flow_parameters,only:wx1
use flow_parameters,only:wx2
use flow_parameters,only:uy1
use flow_parameters,only:uy2
use flow_parameters,only:vy1
use flow_parameters,only:vy2
use flow_parameters,only:wy1
use flow_parameters,only:wy2
use flow_parameters,only:uz1
use flow_parameters,only:uz2
use flow_parameters,only:vz1
use flow_parameters,only:vz2
use flow_parameters,only:wz1
use flow_parameters,only:wz2
use flow_parameters,only:freq_ux1
use flow_parameters,only:freq_vx1
use flow_parameters,only:freq_wx1
use flow_parameters,only:freq_ux2
use flow_parameters,only:freq_vx2
use flow_parameters,only:freq_wx2
use flow_parameters,only:freq_uy1
use flow_parameters,only:freq_vy1
use flow_parameters,only:freq_wy1
use flow_parameters,only:freq_uy2
use flow_parameters,only:freq_vy2
use flow_parameters,only:freq_wy2
use flow_parameters,only:freq_uz1
use flow_parameters,only:freq_vz1
use flow_parameters,only:freq_wz1
use flow_parameters,only:freq_uz2
use flow_parameters,only:freq_vz2
use flow_parameters,only:freq_wz2
use flow_parameters,only:re
use flow_parameters,only:dt
use flow_parameters,only:reinv
use flow_parameters,only:dtinv
use flow_parameters,only:restol_ad
use flow_parameters,only:restol_gcp
use flow_parameters,only:restol_gcv
use flow_parameters,only:omega_pson
use flow_parameters,only:omega_ad
use flow_parameters,only:restol_pson
use flow_parameters,only:restol_phiprime
use flow_parameters,only:restol_div
use flow_parameters,only:resdiv
use flow_parameters,only:time
use flow_parameters,only:time_start
use flow_parameters,only:bodyinterceptweight
use flow_parameters,only:imagepointweight
use flow_parameters,only:probelengthnormalized
use flow_parameters,only:probelengthnormalizedd
use flow_parameters,only:gcmflag
use flow_parameters,only:membrane_tkns_factor
use flow_parameters,only:membrane_tkns
use flow_parameters,only:area_left
use flow_parameters,only:area_right
use flow_parameters,only:area_bot
use flow_parameters,only:area_top
use flow_parameters,only:area_back
use flow_parameters,only:area_front
use flow_parameters,only:outflow_area
use flow_parameters,only:n_theta
use flow_parameters,only:n_phi
use flow_parameters,only:radiusx
use flow_parameters,only:radiusy
use flow_parameters,only:radiusz
use flow_parameters,only:vxcent
use flow_parameters,only:vycent
use flow_parameters,only:vzcent
use flow_parameters,only:angvx
use flow_parameters,only:angvy
use flow_parameters,only:angvz
use flow_parameters,only:xcent
use flow_parameters,only:ycent
use flow_parameters,only:zcent
use flow_parameters,only:angvx_old
use flow_parameters,only:angvy_old
use flow_parameters,only:angvz_old
use flow_parameters,only:xcentinit
use flow_parameters,only:ycentinit
use flow_parameters,only:zcentinit
use flow_parameters,only:vxcenttrans
use flow_parameters,only:vycenttrans
use flow_parameters,only:vzcenttrans
use flow_parameters,only:ampx
use flow_parameters,only:ampy
use flow_parameters,only:ampz
use flow_parameters,only:freqx
use flow_parameters,only:freqy
use flow_parameters,only:freqz
use flow_parameters,only:angvxinit
use flow_parameters,only:angvyinit
use flow_parameters,only:angvzinit
use flow_parameters,only:ampangx
use flow_parameters,only:ampangy
use flow_parameters,only:ampangz
use flow_parameters,only:freqangx
use flow_parameters,only:freqangy
use flow_parameters,only:freqangz
use flow_parameters,only:angx
use flow_parameters,only:angy
use flow_parameters,only:angz
use flow_parameters,only:angxinit
use flow_parameters,only:angyinit
use flow_parameters,only:angzinit
use flow_parameters,only:xpivot
use flow_parameters,only:ypivot
use flow_parameters,only:zpivot
use flow_parameters,only:vxphase
use flow_parameters,only:vyphase
use flow_parameters,only:vzphase
use flow_parameters,only:angxphase
use flow_parameters,only:angyphase
use flow_parameters,only:angzphase
use flow_parameters,only:forced_motion_spec
use flow_parameters,only:xcentconstr
use flow_parameters,only:ycentconstr
use flow_parameters,only:zcentconstr
use flow_parameters,only:density_fluid
use flow_parameters,only:density_solid
use flow_parameters,only:zoom_factor
use flow_parameters,only:is_pbc_homogeneous
use flow_parameters,only:pbcx1
use flow_parameters,only:pbcx2
use flow_parameters,only:pbcy1
use flow_parameters,only:pbcy2
use flow_parameters,only:pbcz1
use flow_parameters,only:pbcz2
use flow_parameters,only:pppx1
use flow_parameters,only:pppx2
use flow_parameters,only:pppy1
use flow_parameters,only:pppy2
use flow_parameters,only:pppz1
use flow_parameters,only:pppz2
use flow_parameters,only:smooth_divergence
use flow_parameters,only:alfa_upwind
use flow_parameters,only:x1_upwind
use flow_parameters,only:alfa
use flow_parameters,only:vper
use flow_parameters,only:imonitor_probe
use flow_parameters,only:nprobeflow
use flow_parameters,only:iprobe
use flow_parameters,only:jprobe
use flow_parameters,only:kprobe
use flow_parameters,only:imonitor_drag
use flow_parameters,only:nintpdatamax
use flow_parameters,only:irowmax
use flow_parameters,only:rk_step
use flow_parameters,only:rk_bet
use flow_parameters,only:rk_gam
use flow_parameters,only:rk_zet
use flow_parameters,only:i_xx
use flow_parameters,only:i_yy
use flow_parameters,only:i_zz
use flow_parameters,only:i_xy
use flow_parameters,only:i_yz
use flow_parameters,only:i_xz
use flow_parameters,only:moment_x
use flow_parameters,only:moment_y
use flow_parameters,only:moment_z
use flow_parameters,only:force_x
use flow_parameters,only:force_y
use flow_parameters,only:force_z
use flow_parameters,only:cxt
use flow_parameters,only:cyt
use flow_parameters,only:czt
use flow_parameters,only:cxs
use flow_parameters,only:cys
use flow_parameters,only:czs
use flow_parameters,only:cmxt
use flow_parameters,only:cmyt
use flow_parameters,only:cmzt
use flow_parameters,only:cmxs
use flow_parameters,only:cmys
use flow_parameters,only:cmzs
use flow_parameters,only:vq1_i
use flow_parameters,only:vq1
use flow_parameters,only:vq2_i
use flow_parameters,only:vq2
use flow_parameters,only:vq3_i
use flow_parameters,only:vq3
use flow_parameters,only:vq4_i
use flow_parameters,only:vq4
use flow_parameters,only:goa_area_i
use flow_parameters,only:goa_area
integer(4),parameter::max_grid_level=10_4
integer(4)::mgrids_x
integer(4)::mgrids_y
integer(4)::mgrids_z
integer(4)::mlevel
integer(4)::mgcyclex
integer(4)::mgcycley
integer(4)::mgcyclez
integer(4)::infoconv
integer(4)::incrlev
integer(4)::iterfinest
integer(4)::iterinter
integer(4)::itercoarest
integer(4)::ittt1
integer(4)::iredblack
integer(4)::tncolorx
integer(4)::tncolory
integer(4)::tncolorz
integer(4)::istep
integer(4)::jstep
integer(4)::kstep
real(8)::pmgx1
real(8)::pmgx2
real(8)::pmgy1
real(8)::pmgy2
real(8)::pmgz1
real(8)::pmgz2
integer(4),allocatable::mgrid_i(:)
integer(4),allocatable::mgrid_j(:)
integer(4),allocatable::mgrid_k(:)
integer(4),allocatable::zc_start_mg(:)
integer(4),allocatable::zc_end_mg(:)
integer(4),allocatable::z_start_mg(:)
integer(4),allocatable::z_end_mg(:)
integer(4),allocatable::zb1_mg(:)
integer(4),allocatable::zb2_mg(:)
integer(4),allocatable::num_slicesz_mg(:)
integer(4),allocatable::l_proc_mg(:)
integer(4),allocatable::r_proc_mg(:)
integer(4),allocatable::leftmost_mg(:)
integer(4),allocatable::rightmost_mg(:)
integer(4),allocatable::all_slicesy_mg(:,:)
integer(4),allocatable::all_slicesz_mg(:,:)
type::mg_array_type
real(8),allocatable::dxcinv(:)
real(8),allocatable::dycinv(:)
real(8),allocatable::dzcinv(:)
real(8),allocatable::dxinv(:)
real(8),allocatable::dyinv(:)
real(8),allocatable::dzinv(:)
real(8),allocatable::x(:)
real(8),allocatable::xc(:)
real(8),allocatable::y(:)
real(8),allocatable::yc(:)
real(8),allocatable::z(:)
real(8),allocatable::zc(:)
real(8),allocatable::rhs(:,:,:)
real(8),allocatable::phi(:,:,:)
real(8),allocatable::res(:,:,:)
integer(1),allocatable::ibk(:,:,:)
integer(1),allocatable::igk(:,:,:)
end type
type(mg_array_type),allocatable::mg_array(:)
integer(4),allocatable::yc_start_mg(:)
integer(4),allocatable::yc_end_mg(:)
integer(4),allocatable::y_start_mg(:)
integer(4),allocatable::y_end_mg(:)
integer(4),allocatable::yb1_mg(:)
integer(4),allocatable::yb2_mg(:)
integer(4),allocatable::num_slicesy_mg(:)
integer(4),allocatable::b_proc_mg(:)
integer(4),allocatable::t_proc_mg(:)
integer(4),allocatable::btommost_mg(:)
integer(4),allocatable::topmost_mg(:)
end
