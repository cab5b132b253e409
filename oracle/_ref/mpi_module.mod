﻿!mod$ v1 sum:975adc689f400ddf
!need$ aa821a1d41b92a25 n global_parameters
module mpi_module
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
integer(4),parameter::mpi_source=3_4
integer(4),parameter::mpi_tag=4_4
integer(4),parameter::mpi_error=5_4
integer(4),parameter::mpi_status_size=5_4
integer(4)::mpi_status_ignore(1_8:5_8)
integer(4)::mpi_statuses_ignore(1_8:5_8,1_8:1_8)
integer(4)::mpi_errcodes_ignore(1_8:1_8)
character(1_8,1)::mpi_argvs_null(1_8:1_8,1_8:1_8)
character(1_8,1)::mpi_argv_null(1_8:1_8)
integer(4),parameter::mpi_success=0_4
integer(4),parameter::mpi_err_other=15_4
integer(4),parameter::mpi_err_count=2_4
integer(4),parameter::mpi_err_spawn=42_4
integer(4),parameter::mpi_err_locktype=47_4
integer(4),parameter::mpi_err_op=9_4
integer(4),parameter::mpi_err_dup_datarep=24_4
integer(4),parameter::mpi_err_unsupported_datarep=43_4
integer(4),parameter::mpi_err_truncate=14_4
integer(4),parameter::mpi_err_info_nokey=31_4
integer(4),parameter::mpi_err_assert=53_4
integer(4),parameter::mpi_err_file_exists=25_4
integer(4),parameter::mpi_err_pending=18_4
integer(4),parameter::mpi_err_comm=5_4
integer(4),parameter::mpi_err_keyval=48_4
integer(4),parameter::mpi_err_name=33_4
integer(4),parameter::mpi_err_request=19_4
integer(4),parameter::mpi_err_type=3_4
integer(4),parameter::mpi_err_info_value=30_4
integer(4),parameter::mpi_err_rma_sync=50_4
integer(4),parameter::mpi_err_no_mem=34_4
integer(4),parameter::mpi_err_bad_file=22_4
integer(4),parameter::mpi_err_quota=39_4
integer(4),parameter::mpi_err_root=7_4
integer(4),parameter::mpi_err_service=41_4
integer(4),parameter::mpi_err_io=32_4
integer(4),parameter::mpi_err_rma_flavor=58_4
integer(4),parameter::mpi_err_access=20_4
integer(4),parameter::mpi_err_no_space=36_4
integer(4),parameter::mpi_err_conversion=23_4
integer(4),parameter::mpi_err_win=45_4
integer(4),parameter::mpi_err_file=27_4
integer(4),parameter::mpi_err_rma_shared=57_4
integer(4),parameter::mpi_err_base=46_4
integer(4),parameter::mpi_err_rma_conflict=49_4
integer(4),parameter::mpi_err_in_status=17_4
integer(4),parameter::mpi_err_info_key=29_4
integer(4),parameter::mpi_err_arg=12_4
integer(4),parameter::mpi_err_read_only=40_4
integer(4),parameter::mpi_err_size=51_4
integer(4),parameter::mpi_err_buffer=1_4
integer(4),parameter::mpi_err_lastcode=1073741823_4
integer(4),parameter::mpi_err_disp=52_4
integer(4),parameter::mpi_err_port=38_4
integer(4),parameter::mpi_err_group=8_4
integer(4),parameter::mpi_err_topology=10_4
integer(4),parameter::mpi_err_tag=4_4
integer(4),parameter::mpi_err_not_same=35_4
integer(4),parameter::mpi_err_info=28_4
integer(4),parameter::mpi_err_unknown=13_4
integer(4),parameter::mpi_err_file_in_use=26_4
integer(4),parameter::mpi_err_rma_attach=56_4
integer(4),parameter::mpi_err_unsupported_operation=44_4
integer(4),parameter::mpi_err_amode=21_4
integer(4),parameter::mpi_err_rank=6_4
integer(4),parameter::mpi_err_dims=11_4
integer(4),parameter::mpi_err_no_such_file=37_4
integer(4),parameter::mpi_err_rma_range=55_4
integer(4),parameter::mpi_err_intern=16_4
integer(4),parameter::mpi_errors_are_fatal=1409286144_4
integer(4),parameter::mpi_errors_return=1409286145_4
integer(4),parameter::mpi_ident=0_4
integer(4),parameter::mpi_congruent=1_4
integer(4),parameter::mpi_similar=2_4
integer(4),parameter::mpi_unequal=3_4
integer(4),parameter::mpi_win_flavor_create=1_4
integer(4),parameter::mpi_win_flavor_allocate=2_4
integer(4),parameter::mpi_win_flavor_dynamic=3_4
integer(4),parameter::mpi_win_flavor_shared=4_4
integer(4),parameter::mpi_win_separate=1_4
integer(4),parameter::mpi_win_unified=2_4
integer(4),parameter::mpi_max=1476395009_4
integer(4),parameter::mpi_min=1476395010_4
integer(4),parameter::mpi_sum=1476395011_4
integer(4),parameter::mpi_prod=1476395012_4
integer(4),parameter::mpi_land=1476395013_4
integer(4),parameter::mpi_band=1476395014_4
integer(4),parameter::mpi_lor=1476395015_4
integer(4),parameter::mpi_bor=1476395016_4
integer(4),parameter::mpi_lxor=1476395017_4
integer(4),parameter::mpi_bxor=1476395018_4
integer(4),parameter::mpi_minloc=1476395019_4
integer(4),parameter::mpi_maxloc=1476395020_4
integer(4),parameter::mpi_replace=1476395021_4
integer(4),parameter::mpi_no_op=1476395022_4
integer(4),parameter::mpi_comm_world=1140850688_4
integer(4),parameter::mpi_comm_self=1140850689_4
integer(4),parameter::mpi_group_empty=1207959552_4
integer(4),parameter::mpi_comm_null=67108864_4
integer(4),parameter::mpi_win_null=536870912_4
integer(4),parameter::mpi_file_null=0_4
integer(4),parameter::mpi_group_null=134217728_4
integer(4),parameter::mpi_op_null=402653184_4
integer(4),parameter::mpi_datatype_null=201326592_4
integer(4),parameter::mpi_request_null=738197504_4
integer(4),parameter::mpi_errhandler_null=335544320_4
integer(4),parameter::mpi_info_null=469762048_4
integer(4),parameter::mpi_info_env=1543503873_4
integer(4),parameter::mpi_tag_ub=1681915906_4
integer(4),parameter::mpi_host=1681915908_4
integer(4),parameter::mpi_io=1681915910_4
integer(4),parameter::mpi_wtime_is_global=1681915912_4
integer(4),parameter::mpi_universe_size=1681915914_4
integer(4),parameter::mpi_lastusedcode=1681915916_4
integer(4),parameter::mpi_appnum=1681915918_4
integer(4),parameter::mpi_win_base=1711276034_4
integer(4),parameter::mpi_win_size=1711276036_4
integer(4),parameter::mpi_win_disp_unit=1711276038_4
integer(4),parameter::mpi_win_create_flavor=1711276040_4
integer(4),parameter::mpi_win_model=1711276042_4
integer(4),parameter::mpi_max_error_string=511_4
integer(4),parameter::mpi_max_port_name=255_4
integer(4),parameter::mpi_max_object_name=127_4
integer(4),parameter::mpi_max_info_key=254_4
integer(4),parameter::mpi_max_info_val=1023_4
integer(4),parameter::mpi_max_processor_name=127_4
integer(4),parameter::mpi_max_datarep_string=127_4
integer(4),parameter::mpi_max_library_version_string=8191_4
integer(4),parameter::mpi_undefined=(-32766_4)
integer(4),parameter::mpi_keyval_invalid=603979776_4
integer(4),parameter::mpi_bsend_overhead=96_4
integer(4),parameter::mpi_proc_null=-1_4
integer(4),parameter::mpi_any_source=-2_4
integer(4),parameter::mpi_any_tag=-1_4
integer(4),parameter::mpi_root=-3_4
integer(4),parameter::mpi_graph=1_4
integer(4),parameter::mpi_cart=2_4
integer(4),parameter::mpi_dist_graph=3_4
integer(4),parameter::mpi_version=3_4
integer(4),parameter::mpi_subversion=1_4
integer(4),parameter::mpi_lock_exclusive=234_4
integer(4),parameter::mpi_lock_shared=235_4
integer(4),parameter::mpi_complex=1275070494_4
integer(4),parameter::mpi_double_complex=1275072546_4
integer(4),parameter::mpi_logical=1275069469_4
integer(4),parameter::mpi_real=1275069468_4
integer(4),parameter::mpi_double_precision=1275070495_4
integer(4),parameter::mpi_integer=1275069467_4
integer(4),parameter::mpi_2integer=1275070496_4
integer(4),parameter::mpi_2double_precision=1275072547_4
integer(4),parameter::mpi_2real=1275070497_4
integer(4),parameter::mpi_character=1275068698_4
integer(4),parameter::mpi_byte=1275068685_4
integer(4),parameter::mpi_ub=1275068433_4
integer(4),parameter::mpi_lb=1275068432_4
integer(4),parameter::mpi_packed=1275068687_4
integer(4),parameter::mpi_integer1=1275068717_4
integer(4),parameter::mpi_integer2=1275068975_4
integer(4),parameter::mpi_integer4=1275069488_4
integer(4),parameter::mpi_integer8=1275070513_4
integer(4),parameter::mpi_integer16=201326592_4
integer(4),parameter::mpi_real4=1275069479_4
integer(4),parameter::mpi_real8=1275070505_4
integer(4),parameter::mpi_real16=1275072555_4
integer(4),parameter::mpi_complex8=1275070504_4
integer(4),parameter::mpi_complex16=1275072554_4
integer(4),parameter::mpi_complex32=1275076652_4
integer(4),parameter::mpi_address_kind=8_4
integer(4),parameter::mpi_offset_kind=8_4
integer(4),parameter::mpi_count_kind=8_4
integer(4),parameter::mpi_integer_kind=4_4
integer(4),parameter::mpi_char=1275068673_4
integer(4),parameter::mpi_signed_char=1275068696_4
integer(4),parameter::mpi_unsigned_char=1275068674_4
integer(4),parameter::mpi_wchar=1275069454_4
integer(4),parameter::mpi_short=1275068931_4
integer(4),parameter::mpi_unsigned_short=1275068932_4
integer(4),parameter::mpi_int=1275069445_4
integer(4),parameter::mpi_unsigned=1275069446_4
integer(4),parameter::mpi_long=1275070471_4
integer(4),parameter::mpi_unsigned_long=1275070472_4
integer(4),parameter::mpi_float=1275069450_4
integer(4),parameter::mpi_double=1275070475_4
integer(4),parameter::mpi_long_double=1275072524_4
integer(4),parameter::mpi_long_long_int=1275070473_4
integer(4),parameter::mpi_unsigned_long_long=1275070489_4
integer(4),parameter::mpi_long_long=1275070473_4
integer(4),parameter::mpi_float_int=-1946157056_4
integer(4),parameter::mpi_double_int=-1946157055_4
integer(4),parameter::mpi_long_int=-1946157054_4
integer(4),parameter::mpi_short_int=-1946157053_4
integer(4),parameter::mpi_2int=1275070486_4
integer(4),parameter::mpi_long_double_int=-1946157052_4
integer(4),parameter::mpi_int8_t=1275068727_4
integer(4),parameter::mpi_int16_t=1275068984_4
integer(4),parameter::mpi_int32_t=1275069497_4
integer(4),parameter::mpi_int64_t=1275070522_4
integer(4),parameter::mpi_uint8_t=1275068731_4
integer(4),parameter::mpi_uint16_t=1275068988_4
integer(4),parameter::mpi_uint32_t=1275069501_4
integer(4),parameter::mpi_uint64_t=1275070526_4
integer(4),parameter::mpi_c_bool=1275068735_4
integer(4),parameter::mpi_c_float_complex=1275070528_4
integer(4),parameter::mpi_c_complex=1275070528_4
integer(4),parameter::mpi_c_double_complex=1275072577_4
integer(4),parameter::mpi_c_long_double_complex=1275076674_4
integer(4),parameter::mpi_aint=1275070531_4
integer(4),parameter::mpi_offset=1275070532_4
integer(4),parameter::mpi_count=1275070533_4
integer(4),parameter::mpi_cxx_bool=1275068723_4
integer(4),parameter::mpi_cxx_float_complex=1275070516_4
integer(4),parameter::mpi_cxx_double_complex=1275072565_4
integer(4),parameter::mpi_cxx_long_double_complex=1275076662_4
integer(4),parameter::mpi_combiner_named=1_4
integer(4),parameter::mpi_combiner_dup=2_4
integer(4),parameter::mpi_combiner_contiguous=3_4
integer(4),parameter::mpi_combiner_vector=4_4
integer(4),parameter::mpi_combiner_hvector_integer=5_4
integer(4),parameter::mpi_combiner_hvector=6_4
integer(4),parameter::mpi_combiner_indexed=7_4
integer(4),parameter::mpi_combiner_hindexed_integer=8_4
integer(4),parameter::mpi_combiner_hindexed=9_4
integer(4),parameter::mpi_combiner_indexed_block=10_4
integer(4),parameter::mpi_combiner_struct_integer=11_4
integer(4),parameter::mpi_combiner_struct=12_4
integer(4),parameter::mpi_combiner_subarray=13_4
integer(4),parameter::mpi_combiner_darray=14_4
integer(4),parameter::mpi_combiner_f90_real=15_4
integer(4),parameter::mpi_combiner_f90_complex=16_4
integer(4),parameter::mpi_combiner_f90_integer=17_4
integer(4),parameter::mpi_combiner_resized=18_4
integer(4),parameter::mpi_combiner_hindexed_block=19_4
integer(4),parameter::mpi_typeclass_real=1_4
integer(4),parameter::mpi_typeclass_integer=2_4
integer(4),parameter::mpi_typeclass_complex=3_4
integer(4),parameter::mpi_mode_nocheck=1024_4
integer(4),parameter::mpi_mode_nostore=2048_4
integer(4),parameter::mpi_mode_noput=4096_4
integer(4),parameter::mpi_mode_noprecede=8192_4
integer(4),parameter::mpi_mode_nosucceed=16384_4
integer(4),parameter::mpi_comm_type_shared=1_4
integer(4),parameter::mpi_message_null=738197504_4
integer(4),parameter::mpi_message_no_proc=1811939328_4
integer(4),parameter::mpi_thread_single=0_4
integer(4),parameter::mpi_thread_funneled=1_4
integer(4),parameter::mpi_thread_serialized=2_4
integer(4),parameter::mpi_thread_multiple=3_4
integer(4),parameter::mpi_mode_rdonly=2_4
integer(4),parameter::mpi_mode_rdwr=8_4
integer(4),parameter::mpi_mode_wronly=4_4
integer(4),parameter::mpi_mode_delete_on_close=16_4
integer(4),parameter::mpi_mode_unique_open=32_4
integer(4),parameter::mpi_mode_create=1_4
integer(4),parameter::mpi_mode_excl=64_4
integer(4),parameter::mpi_mode_append=128_4
integer(4),parameter::mpi_mode_sequential=256_4
integer(4),parameter::mpi_seek_set=600_4
integer(4),parameter::mpi_seek_cur=602_4
integer(4),parameter::mpi_seek_end=604_4
integer(4),parameter::mpi_order_c=56_4
integer(4),parameter::mpi_order_fortran=57_4
integer(4),parameter::mpi_distribute_block=121_4
integer(4),parameter::mpi_distribute_cyclic=122_4
integer(4),parameter::mpi_distribute_none=123_4
integer(4),parameter::mpi_distribute_dflt_darg=-49767_4
integer(8),parameter::mpi_displacement_current=-54278278_8
logical(4),parameter::mpi_subarrays_supported=.false._4
logical(4),parameter::mpi_async_protects_nonblocking=.false._4
integer(4)::mpi_bottom
integer(4)::mpi_in_place
integer(4)::mpi_unweighted
integer(4)::mpi_weights_empty
procedure()::mpi_dup_fn
procedure()::mpi_null_delete_fn
procedure()::mpi_null_copy_fn
procedure()::mpi_comm_dup_fn
procedure()::mpi_comm_null_delete_fn
procedure()::mpi_comm_null_copy_fn
procedure()::mpi_win_dup_fn
procedure()::mpi_win_null_delete_fn
procedure()::mpi_win_null_copy_fn
procedure()::mpi_type_dup_fn
procedure()::mpi_type_null_delete_fn
procedure()::mpi_type_null_copy_fn
procedure()::mpi_conversion_fn_null
procedure(real(8))::mpi_wtime
procedure(real(8))::mpi_wtick
procedure(real(8))::pmpi_wtime
procedure(real(8))::pmpi_wtick
integer(4)::nprocs_g
integer(4)::nprocs
integer(4)::lproc_g
integer(4)::lproc
integer(4)::ierr
integer(4)::istatus(1_8:5_8)
integer(4)::proc_m
integer(4)::proc_s(1_8:8_8)
integer(4)::nelastic
integer(4)::world_group
integer(4)::flow_group
integer(4)::flow_comm
integer(4)::z_slab_unif
integer(4)::kslices
integer(4)::zc_start
integer(4)::zc_end
integer(4)::z_start
integer(4)::z_end
integer(4)::zb1
integer(4)::zb2
integer(4)::lproc_leftmost
integer(4)::lproc_rightmost
integer(4)::lproc_left
integer(4)::lproc_right
integer(4),allocatable::numslices(:)
integer(4),parameter::uniform_slab=1_4
integer(4),parameter::nonuniform_slab=2_4
integer(4)::toleft
integer(4)::torght
integer(4)::isend_rq_toleft
integer(4)::isend_rq_torght
integer(4)::irecv_rq_fromleft
integer(4)::irecv_rq_fromrght
integer(4)::isend_stat_tl(1_8:5_8)
integer(4)::isend_stat_tr(1_8:5_8)
integer(4)::irecv_stat_fl(1_8:5_8)
integer(4)::irecv_stat_fr(1_8:5_8)
integer(1),allocatable::i_lbuff(:,:,:)
integer(1),allocatable::i_rbuff(:,:,:)
integer(4)::y_slab_unif
integer(4)::jproc
integer(4)::kproc
integer(4)::jslices
integer(4),allocatable::numslicesy(:)
integer(4),allocatable::numslicesz(:)
integer(4)::yc_start
integer(4)::yc_end
integer(4)::y_start
integer(4)::y_end
integer(4)::yb1
integer(4)::yb2
integer(4)::nprocy
integer(4)::nprocz
integer(4)::lproc_btommost
integer(4)::lproc_topmost
integer(4)::lproc_top
integer(4)::lproc_bottom
integer(4)::tobtom
integer(4)::totop
integer(4)::isend_rq_tobtom
integer(4)::isend_rq_totop
integer(4)::irecv_rq_frombtom
integer(4)::irecv_rq_fromtop
integer(4)::isend_stat_tb(1_8:5_8)
integer(4)::isend_stat_tt(1_8:5_8)
integer(4)::irecv_stat_fb(1_8:5_8)
integer(4)::irecv_stat_ft(1_8:5_8)
common/mpifcmb5/mpi_unweighted
common/mpifcmb9/mpi_weights_empty
common/mpipriv1/mpi_bottom,mpi_in_place,mpi_status_ignore
common/mpipriv2/mpi_statuses_ignore,mpi_errcodes_ignore
common/mpiprivc/mpi_argvs_null,mpi_argv_null
end
