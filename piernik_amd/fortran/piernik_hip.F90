!
! piernik_hip.F90 -- ISO_C_BINDING interface to libpiernik_hip.so
!
! The Fortran view of include/piernik_hip.h.  PIERNIK's host code uses this
! module from an adapter with the solve_cg_sub signature
! (src/scheme/sweeps.F90:140-151); see the adapter listed in the
! repository's integration notes.  Every function returns 0 on success; on a
! nonzero code the adapter passes pk_error_message() to die(), as the
! reference does for any error.
!
module piernik_hip

   use, intrinsic :: iso_c_binding, only: c_int32_t, c_int64_t, c_double, c_float, c_ptr, c_char, c_null_char, &
        &                                  c_f_pointer, c_associated

   implicit none

   private
   public :: pk_fluid, pk_config, pk_stage_report, pk_profile_stats
   public :: pk_init, pk_finalize, pk_register_block, pk_unregister_block, pk_upload, pk_download, &
        &    pk_upload_from, pk_download_to, pk_solve_cg, pk_sweep, pk_fill_ghosts, pk_set_neighbour, pk_glm_damp, &
        &    pk_timestep_fluid, pk_fluid_update, pk_run, pk_get_report, pk_synchronize, &
        &    pk_autotune, pk_profile_begin, pk_profile_end, pk_error_message, &
        &    pk_upload_device, pk_download_device, pk_timestep_all, pk_fluid_update_all, pk_run_all, &
        &    pk_autotune_all, pk_list_tiles, pk_set_list_tiles, pk_set_list_split, &
        &    pk_adjoint_setup, pk_state_save, pk_state_load, &
        &    pk_adjoint_set, pk_adjoint_get, pk_fluid_update_record, pk_sweep_adjoint, &
        &    pk_fluid_update_adjoint, pk_timestep_adjoint, pk_ensemble_ids, pk_ensemble_step, pk_ensemble_run, &
        &    pk_ensemble_state_save, pk_ensemble_state_load, pk_ensemble_record_step, &
        &    pk_ensemble_step_adjoint, pk_ensemble_timestep_adjoint, &
        &    pk_unsplit_stage, pk_fluid_update_unsplit, pk_run_unsplit, pk_set_cfl_control, pk_redone_steps, &
        &    pk_register_fcflx, pk_datafield, pk_datafield_f32, pk_set_time_control
   public :: PK_EULER, PK_RK2_1, PK_RK2_2, PK_LIM_VANLEER, PK_LIM_MINMOD, PK_LIM_MONCEN, PK_LIM_SUPERBEE
   public :: PK_FC_FINEBND, PK_FC_COARSEBND
   public :: PK_CFL_WARN, PK_CFL_REDO, PK_CFL_FLEXIBLE, PK_CFL_AUTO
   public :: PK_BATCH_GRAPH, PK_BATCH_PERSISTENT, PK_BATCH_PER_BLOCK

   ! cflcontrol of pk_set_cfl_control: 'warn'/'none', 'redo', 'flexible', 'auto' (timestep.F90:73-90)
   integer(c_int32_t), parameter :: PK_CFL_WARN = 0, PK_CFL_REDO = 1, PK_CFL_FLEXIBLE = 2, PK_CFL_AUTO = 3

   ! kinds of pk_register_fcflx: cg%finebnd, cg%coarsebnd (grid_container_fcflx.F90:40-41)
   integer(c_int32_t), parameter :: PK_FC_FINEBND = 0, PK_FC_COARSEBND = 1

   ! the reference's stage enumerators (src/base/constants.F90:230-232)
   integer(c_int32_t), parameter :: PK_EULER = 10000, PK_RK2_1 = 10001, PK_RK2_2 = 10002
   integer(c_int32_t), parameter :: PK_LIM_VANLEER = 0, PK_LIM_MINMOD = 1, PK_LIM_MONCEN = 2, PK_LIM_SUPERBEE = 3
   integer(c_int32_t), parameter :: PK_BATCH_GRAPH = 0, PK_BATCH_PERSISTENT = 1, PK_BATCH_PER_BLOCK = 2

   type, bind(C) :: pk_fluid
      real(c_double)     :: gamma
      integer(c_int32_t) :: magnetized
      integer(c_int32_t) :: reserved
   end type pk_fluid

   type, bind(C) :: pk_config
      integer(c_int32_t) :: nfluids
      integer(c_int32_t) :: limiter
      integer(c_int32_t) :: limiter_b
      integer(c_int32_t) :: integration_order
      type(pk_fluid)     :: fluid(2)
      real(c_double)     :: smalld
      real(c_double)     :: smallei
      integer(c_int32_t) :: use_smalld
      integer(c_int32_t) :: use_smallei
      real(c_double)     :: cfl
      real(c_double)     :: cfl_glm
      real(c_double)     :: glm_alpha
      integer(c_int32_t) :: has_dir(3)
      integer(c_int32_t) :: tile_x(2)
      integer(c_int32_t) :: tile_yz(2)
      integer(c_int32_t) :: fast_math
      integer(c_int32_t) :: reserved
   end type pk_config

   type, bind(C) :: pk_stage_report
      integer(c_int32_t) :: dn_negative
      integer(c_int32_t) :: ei_negative
      real(c_double)     :: magic_mass(2)   ! per fluid
   end type pk_stage_report

   type, bind(C) :: pk_profile_stats
      integer(c_int64_t) :: launches(4)
      real(c_double)     :: ms(4)
      real(c_double)     :: algo_bytes(4)
   end type pk_profile_stats

   interface

      integer(c_int32_t) function pk_init(cfg) bind(C, name="pk_init")
         import :: c_int32_t, pk_config
         type(pk_config), intent(in) :: cfg
      end function pk_init

      integer(c_int32_t) function pk_finalize() bind(C, name="pk_finalize")
         import :: c_int32_t
      end function pk_finalize

      type(c_ptr) function pk_last_error() bind(C, name="pk_last_error")
         import :: c_ptr
      end function pk_last_error

      ! u(nvar, nxg, nyg, nzg), b(3, ...), psi(...): the reference's arrays, passed as c_loc()
      integer(c_int32_t) function pk_register_block(cg_id, n, nb, dl, u, uh, b, bh, psi, psih) &
           &                                        bind(C, name="pk_register_block")
         import :: c_int32_t, c_int64_t, c_double, c_ptr
         integer(c_int64_t), value      :: cg_id
         integer(c_int32_t), intent(in) :: n(3)
         integer(c_int32_t), value      :: nb
         real(c_double),     intent(in) :: dl(3)
         type(c_ptr),        value      :: u, uh, b, bh, psi, psih
      end function pk_register_block

      integer(c_int32_t) function pk_unregister_block(cg_id) bind(C, name="pk_unregister_block")
         import :: c_int32_t, c_int64_t
         integer(c_int64_t), value :: cg_id
      end function pk_unregister_block

      integer(c_int32_t) function pk_upload(cg_id) bind(C, name="pk_upload")
         import :: c_int32_t, c_int64_t
         integer(c_int64_t), value :: cg_id
      end function pk_upload

      integer(c_int32_t) function pk_download(cg_id) bind(C, name="pk_download")
         import :: c_int32_t, c_int64_t
         integer(c_int64_t), value :: cg_id
      end function pk_download

      integer(c_int32_t) function pk_upload_from(cg_id, u, b, psi) bind(C, name="pk_upload_from")
         import :: c_int32_t, c_int64_t, c_ptr
         integer(c_int64_t), value :: cg_id
         type(c_ptr),        value :: u, b, psi
      end function pk_upload_from

      integer(c_int32_t) function pk_download_to(cg_id, u, b, psi) bind(C, name="pk_download_to")
         import :: c_int32_t, c_int64_t, c_ptr
         integer(c_int64_t), value :: cg_id
         type(c_ptr),        value :: u, b, psi
      end function pk_download_to

      ! solve_cg_sub(cg, ddim, istep): ddim here is 0-based (xdim - 1 .. zdim - 1)
      integer(c_int32_t) function pk_solve_cg(cg_id, ddim, istep, dt, chspeed) bind(C, name="pk_solve_cg")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value :: cg_id
         integer(c_int32_t), value :: ddim, istep
         real(c_double),     value :: dt, chspeed
      end function pk_solve_cg

      integer(c_int32_t) function pk_sweep(cg_id, ddim, dt, chspeed, psi_fac) bind(C, name="pk_sweep")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value :: cg_id
         integer(c_int32_t), value :: ddim
         real(c_double),     value :: dt, chspeed, psi_fac
      end function pk_sweep

      integer(c_int32_t) function pk_fill_ghosts(cg_id) bind(C, name="pk_fill_ghosts")
         import :: c_int32_t, c_int64_t
         integer(c_int64_t), value :: cg_id
      end function pk_fill_ghosts

      integer(c_int32_t) function pk_set_cfl_control(mode, cfl_max, dt_shrink, max_attempts, disallow_negatives) &
           &                                         bind(C, name="pk_set_cfl_control")
         import :: c_int32_t, c_double
         integer(c_int32_t), value :: mode, max_attempts, disallow_negatives
         real(c_double),     value :: cfl_max, dt_shrink
      end function pk_set_cfl_control

      integer(c_int32_t) function pk_redone_steps(count) bind(C, name="pk_redone_steps")
         import :: c_int32_t
         integer(c_int32_t), intent(out) :: count
      end function pk_redone_steps

      integer(c_int32_t) function pk_unsplit_stage(cg_id, istep, dt, chspeed) bind(C, name="pk_unsplit_stage")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value :: cg_id
         integer(c_int32_t), value :: istep
         real(c_double),     value :: dt, chspeed
      end function pk_unsplit_stage

      integer(c_int32_t) function pk_fluid_update_unsplit(cg_id, dt) bind(C, name="pk_fluid_update_unsplit")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value :: cg_id
         real(c_double),     value :: dt
      end function pk_fluid_update_unsplit

      integer(c_int32_t) function pk_run_unsplit(cg_id, tend, nmax, t_out, nsteps_out) bind(C, name="pk_run_unsplit")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value :: cg_id
         real(c_double),     value :: tend
         integer(c_int32_t), value :: nmax
         real(c_double),     intent(out) :: t_out
         integer(c_int32_t), intent(out) :: nsteps_out
      end function pk_run_unsplit

      ! name: NUL-terminated short name (hdf_vars); out: n_b(1)*n_b(2)*n_b(3) values
      integer(c_int32_t) function pk_set_time_control(dt_initial, dt_max_grow, dt_min, dt_max, tend) &
           &                                        bind(C, name="pk_set_time_control")
         import :: c_int32_t, c_double
         real(c_double), value :: dt_initial, dt_max_grow, dt_min, dt_max, tend
      end function pk_set_time_control

      integer(c_int32_t) function pk_datafield(cg_id, name, out) bind(C, name="pk_datafield")
         import :: c_int64_t, c_int32_t, c_char, c_double
         integer(c_int64_t), value                         :: cg_id
         character(kind=c_char), dimension(*), intent(in)  :: name
         real(c_double),         dimension(*), intent(out) :: out
      end function pk_datafield

      integer(c_int32_t) function pk_datafield_f32(cg_id, name, out) bind(C, name="pk_datafield_f32")
         import :: c_int64_t, c_int32_t, c_char, c_float
         integer(c_int64_t), value                         :: cg_id
         character(kind=c_char), dimension(*), intent(in)  :: name
         real(c_float),          dimension(*), intent(out) :: out
      end function pk_datafield_f32

      integer(c_int32_t) function pk_register_fcflx(cg_id, ddim, side, kind, first, index, uflx, bflx) &
           &                                       bind(C, name="pk_register_fcflx")
         import :: c_int32_t, c_int64_t, c_ptr
         integer(c_int64_t), value :: cg_id, first
         integer(c_int32_t), value :: ddim, side, kind
         type(c_ptr),        value :: index, uflx, bflx
      end function pk_register_fcflx

      integer(c_int32_t) function pk_set_neighbour(cg_id, dir, side, nb_id) bind(C, name="pk_set_neighbour")
         import :: c_int32_t, c_int64_t
         integer(c_int64_t), value :: cg_id, nb_id
         integer(c_int32_t), value :: dir, side
      end function pk_set_neighbour

      integer(c_int32_t) function pk_glm_damp(cg_id, fac) bind(C, name="pk_glm_damp")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value :: cg_id
         real(c_double),     value :: fac
      end function pk_glm_damp

      integer(c_int32_t) function pk_timestep_fluid(cg_id, dt_out, c_out) bind(C, name="pk_timestep_fluid")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value       :: cg_id
         real(c_double),     intent(out) :: dt_out, c_out
      end function pk_timestep_fluid

      integer(c_int32_t) function pk_fluid_update(cg_id, dt) bind(C, name="pk_fluid_update")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value :: cg_id
         real(c_double),     value :: dt
      end function pk_fluid_update

      integer(c_int32_t) function pk_run(cg_id, nsteps, t_out, dt_out) bind(C, name="pk_run")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value       :: cg_id
         integer(c_int32_t), value       :: nsteps
         real(c_double),     intent(out) :: t_out, dt_out
      end function pk_run

      integer(c_int32_t) function pk_get_report(cg_id, rep, reset) bind(C, name="pk_get_report")
         import :: c_int32_t, c_int64_t, pk_stage_report
         integer(c_int64_t),    value       :: cg_id
         type(pk_stage_report), intent(out) :: rep
         integer(c_int32_t),    value       :: reset
      end function pk_get_report

      integer(c_int32_t) function pk_synchronize() bind(C, name="pk_synchronize")
         import :: c_int32_t
      end function pk_synchronize

      integer(c_int32_t) function pk_autotune(cg_id, dt, chspeed, tiles_out) bind(C, name="pk_autotune")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value       :: cg_id
         real(c_double),     value       :: dt, chspeed
         integer(c_int32_t), intent(out) :: tiles_out(6)
      end function pk_autotune

      ! device arrays (e.g. from a GPU allocation), reference layout
      integer(c_int32_t) function pk_upload_device(cg_id, u, b, psi) bind(C, name="pk_upload_device")
         import :: c_int32_t, c_int64_t, c_ptr
         integer(c_int64_t), value :: cg_id
         type(c_ptr),        value :: u, b, psi
      end function pk_upload_device

      integer(c_int32_t) function pk_download_device(cg_id, u, b, psi) bind(C, name="pk_download_device")
         import :: c_int32_t, c_int64_t, c_ptr
         integer(c_int64_t), value :: cg_id
         type(c_ptr),        value :: u, b, psi
      end function pk_download_device

      ! the whole registered block list (the block loops of sweeps.F90:230-260)
      integer(c_int32_t) function pk_timestep_all(dt_out, c_out) bind(C, name="pk_timestep_all")
         import :: c_int32_t, c_double
         real(c_double), intent(out) :: dt_out, c_out
      end function pk_timestep_all

      integer(c_int32_t) function pk_fluid_update_all(mode, dt) bind(C, name="pk_fluid_update_all")
         import :: c_int32_t, c_double
         integer(c_int32_t), value :: mode
         real(c_double),     value :: dt
      end function pk_fluid_update_all

      integer(c_int32_t) function pk_run_all(mode, nsteps, t_out, dt_out) bind(C, name="pk_run_all")
         import :: c_int32_t, c_double
         integer(c_int32_t), value       :: mode, nsteps
         real(c_double),     intent(out) :: t_out, dt_out
      end function pk_run_all

      integer(c_int32_t) function pk_autotune_all(dt, chspeed) bind(C, name="pk_autotune_all")
         import :: c_int32_t, c_double
         real(c_double), value :: dt, chspeed
      end function pk_autotune_all

      integer(c_int32_t) function pk_list_tiles(ddim, ifluid, cg_id, tiles) bind(C, name="pk_list_tiles")
         import :: c_int32_t, c_int64_t
         integer(c_int32_t), value       :: ddim, ifluid
         integer(c_int64_t), value       :: cg_id
         integer(c_int32_t), intent(out) :: tiles(3)
      end function pk_list_tiles

      integer(c_int32_t) function pk_set_list_tiles(ddim, ifluid, lanes, ts) bind(C, name="pk_set_list_tiles")
         import :: c_int32_t
         integer(c_int32_t), value :: ddim, ifluid, lanes, ts
      end function pk_set_list_tiles

      integer(c_int32_t) function pk_set_list_split(ddim, ifluid, shapes, nshapes, ts) bind(C, name="pk_set_list_split")
         import :: c_int32_t
         integer(c_int32_t), value      :: ddim, ifluid, nshapes, ts
         integer(c_int32_t), intent(in) :: shapes(3, *)
      end function pk_set_list_split

      ! reverse mode of the fused sweep (one fluid, periodic blocks)
      integer(c_int32_t) function pk_adjoint_setup(cg_id, nslots) bind(C, name="pk_adjoint_setup")
         import :: c_int32_t, c_int64_t
         integer(c_int64_t), value :: cg_id
         integer(c_int32_t), value :: nslots
      end function pk_adjoint_setup

      integer(c_int32_t) function pk_state_save(cg_id, slot) bind(C, name="pk_state_save")
         import :: c_int32_t, c_int64_t
         integer(c_int64_t), value :: cg_id
         integer(c_int32_t), value :: slot
      end function pk_state_save

      integer(c_int32_t) function pk_state_load(cg_id, slot) bind(C, name="pk_state_load")
         import :: c_int32_t, c_int64_t
         integer(c_int64_t), value :: cg_id
         integer(c_int32_t), value :: slot
      end function pk_state_load

      integer(c_int32_t) function pk_adjoint_set(cg_id, lam_u, lam_b, lam_psi) bind(C, name="pk_adjoint_set")
         import :: c_int32_t, c_int64_t, c_ptr
         integer(c_int64_t), value :: cg_id
         type(c_ptr),        value :: lam_u, lam_b, lam_psi
      end function pk_adjoint_set

      integer(c_int32_t) function pk_adjoint_get(cg_id, lam_u, lam_b, lam_psi) bind(C, name="pk_adjoint_get")
         import :: c_int32_t, c_int64_t, c_ptr
         integer(c_int64_t), value :: cg_id
         type(c_ptr),        value :: lam_u, lam_b, lam_psi
      end function pk_adjoint_get

      integer(c_int32_t) function pk_fluid_update_record(cg_id, slot0, dt) bind(C, name="pk_fluid_update_record")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value :: cg_id
         integer(c_int32_t), value :: slot0
         real(c_double),     value :: dt
      end function pk_fluid_update_record

      integer(c_int32_t) function pk_sweep_adjoint(cg_id, slot, ddim, dt, chspeed, psi_fac, dt_bar, ch_bar) &
           &                                       bind(C, name="pk_sweep_adjoint")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value         :: cg_id
         integer(c_int32_t), value         :: slot, ddim
         real(c_double),     value         :: dt, chspeed, psi_fac
         real(c_double),     intent(inout) :: dt_bar, ch_bar
      end function pk_sweep_adjoint

      integer(c_int32_t) function pk_fluid_update_adjoint(cg_id, slot0, dt, dt_bar) &
           &                                              bind(C, name="pk_fluid_update_adjoint")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value         :: cg_id
         integer(c_int32_t), value         :: slot0
         real(c_double),     value         :: dt
         real(c_double),     intent(inout) :: dt_bar
      end function pk_fluid_update_adjoint

      integer(c_int32_t) function pk_timestep_adjoint(cg_id, slot, dt_bar) bind(C, name="pk_timestep_adjoint")
         import :: c_int32_t, c_int64_t, c_double
         integer(c_int64_t), value :: cg_id
         integer(c_int32_t), value :: slot
         real(c_double),     value :: dt_bar
      end function pk_timestep_adjoint

      ! ensembles of independent instances (arrays: one entry per registered block)
      integer(c_int32_t) function pk_ensemble_ids(ids, n) bind(C, name="pk_ensemble_ids")
         import :: c_int32_t, c_int64_t
         integer(c_int64_t), intent(out) :: ids(*)
         integer(c_int32_t), value       :: n
      end function pk_ensemble_ids

      integer(c_int32_t) function pk_ensemble_step(mode, first, dt_out, capped_out) bind(C, name="pk_ensemble_step")
         import :: c_int32_t, c_double
         integer(c_int32_t), value       :: mode, first
         real(c_double),     intent(out) :: dt_out(*)
         integer(c_int32_t), intent(out) :: capped_out(*)
      end function pk_ensemble_step

      integer(c_int32_t) function pk_ensemble_run(mode, nsteps, t_out, dt_out) bind(C, name="pk_ensemble_run")
         import :: c_int32_t, c_double
         integer(c_int32_t), value       :: mode, nsteps
         real(c_double),     intent(out) :: t_out(*), dt_out(*)
      end function pk_ensemble_run

      integer(c_int32_t) function pk_ensemble_state_save(slot) bind(C, name="pk_ensemble_state_save")
         import :: c_int32_t
         integer(c_int32_t), value :: slot
      end function pk_ensemble_state_save

      integer(c_int32_t) function pk_ensemble_state_load(slot) bind(C, name="pk_ensemble_state_load")
         import :: c_int32_t
         integer(c_int32_t), value :: slot
      end function pk_ensemble_state_load

      integer(c_int32_t) function pk_ensemble_record_step(dts, slot0) bind(C, name="pk_ensemble_record_step")
         import :: c_int32_t, c_double
         real(c_double),     intent(in) :: dts(*)
         integer(c_int32_t), value      :: slot0
      end function pk_ensemble_record_step

      integer(c_int32_t) function pk_ensemble_step_adjoint(dts, slot0, dt_bar) bind(C, name="pk_ensemble_step_adjoint")
         import :: c_int32_t, c_double
         real(c_double),     intent(in)    :: dts(*)
         integer(c_int32_t), value         :: slot0
         real(c_double),     intent(inout) :: dt_bar(*)
      end function pk_ensemble_step_adjoint

      integer(c_int32_t) function pk_ensemble_timestep_adjoint(slot, dt_bar) &
           &                                                   bind(C, name="pk_ensemble_timestep_adjoint")
         import :: c_int32_t, c_double
         integer(c_int32_t), value      :: slot
         real(c_double),     intent(in) :: dt_bar(*)
      end function pk_ensemble_timestep_adjoint

      integer(c_int32_t) function pk_profile_begin() bind(C, name="pk_profile_begin")
         import :: c_int32_t
      end function pk_profile_begin

      integer(c_int32_t) function pk_profile_end(st) bind(C, name="pk_profile_end")
         import :: c_int32_t, pk_profile_stats
         type(pk_profile_stats), intent(out) :: st
      end function pk_profile_end

   end interface

contains

   !> pk_last_error() as a Fortran string, for die()
   function pk_error_message() result(msg)

      implicit none

      character(len=:), allocatable :: msg
      character(kind=c_char), dimension(:), pointer :: p
      type(c_ptr) :: cp
      integer :: n

      cp = pk_last_error()
      if (.not. c_associated(cp)) then
         msg = ""
         return
      endif
      call c_f_pointer(cp, p, [1024])
      n = 0
      do while (n < 1024)
         if (p(n + 1) == c_null_char) exit
         n = n + 1
      enddo
      allocate(character(len=n) :: msg)
      msg = transfer(p(1:n), msg)

   end function pk_error_message

end module piernik_hip
