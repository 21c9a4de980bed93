! Fortran interfaces of libgcomcot_hip (include/gcomcot_hip.h).
!
! The reference's Fortran calls Part 1 of the library (mass_launch_,
! momt_launch_, all_grid_launch_, ...) through implicit interfaces with
! every argument by reference; it links against libgcomcot_hip.so without
! this module.  New Fortran code that wants checked interfaces, or the
! extensions of Parts 2 and 3 (nonlinear and Cartesian layers, the
! whole-run device loop, wet maxima and arrival, the scenario batch
! engine), uses this module.  Arrays are column-major (NX,NY) REAL planes;
! integers are default INTEGER (C int); step counts of the batch engine
! are INTEGER(C_LONG_LONG).
module gcomcot_hip
  use, intrinsic :: iso_c_binding
  implicit none

  interface
    ! ---------------------------------------------------------- Part 1
    subroutine gcomcot_init_gpu() bind(C, name="gcomcot_init_gpu_")
    end subroutine

    subroutine gcomcot_init_layer(id, parent, level, rel_size, rel_time, corners, grx, gry, &
                                  r1, r2, r3, r4, r5, r6, r11, h, z, nx, ny) &
        bind(C, name="gcomcot_init_layer_")
      import :: c_int, c_float
      integer(c_int) :: id, parent, level, rel_size, rel_time, corners(4), nx, ny
      real(c_float) :: grx, gry
      real(c_float) :: r1(*), r2(*), r3(*), r4(*), r5(*), r6(*), r11(*), h(*), z(*)
    end subroutine

    subroutine gcomcot_show_grids() bind(C, name="gcomcot_show_grids_")
    end subroutine

    subroutine mass_launch(z1, z2, h, lid) bind(C, name="mass_launch_")
      import :: c_int, c_float
      real(c_float) :: z1(*), z2(*), h(*)
      integer(c_int) :: lid
    end subroutine

    subroutine openbd_launch(z2) bind(C, name="openbd_launch_")
      import :: c_float
      real(c_float) :: z2(*)
    end subroutine

    subroutine momt_launch(m2, n2, z2, lid) bind(C, name="momt_launch_")
      import :: c_int, c_float
      real(c_float) :: m2(*), n2(*), z2(*)
      integer(c_int) :: lid
    end subroutine

    subroutine all_grid_launch() bind(C, name="all_grid_launch_")
    end subroutine

    subroutine maxamp_launch(lid) bind(C, name="maxamp_launch_")
      import :: c_int
      integer(c_int) :: lid
    end subroutine

    subroutine cuda_update_layer(lid) bind(C, name="cuda_update_layer_")
      import :: c_int
      integer(c_int) :: lid
    end subroutine

    subroutine cuda_getz(z, lid) bind(C, name="cuda_getz_")
      import :: c_int, c_float
      real(c_float) :: z(*)
      integer(c_int) :: lid
    end subroutine
    subroutine cuda_setz(z, lid) bind(C, name="cuda_setz_")
      import :: c_int, c_float
      real(c_float) :: z(*)
      integer(c_int) :: lid
    end subroutine
    subroutine cuda_getz1(z, lid) bind(C, name="cuda_getz1_")
      import :: c_int, c_float
      real(c_float) :: z(*)
      integer(c_int) :: lid
    end subroutine
    subroutine cuda_setz1(z, lid) bind(C, name="cuda_setz1_")
      import :: c_int, c_float
      real(c_float) :: z(*)
      integer(c_int) :: lid
    end subroutine
    subroutine cuda_getzmax(z, lid) bind(C, name="cuda_getzmax_")
      import :: c_int, c_float
      real(c_float) :: z(*)
      integer(c_int) :: lid
    end subroutine

    subroutine cuda_getmn(m, n, lid) bind(C, name="cuda_getmn_")
      import :: c_int, c_float
      real(c_float) :: m(*), n(*)
      integer(c_int) :: lid
    end subroutine
    subroutine cuda_setmn(m, n, lid) bind(C, name="cuda_setmn_")
      import :: c_int, c_float
      real(c_float) :: m(*), n(*)
      integer(c_int) :: lid
    end subroutine
    subroutine cuda_getmn1(m, n, lid) bind(C, name="cuda_getmn1_")
      import :: c_int, c_float
      real(c_float) :: m(*), n(*)
      integer(c_int) :: lid
    end subroutine
    subroutine cuda_setmn1(m, n, lid) bind(C, name="cuda_setmn1_")
      import :: c_int, c_float
      real(c_float) :: m(*), n(*)
      integer(c_int) :: lid
    end subroutine

    subroutine gcomcot_get_z(lid, l1, l2, l3, l4, i, j) bind(C, name="gcomcot_get_z_")
      import :: c_int, c_float
      integer(c_int) :: lid, i, j
      real(c_float) :: l1, l2, l3, l4
    end subroutine

    subroutine cuda_shutdown() bind(C, name="cuda_shutdown_")
    end subroutine

    ! ---------------------------------------------------------- Part 2
    subroutine gcomcot_para_rows(laygov, ny, dt, dx, y, dely, coef) &
        bind(C, name="gcomcot_para_rows_")
      import :: c_int, c_float
      integer(c_int) :: laygov, ny
      real(c_float) :: dt, dx, y(*), dely(*), coef(*)
    end subroutine

    integer(c_int) function gcomcot_init_layer_ex(id, parent, level, laycord, laygov, &
        rel_size, rel_time, corners, fric_switch, fric_coef, dt, dx, dy, coef, h, z, nx, ny) &
        bind(C, name="gcomcot_init_layer_ex_")
      import :: c_int, c_float
      integer(c_int) :: id, parent, level, laycord, laygov, rel_size, rel_time, corners(4)
      integer(c_int) :: fric_switch, nx, ny
      real(c_float) :: fric_coef, dt, dx, dy, coef(*), h(*), z(*)
    end function

    ! a Cartesian child of the spherical layer 1 with SC_OPTION = 1: POS, CXY
    ! and CORNERS as SUBGRID_MATCHING leaves them
    integer(c_int) function gcomcot_init_layer_sc(id, parent, level, laygov, rel_size, &
        rel_time, corners, fric_switch, fric_coef, dt, dx, dy, h, z, nx, ny, pos, cxy) &
        bind(C, name="gcomcot_init_layer_sc_")
      import :: c_int, c_float
      integer(c_int) :: id, parent, level, laygov, rel_size, rel_time, corners(4)
      integer(c_int) :: fric_switch, nx, ny, pos(*)
      real(c_float) :: fric_coef, dt, dx, dy, h(*), z(*), cxy(*)
    end function

    integer(c_int) function gcomcot_run_steps(nsteps, use_graph) &
        bind(C, name="gcomcot_run_steps_")
      import :: c_int
      integer(c_int) :: nsteps, use_graph
    end function

    subroutine gcomcot_get_max(lid, zmax_wet, arrival) bind(C, name="gcomcot_get_max_")
      import :: c_int, c_float
      integer(c_int) :: lid, arrival(*)
      real(c_float) :: zmax_wet(*)
    end subroutine

    subroutine gcomcot_set_arrival_threshold(thr) &
        bind(C, name="gcomcot_set_arrival_threshold_")
      import :: c_float
      real(c_float) :: thr
    end subroutine

    subroutine gcomcot_set_error_mode(fatal) bind(C, name="gcomcot_set_error_mode_")
      import :: c_int
      integer(c_int) :: fatal
    end subroutine

    integer(c_int) function gcomcot_last_error(buf, length) bind(C, name="gcomcot_last_error_")
      import :: c_int, c_char
      character(kind=c_char) :: buf(*)
      integer(c_int) :: length
    end function

    integer(c_int) function gcomcot_time_steps(nsteps, ms_total, ms_each) &
        bind(C, name="gcomcot_time_steps_")
      import :: c_int, c_float
      integer(c_int) :: nsteps
      real(c_float) :: ms_total, ms_each
    end function

    integer(c_int) function gcomcot_autotune(reps, variant, blocking, ms_per_step) &
        bind(C, name="gcomcot_autotune_")
      import :: c_int, c_float
      integer(c_int) :: reps, variant, blocking
      real(c_float) :: ms_per_step
    end function

    integer(c_int) function gcomcot_set_temporal_blocking(mode) &
        bind(C, name="gcomcot_set_temporal_blocking_")
      import :: c_int
      integer(c_int) :: mode
    end function

    integer(c_int) function gcomcot_set_fused_halves(on) &
        bind(C, name="gcomcot_set_fused_halves_")
      import :: c_int
      integer(c_int) :: on
    end function

    integer(c_int) function gcomcot_set_store_policy(mode) &
        bind(C, name="gcomcot_set_store_policy_")
      import :: c_int
      integer(c_int) :: mode
    end function

    ! Exact checkpoints of the device state
    integer(c_int) function gcomcot_checkpoint_save(path, pathlen) &
        bind(C, name="gcomcot_checkpoint_save_")
      import :: c_int, c_char
      character(kind=c_char) :: path(*)
      integer(c_int) :: pathlen
    end function

    integer(c_int) function gcomcot_checkpoint_load(path, pathlen) &
        bind(C, name="gcomcot_checkpoint_load_")
      import :: c_int, c_char
      character(kind=c_char) :: path(*)
      integer(c_int) :: pathlen
    end function

    ! The reference's hot-start files (HOT_START START_TYPE 0 / 1 / 2)
    integer(c_int) function gcomcot_hotstart_write(start_step, dir, dirlen) &
        bind(C, name="gcomcot_hotstart_write_")
      import :: c_int, c_char
      integer(c_int) :: start_step, dirlen
      character(kind=c_char) :: dir(*)
    end function

    integer(c_int) function gcomcot_hotstart_read(start_type, start_step, dir, dirlen, &
                                                  nmissing) &
        bind(C, name="gcomcot_hotstart_read_")
      import :: c_int, c_char
      integer(c_int) :: start_type, start_step, dirlen, nmissing
      character(kind=c_char) :: dir(*)
    end function

    ! Boundary treatments of the outermost layer (SPONGE_COEF/SPONGE_LAYER, BC_WALL)
    subroutine gcomcot_sponge_coef(laycord, nx, ny, h, x, y, cx, cy) &
        bind(C, name="gcomcot_sponge_coef_")
      import :: c_int, c_float
      integer(c_int) :: laycord, nx, ny
      real(c_float) :: h(*), x(*), y(*), cx(*), cy(*)
    end subroutine

    integer(c_int) function gcomcot_sponge_set(on) bind(C, name="gcomcot_sponge_set_")
      import :: c_int
      integer(c_int) :: on
    end function

    subroutine gcomcot_bc_wall(h, nx, ny) bind(C, name="gcomcot_bc_wall_")
      import :: c_int, c_float
      real(c_float) :: h(*)
      integer(c_int) :: nx, ny
    end subroutine

    ! Fault sources on the device (GET_FLOOR_DEFORM's Okada branch)
    integer(c_int) function gcomcot_fault_deform(params, fswitch) &
        bind(C, name="gcomcot_fault_deform_")
      import :: c_int, c_float
      real(c_float) :: params(11)
      integer(c_int) :: fswitch
    end function

    integer(c_int) function gcomcot_fault_schedule(params, fswitch, t0) &
        bind(C, name="gcomcot_fault_schedule_")
      import :: c_int, c_float
      real(c_float) :: params(11), t0
      integer(c_int) :: fswitch
    end function

    integer(c_int) function gcomcot_get_deform(lid, deform) bind(C, name="gcomcot_get_deform_")
      import :: c_int, c_float
      integer(c_int) :: lid
      real(c_float) :: deform(*)
    end function

    integer(c_int) function gcomcot_get_depth(lid, h) bind(C, name="gcomcot_get_depth_")
      import :: c_int, c_float
      integer(c_int) :: lid
      real(c_float) :: h(*)
    end function

    ! Reference-format output files of the device loop (DATA_PRT, MAX_AMP)
    integer(c_int) function gcomcot_output_set(dir, dirlen, iprt, out_opt, tend, tide) &
        bind(C, name="gcomcot_output_set_")
      import :: c_int, c_float, c_char
      character(kind=c_char) :: dir(*)
      integer(c_int) :: dirlen, iprt, out_opt
      real(c_float) :: tend, tide
    end function

    integer(c_int) function gcomcot_output_layer(lid, fluxswitch) &
        bind(C, name="gcomcot_output_layer_")
      import :: c_int
      integer(c_int) :: lid, fluxswitch
    end function

    integer(c_int) function gcomcot_output_flush(nfiles) bind(C, name="gcomcot_output_flush_")
      import :: c_int, c_long_long
      integer(c_long_long) :: nfiles
    end function

    ! Gauges sampled on the device (the GET_TS series, output.f90:1035-1155)
    integer(c_int) function gcomcot_set_layer_coords(lid, x, y) &
        bind(C, name="gcomcot_set_layer_coords_")
      import :: c_int, c_float
      integer(c_int) :: lid
      real(c_float) :: x(*), y(*)
    end function

    integer(c_int) function gcomcot_gauges_set(n, x0, y0, lid, capacity) &
        bind(C, name="gcomcot_gauges_set_")
      import :: c_int, c_float
      integer(c_int) :: n, lid(*), capacity
      real(c_float) :: x0(*), y0(*)
    end function

    integer(c_int) function gcomcot_gauges_read(series, max_steps, nout) &
        bind(C, name="gcomcot_gauges_read_")
      import :: c_int, c_float
      real(c_float) :: series(*)
      integer(c_int) :: max_steps, nout
    end function

    ! ---------------------------------------------------------- Part 3
    integer(c_int) function gcomcot_batch_reset() bind(C, name="gcomcot_batch_reset_")
      import :: c_int
    end function

    integer(c_int) function gcomcot_batch_new_scenario(steps, sid) &
        bind(C, name="gcomcot_batch_new_scenario_")
      import :: c_int, c_long_long
      integer(c_long_long) :: steps
      integer(c_int) :: sid
    end function

    integer(c_int) function gcomcot_batch_add_layer(sid, id, parent, laycord, laygov, &
        rel_size, rel_time, corners, fric_switch, fric_coef, dt, dx, dy, coef, h, z, nx, ny) &
        bind(C, name="gcomcot_batch_add_layer_")
      import :: c_int, c_float
      integer(c_int) :: sid, id, parent, laycord, laygov, rel_size, rel_time, corners(4)
      integer(c_int) :: fric_switch, nx, ny
      real(c_float) :: fric_coef, dt, dx, dy, coef(*), h(*), z(*)
    end function

    integer(c_int) function gcomcot_batch_add_layer_sc(sid, id, parent, laygov, rel_size, &
        rel_time, corners, fric_switch, fric_coef, dt, dx, dy, h, z, nx, ny, pos, cxy) &
        bind(C, name="gcomcot_batch_add_layer_sc_")
      import :: c_int, c_float
      integer(c_int) :: sid, id, parent, laygov, rel_size, rel_time, corners(4)
      integer(c_int) :: fric_switch, nx, ny, pos(*)
      real(c_float) :: fric_coef, dt, dx, dy, h(*), z(*), cxy(*)
    end function

    integer(c_int) function gcomcot_batch_run(nsteps, nfinished) &
        bind(C, name="gcomcot_batch_run_")
      import :: c_int, c_long_long
      integer(c_long_long) :: nsteps
      integer(c_int) :: nfinished
    end function

    integer(c_int) function gcomcot_batch_get_max(sid, lid, zmax_wet, arrival) &
        bind(C, name="gcomcot_batch_get_max_")
      import :: c_int, c_float
      integer(c_int) :: sid, lid, arrival(*)
      real(c_float) :: zmax_wet(*)
    end function

    integer(c_int) function gcomcot_batch_remove(sid) bind(C, name="gcomcot_batch_remove_")
      import :: c_int
      integer(c_int) :: sid
    end function

    integer(c_int) function gcomcot_batch_set_fused_halves(on) &
        bind(C, name="gcomcot_batch_set_fused_halves_")
      import :: c_int
      integer(c_int) :: on
    end function
  end interface
end module gcomcot_hip
