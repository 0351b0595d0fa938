! Fortran binding of the C-ABI in include/echo_amd.h (iso_c_binding).
!
! A Fortran caller such as the reference's echo.f90 uses this module in place
! of the per-stage calls into evolve.f90: echo_create replaces the set-up of
! echo.f90:73-109, echo_upload_state hands over the initial v/u arrays
! (common.f90:203-214, same column-major layout v(ix,iy,iz,k)), echo_steps
! replaces the body of the main loop echo.f90:185-241 and
! echo_download_state refreshes v/u on the output cadence (echo.f90:243-260).
module echo_amd_f
  use iso_c_binding
  implicit none
  private
  public :: echo_config, echo_create, echo_destroy, echo_nv, echo_init_from_prim, &
            echo_upload_state, echo_download_state, echo_steps, echo_rk_stage, echo_end_step, &
            echo_get_scalars, echo_set_tout, echo_cons2prim_batch, echo_set_timing, &
            echo_get_timing, echo_last_error, echo_error_message, echo_set_stop, echo_get_stop, &
            echo_write_frame, echo_wait_frames, echo_fo_setup, echo_fo_slice, echo_fo_width, &
            echo_fo_elements, echo_fo_write, echo_set_dump, echo_dump, echo_restart, &
            echo_glauber_params, echo_bfield_params, echo_nucleons, echo_init_glauber, &
            echo_init_glauber_mhd, echo_glauber_sampler, echo_glauber_event, echo_glauber_generate, &
            echo_glauber_read_event, echo_set_bulk_peak

  ! struct echo_config: 16 int32 then 19 double (param.dat keys in brackets)
  type, bind(C) :: echo_config
    integer(c_int32_t) :: nx, ny, nz            ! NX, NY, NZ
    integer(c_int32_t) :: coordinates           ! COORD
    integer(c_int32_t) :: viscosity             ! VISCOUS
    integer(c_int32_t) :: bulk                  ! BULK
    integer(c_int32_t) :: mhd                   ! MHD
    integer(c_int32_t) :: divclean              ! DIVCLEAN
    integer(c_int32_t) :: eqstate               ! EOS
    integer(c_int32_t) :: recal                 ! REC_ALGO (1 TVD2 .. 7 MPE5)
    integer(c_int32_t) :: nrk                   ! INT_ALGO (2 RK2, 3 RK3)
    integer(c_int32_t) :: boundary_cond         ! bound_con
    integer(c_int32_t) :: algo_solver           ! ALG_SOLV
    integer(c_int32_t) :: obtained_zz           ! TRACE_IMP (1 'zz', 0 'no')
    integer(c_int32_t) :: init_type             ! INIT_TYPE
    integer(c_int32_t) :: derivatives_out       ! derivativ
    real(c_double) :: x1, x2, y1, y2, z1, z2    ! XMIN .. ZMAX
    real(c_double) :: tmin                      ! TSTART
    real(c_double) :: cfl                       ! CFL
    real(c_double) :: maxdt                     ! MAXDT
    real(c_double) :: maxspeed                  ! MAXSPEED
    real(c_double) :: eta_over_s                ! ETA_S
    real(c_double) :: tau_pi_coeff              ! TAU_PI_C
    real(c_double) :: enezero                   ! ENEZERO
    real(c_double) :: glm_alpha                 ! GLM_PARAM
    real(c_double) :: sigma_el                  ! EL_COND
    real(c_double) :: sigma_chi                 ! CHIR_COND
    real(c_double) :: dtout                     ! DTOUT
    real(c_double) :: smooth_temp               ! CUT_TEMP
    real(c_double) :: tmax                      ! TSTOP
  end type echo_config

  ! struct echo_glauber_params: parameters of generate_energy_density_profile
  type, bind(C) :: echo_glauber_params
    real(c_double) :: kappa, ah, sig_mc, ybeam, deta, sigeta, rhocenter
  end type echo_glauber_params

  ! struct echo_glauber_sampler: generate_events (NUCLEUS data, SIGMA_IN in mb,
  ! B, FIXED_B, NCONF, NBCOLL, MIN_PART)
  type, bind(C) :: echo_glauber_sampler
    integer(c_int32_t) :: A, Z, nconf, nbcoll, min_participants, fixed_b
    real(c_double) :: radius, diffuse, rho0, sigma_nn, bimpact, bmax
  end type echo_glauber_sampler

  ! struct echo_glauber_event: one event as read back (caller's arrays)
  type, bind(C) :: echo_glauber_event
    real(c_double) :: b
    integer(c_int32_t) :: n_part_a, n_part_b, n_spec_a, n_spec_b, n_coll
    type(c_ptr) :: part_xy, part_type, spec_xy, spec_type, coll_xy
  end type echo_glauber_event

  ! struct echo_bfield_params: generate_B_field_profile (gamma_col, B_amp_fac, PW, MAGFIELDT)
  type, bind(C) :: echo_bfield_params
    real(c_double) :: gamma_col, b_amplify, pw
    integer(c_int32_t) :: magfield_type
  end type echo_bfield_params

  ! struct echo_nucleons: (x, y) pairs and types (0 proton, 1 neutron) of one list,
  ! e.g. c_loc of xspecA/yspecA packed in pairs and of spec_typeA
  type, bind(C) :: echo_nucleons
    type(c_ptr) :: xy
    type(c_ptr) :: ptype
    integer(c_int32_t) :: n
  end type echo_nucleons

  interface
    integer(c_int) function echo_create(cfg, eos_rows, n_rows, device, grid) bind(C, name="echo_create")
      import :: c_int, c_int32_t, c_double, c_ptr, echo_config
      type(echo_config), intent(in) :: cfg
      real(c_double), intent(in) :: eos_rows(*)
      integer(c_int32_t), value :: n_rows, device
      type(c_ptr), intent(out) :: grid
    end function
    subroutine echo_destroy(grid) bind(C, name="echo_destroy")
      import :: c_ptr
      type(c_ptr), value :: grid
    end subroutine
    integer(c_int) function echo_nv(grid) bind(C, name="echo_nv")
      import :: c_int, c_ptr
      type(c_ptr), value :: grid
    end function
    integer(c_int) function echo_init_from_prim(grid, v) bind(C, name="echo_init_from_prim")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: grid
      real(c_double), intent(in) :: v(*)
    end function
    integer(c_int) function echo_upload_state(grid, v, u) bind(C, name="echo_upload_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: grid
      real(c_double), intent(in) :: v(*), u(*)
    end function
    integer(c_int) function echo_download_state(grid, v, u) bind(C, name="echo_download_state")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: grid
      real(c_double), intent(out) :: v(*), u(*)
    end function
    integer(c_int) function echo_steps(grid, nsteps) bind(C, name="echo_steps")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: grid
      integer(c_int32_t), value :: nsteps
    end function
    integer(c_int) function echo_rk_stage(grid, irk) bind(C, name="echo_rk_stage")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: grid
      integer(c_int32_t), value :: irk
    end function
    integer(c_int) function echo_end_step(grid) bind(C, name="echo_end_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: grid
    end function
    integer(c_int) function echo_get_scalars(grid, s) bind(C, name="echo_get_scalars")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: grid
      real(c_double), intent(out) :: s(10)
    end function
    integer(c_int) function echo_set_tout(grid, tout) bind(C, name="echo_set_tout")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: grid
      real(c_double), value :: tout
    end function
    integer(c_int) function echo_cons2prim_batch(grid, u, v, status, n) bind(C, name="echo_cons2prim_batch")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: grid
      real(c_double), intent(in) :: u(*)
      real(c_double), intent(inout) :: v(*)
      integer(c_int32_t), intent(out) :: status(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function echo_set_timing(grid, on) bind(C, name="echo_set_timing")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: grid
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function echo_get_timing(grid, t) bind(C, name="echo_get_timing")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: grid
      real(c_double), intent(out) :: t(4)
    end function
    type(c_ptr) function echo_last_error() bind(C, name="echo_last_error")
      import :: c_ptr
    end function
    ! check_temp stop criterion (out.f90:1487-1557)
    integer(c_int) function echo_set_stop(grid, temp_end, freeze_type) bind(C, name="echo_set_stop")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: grid
      real(c_double), value :: temp_end
      integer(c_int32_t), value :: freeze_type
    end function
    integer(c_int) function echo_get_stop(grid, stopped, max_value) bind(C, name="echo_get_stop")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: grid
      integer(c_int32_t), intent(out) :: stopped
      real(c_double), intent(out) :: max_value
    end function
    ! out_vars frames (out.f90:232-591); dir is a C string (trim(d)//c_null_char)
    integer(c_int) function echo_write_frame(grid, sel, precision, iout, dir) bind(C, name="echo_write_frame")
      import :: c_int, c_int32_t, c_ptr, c_char
      type(c_ptr), value :: grid
      integer(c_int32_t), intent(in) :: sel(28)
      integer(c_int32_t), value :: precision, iout
      character(kind=c_char), intent(in) :: dir(*)
    end function
    integer(c_int) function echo_wait_frames(grid) bind(C, name="echo_wait_frames")
      import :: c_int, c_ptr
      type(c_ptr), value :: grid
    end function
    ! restart dump: out_dump / out_restart (out.f90:1313-1483); path as a C string
    integer(c_int) function echo_set_dump(grid, on) bind(C, name="echo_set_dump")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: grid
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function echo_dump(grid, path, iout, thyp) bind(C, name="echo_dump")
      import :: c_int, c_int32_t, c_ptr, c_char, c_double
      type(c_ptr), value :: grid
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: iout
      real(c_double), value :: thyp
    end function
    integer(c_int) function echo_restart(grid, path, iout, thyp) bind(C, name="echo_restart")
      import :: c_int, c_int32_t, c_ptr, c_char, c_double
      type(c_ptr), value :: grid
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), intent(out) :: iout
      real(c_double), intent(out) :: thyp
    end function
    ! freeze-out hypersurface: get_hysufile_ready + out_hyp (out.f90:1568-1645)
    integer(c_int) function echo_fo_setup(grid, freeze_type, freeze_value, hdr) bind(C, name="echo_fo_setup")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: grid
      integer(c_int32_t), value :: freeze_type
      real(c_double), value :: freeze_value
      real(c_double), intent(out) :: hdr(2)
    end function
    integer(c_int) function echo_fo_slice(grid, n_elements) bind(C, name="echo_fo_slice")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: grid
      integer(c_int64_t), intent(out) :: n_elements
    end function
    integer(c_int) function echo_fo_width(grid) bind(C, name="echo_fo_width")
      import :: c_int, c_ptr
      type(c_ptr), value :: grid
    end function
    integer(c_int) function echo_fo_elements(grid, out, n) bind(C, name="echo_fo_elements")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value :: grid
      real(c_double), intent(out) :: out(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function echo_fo_write(grid, path, header) bind(C, name="echo_fo_write")
      import :: c_int, c_int32_t, c_ptr, c_char
      type(c_ptr), value :: grid
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: header
    end function
    integer(c_int) function echo_glauber_generate(smp, seed, dir, n_events) bind(C, name="echo_glauber_generate")
      import :: c_int, c_int32_t, c_int64_t, c_char, echo_glauber_sampler
      type(echo_glauber_sampler), intent(in) :: smp
      integer(c_int64_t), value :: seed
      character(kind=c_char), intent(in) :: dir(*)
      integer(c_int32_t), intent(out) :: n_events
    end function
    integer(c_int) function echo_glauber_read_event(dir, iev, ev, cap_part, cap_spec, cap_coll) &
        bind(C, name="echo_glauber_read_event")
      import :: c_int, c_int32_t, c_char, echo_glauber_event
      character(kind=c_char), intent(in) :: dir(*)
      integer(c_int32_t), value :: iev, cap_part, cap_spec, cap_coll
      type(echo_glauber_event), intent(inout) :: ev
    end function
    integer(c_int) function echo_set_bulk_peak(grid, zs_max, t_c, width) bind(C, name="echo_set_bulk_peak")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: grid
      real(c_double), value :: zs_max, t_c, width
    end function
    integer(c_int) function echo_init_glauber(grid, p, part_a, n_a, part_b, n_b, coll, n_coll) &
        bind(C, name="echo_init_glauber")
      import :: c_int, c_int32_t, c_double, c_ptr, echo_glauber_params
      type(c_ptr), value :: grid
      type(echo_glauber_params), intent(in) :: p
      real(c_double), intent(in) :: part_a(*), part_b(*), coll(*)
      integer(c_int32_t), value :: n_a, n_b, n_coll
    end function
    integer(c_int) function echo_init_glauber_mhd(grid, p, part_a, n_a, part_b, n_b, coll, n_coll, &
        bp, lists) bind(C, name="echo_init_glauber_mhd")
      import :: c_int, c_int32_t, c_double, c_ptr, echo_glauber_params, echo_bfield_params, echo_nucleons
      type(c_ptr), value :: grid
      type(echo_glauber_params), intent(in) :: p
      real(c_double), intent(in) :: part_a(*), part_b(*), coll(*)
      integer(c_int32_t), value :: n_a, n_b, n_coll
      type(echo_bfield_params), intent(in) :: bp
      type(echo_nucleons), intent(in) :: lists(4)
    end function
  end interface

contains

  ! the message of the last failed call, as a Fortran string
  function echo_error_message() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ' '
    call c_f_pointer(echo_last_error(), s, [512])
    do i = 1, 512
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function echo_error_message

end module echo_amd_f
