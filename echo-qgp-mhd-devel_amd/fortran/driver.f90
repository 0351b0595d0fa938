! echo_amd_driver: a Fortran caller of the MI355X grid update, reading the
! reference's param.dat format (10-character flags, common.f90:874-1215) and
! running the time loop of echo.f90:185-262 through the C-ABI.
!
! Usage: echo_amd_driver [param_file]
! Supported reference flags: INIT_TYPE (3 Glauber Monte Carlo, 4 ideal
! Gubser flow), COORD, VISCOUS, BULK, MHD, DIVCLEAN, GLM_PARAM, CUT_TEMP, NX,
! NY, NZ, XMIN..ZMAX, TSTART, TSTOP, TEMP_END, DTOUT, OUTP_PREC, MAXDT, CFL,
! REC_ALGO (default MPE5), INT_ALGO (RK2 default, RK3, SSP on MHD grids),
! MAXSPEED, ENEZERO, ETA_S, TAU_PI_C, TRACE_IMP, EOS, EOS_FILE, ALG_SOLV,
! EL_COND, CHIR_COND, bound_con, the output selection of out_vars (density..=
! .. glm......=, common.f90:1110-1190) with outdir...=, and for INIT_TYPE 3
! NUCLEUS (Au, Pb), RADS, SIGMA_IN, B, FIXED_B, NCONF, NBCOLL, MIN_PART,
! EV_START, KAPPA, AH, SIG, DETA, SIGETA, RHOCENTER, and on MHD grids the
! protons' initial field MAGFIELDT, B_amp_fac, PW (common.f90:826-833);
! other reference flags are ignored.  Driver flags: GUB_T0... (T0 of the Gubser initialiser in
! fm^-1; the reference fixes 2.0 at common.f90:1468), SEED.....= (seed of
! the Glauber-MC generator; the reference takes it from random_seed.dat or
! the clock), GLAUB_DIR= (directory of the event files, default .),
! NSTEPS...= (step limit), CHUNK....= (steps per device call), DUMP_FILE=
! (final v then u, FP64 stream), DEVICE...=.
! Output: frames (echo_write_frame) at the start, at every output time and
! when TEMP_END stops the run (echo.f90:152, 243-247, 264-272).
! Freeze-out hypersurface: HYP_COMPU, FREEZKIND, FREEZEVAL, HYPSURFTI as in
! the reference (common.f90:1062-1076), written to HYP_FILE.= (default
! hypersurface.txt) at the cadence of echo.f90:176-177,252-258; with
! HYP_COMPU=1 the time loop advances one step per device call, as the
! reference checks the cadence after every step.
! Restart: RESTART..=1 writes DUMP_DIR..= (default .)/dump.dat at every
! output time (echo.f90:243-248); the driver flag RESUME...= names a
! dump.dat to start from instead of the Gubser initial state
! (out_restart, echo.f90:162-172).
program echo_amd_driver
  use iso_c_binding
  use echo_amd_f
  implicit none
  type(echo_config) :: cfg
  type(c_ptr) :: grid
  character(len=256) :: pfile, eos_file, dump_file, hyp_file, dump_dir, resume_file
  character(len=10) :: flag
  character(len=120) :: val
  integer :: ios, nsteps, chunk, dev, ierr, nv, iter, n, k
  integer(c_int32_t) :: nrows
  real(8) :: t0gub, s(10), ttot, freeze_value, freeze_dt, thyp, hdr(2), tout_prev
  integer :: hyp_compu, freeze_kind, restart_type
  ! Glauber Monte Carlo (INIT_TYPE 3), output frames, stop criterion
  character(len=5) :: nucleus
  character(len=256) :: glauber_dir, outdir
  real(8) :: rads, sigma_in, bimpact, temp_end
  integer :: fixed_b, nconf, nbcoll, min_part, ev_start, outp_prec
  integer(c_int64_t) :: seed
  integer(c_int32_t) :: sel(28), stopped
  type(echo_glauber_params) :: gpar
  type(echo_bfield_params) :: bpar
  real(8) :: maxv
  integer(c_int32_t) :: iout
  integer(c_int64_t) :: nfo
  real(8), allocatable :: v(:,:,:,:), u(:,:,:,:), rows(:,:)
  integer(8) :: c0, c1, crate
  real(8), parameter :: hbar = 0.197326d0

  pfile = 'param.dat'
  if (command_argument_count() >= 1) call get_command_argument(1, pfile)
  ! defaults of common.f90
  cfg = echo_config(0, 0, 0, 2, 0, 0, 0, 0, 1, 7, 2, 2, 5, 1, 4, 0, &
                    -1d0, 1d0, -1d0, 1d0, -1d0, 1d0, 1d0, 0.4d0, 0.1d0, 0.995d0, 0.08d0, &
                    5d0, 0.001d0, 10d0, 0.0058d0, 0.0015d0, 1d6, 0d0, 1d6)
  t0gub = 2.0d0
  nsteps = huge(1)
  chunk = 10
  dev = 0
  eos_file = 'qcdIEOS0.dat'
  dump_file = ' '
  hyp_file = 'hypersurface.txt'
  dump_dir = '.'
  resume_file = ' '
  restart_type = 0
  iout = 0
  hyp_compu = 0
  freeze_kind = 1
  freeze_value = 0.5d0
  freeze_dt = 0.1d0
  ! Glauber-MC and output defaults of common.f90:574-680
  nucleus = 'Au'
  rads = 200d0
  sigma_in = 42d0
  bimpact = 5d0
  fixed_b = 0
  nconf = 100
  nbcoll = 1
  min_part = 18
  ev_start = 1
  gpar = echo_glauber_params(19d0, 0.15d0, 0.8d0, 0d0, 2d0, 1d0, 1d0)
  bpar = echo_bfield_params(0d0, 1d0, 0.5d0, 1_c_int32_t)
  seed = 1
  glauber_dir = '.'
  temp_end = 0d0
  outp_prec = 8
  outdir = '.'
  sel = 0
  open(29, file=trim(pfile), status='old', iostat=ios)
  if (ios /= 0) then
    write(*,*) 'cannot open ', trim(pfile)
    stop 1
  end if
  do
    read(29, '(a10,a120)', iostat=ios) flag, val
    if (ios < 0) exit
    if (flag(1:1) == '!' .or. flag == '' .or. flag(1:1) == '#') cycle
    select case (flag)
    case ('INIT_TYPE='); read(val, *) cfg%init_type
    case ('COORD....='); read(val, *) cfg%coordinates
    case ('VISCOUS..='); read(val, *) cfg%viscosity
    case ('BULK.....='); read(val, *) cfg%bulk
    case ('MHD......='); read(val, *) cfg%mhd
    case ('DIVCLEAN.='); read(val, *) cfg%divclean
    case ('GLM_PARAM='); read(val, *) cfg%glm_alpha
    case ('CUT_TEMP.='); read(val, *) cfg%smooth_temp
    case ('NX.......='); read(val, *) cfg%nx
    case ('NY.......='); read(val, *) cfg%ny
    case ('NZ.......='); read(val, *) cfg%nz
    case ('XMIN.....='); read(val, *) cfg%x1
    case ('XMAX.....='); read(val, *) cfg%x2
    case ('YMIN.....='); read(val, *) cfg%y1
    case ('YMAX.....='); read(val, *) cfg%y2
    case ('ZMIN.....='); read(val, *) cfg%z1
    case ('ZMAX.....='); read(val, *) cfg%z2
    case ('TSTART...='); read(val, *) cfg%tmin
    case ('TSTOP....='); read(val, *) cfg%tmax
    case ('DTOUT....='); read(val, *) cfg%dtout
    case ('MAXDT....='); read(val, *) cfg%maxdt
    case ('CFL......='); read(val, *) cfg%cfl
    case ('MAXSPEED.='); read(val, *) cfg%maxspeed
    case ('ENEZERO..='); read(val, *) cfg%enezero
    case ('ETA_S....='); read(val, *) cfg%eta_over_s
    case ('TAU_PI_C.='); read(val, *) cfg%tau_pi_coeff
    case ('EOS......='); read(val, *) cfg%eqstate
    case ('EOS_FILE.='); eos_file = adjustl(val)
    case ('ALG_SOLV.='); read(val, *) cfg%algo_solver
    case ('EL_COND..='); read(val, *) cfg%sigma_el
    case ('CHIR_COND='); read(val, *) cfg%sigma_chi
    case ('bound_con='); read(val, *) cfg%boundary_cond
    case ('TRACE_IMP=')
      cfg%obtained_zz = merge(1, 0, adjustl(val) == 'zz')
    case ('INT_ALGO.=')
      select case (trim(adjustl(val)))
      case ('RK2'); cfg%nrk = 2
      case ('RK3'); cfg%nrk = 3
      case ('SSP'); cfg%nrk = 4
      case default
        write(*,*) 'INT_ALGO ', trim(val), ' is not supported (RK2, RK3, SSP)'
        stop 1
      end select
    case ('REC_ALGO.=')
      select case (trim(adjustl(val)))
      case ('TVD2'); cfg%recal = 1
      case ('CENO3'); cfg%recal = 2
      case ('WENO3'); cfg%recal = 3
      case ('WENO5'); cfg%recal = 4
      case ('PPM4'); cfg%recal = 5
      case ('MPE3'); cfg%recal = 6
      case ('MPE5'); cfg%recal = 7
      case default
        write(*,*) 'REC_ALGO ', trim(val), ' is not supported'
        stop 1
      end select
    case ('TEMP_END.='); read(val, *) temp_end
    case ('OUTP_PREC='); read(val, *) outp_prec
    case ('outdir...='); outdir = adjustl(val)
    case ('NUCLEUS..='); nucleus = adjustl(val)
    case ('RADS.....='); read(val, *) rads
    case ('SIGMA_IN.='); read(val, *) sigma_in
    case ('B........='); read(val, *) bimpact
    case ('FIXED_B..='); read(val, *) fixed_b
    case ('NCONF....='); read(val, *) nconf
    case ('NBCOLL...='); read(val, *) nbcoll
    case ('MIN_PART.='); read(val, *) min_part
    case ('EV_START.='); read(val, *) ev_start
    case ('KAPPA....='); read(val, *) gpar%kappa
    case ('AH.......='); read(val, *) gpar%ah
    case ('SIG......='); read(val, *) gpar%sig_mc
    case ('DETA.....='); read(val, *) gpar%deta
    case ('SIGETA...='); read(val, *) gpar%sigeta
    case ('RHOCENTER='); read(val, *) gpar%rhocenter
    case ('SEED.....='); read(val, *) seed
    case ('MAGFIELDT='); read(val, *) bpar%magfield_type
    case ('B_amp_fac='); read(val, *) bpar%b_amplify
    case ('PW.......='); read(val, *) bpar%pw
    case ('GLAUB_DIR='); glauber_dir = adjustl(val)
    case ('density..='); read(val, *) sel(1)
    case ('vx.......='); read(val, *) sel(2)
    case ('vy.......='); read(val, *) sel(3)
    case ('vz.......='); read(val, *) sel(4)
    case ('pressure.='); read(val, *) sel(5)
    case ('ene_dens.='); read(val, *) sel(6)
    case ('temper...='); read(val, *) sel(7)
    case ('entr_dens='); read(val, *) sel(8)
    case ('bulk_visc='); read(val, *) sel(9)
    case ('pi^tt....='); read(val, *) sel(10)
    case ('pi^tx....='); read(val, *) sel(11)
    case ('pi^ty....='); read(val, *) sel(12)
    case ('pi^tz....='); read(val, *) sel(13)
    case ('pi^xy....='); read(val, *) sel(14)
    case ('pi^xz....='); read(val, *) sel(15)
    case ('pi^yz....='); read(val, *) sel(16)
    case ('pi^xx....='); read(val, *) sel(17)
    case ('pi^yy....='); read(val, *) sel(18)
    case ('pi^zz....='); read(val, *) sel(19)
    case ('gamma....='); read(val, *) sel(20)
    case ('bx.......='); read(val, *) sel(21)
    case ('by.......='); read(val, *) sel(22)
    case ('bz.......='); read(val, *) sel(23)
    case ('ex.......='); read(val, *) sel(24)
    case ('ey.......='); read(val, *) sel(25)
    case ('ez.......='); read(val, *) sel(26)
    case ('glm......='); read(val, *) sel(27)
    case ('GUB_T0...='); read(val, *) t0gub
    case ('NSTEPS...='); read(val, *) nsteps
    case ('CHUNK....='); read(val, *) chunk
    case ('DEVICE...='); read(val, *) dev
    case ('DUMP_FILE='); dump_file = adjustl(val)
    case ('HYP_COMPU='); read(val, *) hyp_compu
    case ('FREEZKIND='); read(val, *) freeze_kind
    case ('FREEZEVAL='); read(val, *) freeze_value
    case ('HYPSURFTI='); read(val, *) freeze_dt
    case ('HYP_FILE.='); hyp_file = adjustl(val)
    case ('RESTART..='); read(val, *) restart_type
    case ('DUMP_DIR.='); dump_dir = adjustl(val)
    case ('RESUME...='); resume_file = adjustl(val)
    end select
  end do
  close(29)
  if (cfg%init_type /= 4 .and. cfg%init_type /= 3) then
    write(*,*) 'this driver initialises INIT_TYPE=3 (Glauber Monte Carlo) and 4 (ideal Gubser flow); ', &
               'other initial states are handed over with echo_upload_state'
    stop 2
  end if

  nrows = 0
  allocate(rows(4, 1))
  rows = 0
  if (cfg%eqstate == 3) call read_table(trim(eos_file), rows, nrows)
  ierr = echo_create(cfg, rows, nrows, int(dev, c_int32_t), grid)
  if (ierr /= 0) call fail('echo_create', ierr)
  nv = echo_nv(grid)
  allocate(v(cfg%nx, cfg%ny, cfg%nz, nv), u(cfg%nx, cfg%ny, cfg%nz, nv))
  if (restart_type /= 0) then
    ierr = echo_set_dump(grid, 1_c_int32_t)
    if (ierr /= 0) call fail('echo_set_dump', ierr)
    chunk = 1                        ! dumps follow the step that reaches tout
  end if
  if (hyp_compu == 1) then
    ! get_hysufile_ready (hypersurface.f90:1137-1200): a restarted run appends
    ! to the existing file; set up before the restart, which reads the
    ! freeze-out part of the dump
    ierr = echo_fo_setup(grid, int(freeze_kind, c_int32_t), freeze_value, hdr)
    if (ierr /= 0) call fail('echo_fo_setup', ierr)
    if (len_trim(resume_file) == 0) then
      ierr = echo_fo_write(grid, trim(hyp_file)//c_null_char, 1_c_int32_t)
      if (ierr /= 0) call fail('echo_fo_write', ierr)
    end if
    chunk = 1
  end if
  if (len_trim(resume_file) > 0) then
    ! out_restart; tout = t + dtout is set by echo_restart (echo.f90:164-172)
    ierr = echo_restart(grid, trim(resume_file)//c_null_char, iout, thyp)
    if (ierr /= 0) call fail('echo_restart', ierr)
    iout = iout + 1
    write(*,'(a,a)') ' restarted from ', trim(resume_file)
  else
    if (cfg%init_type == 3) then
      call glauber_init()
    else
      call gubser_init(v)
      ierr = echo_init_from_prim(grid, v)
      if (ierr /= 0) call fail('echo_init_from_prim', ierr)
    end if
    ierr = echo_set_tout(grid, cfg%tmin + cfg%dtout)
    ! out_vars(0) of the initial state, then iout = 1 (echo.f90:152-159)
    call frame(0)
    iout = 1
  end if
  if (temp_end > 0) then
    ierr = echo_set_stop(grid, temp_end, int(freeze_kind, c_int32_t))
    if (ierr /= 0) call fail('echo_set_stop', ierr)
  end if
  ierr = echo_get_scalars(grid, s)
  if (s(1) == cfg%tmin) then
    thyp = cfg%tmin
  else
    thyp = s(1) + freeze_dt          ! echo.f90:174-178 after a restart
  end if

  write(*,'(a,3i6,a,i3,a,i2)') ' grid', cfg%nx, cfg%ny, cfg%nz, '  nv', nv, '  RK', cfg%nrk
  iter = 0
  ttot = 0
  ierr = echo_get_scalars(grid, s)
  do while (s(1) < cfg%tmax .and. iter < nsteps)
    n = min(chunk, nsteps - iter)
    tout_prev = s(7)
    call system_clock(c0, crate)
    ierr = echo_steps(grid, int(n, c_int32_t))
    if (ierr /= 0) call fail('echo_steps', ierr)
    call system_clock(c1)
    ttot = ttot + dble(c1 - c0) / dble(crate)
    iter = iter + n
    ierr = echo_get_scalars(grid, s)
    write(*,'(a,i8,a,es14.6,a,es12.4)') ' iteration', iter, '  t', s(1), '  dt', s(2)
    if (temp_end > 0) then
      ! check_temp at the top of the time loop (echo.f90:189-190): the grid
      ! stops stepping; the frame of the last state (echo.f90:264-272)
      ierr = echo_get_stop(grid, stopped, maxv)
      if (ierr /= 0) call fail('echo_get_stop', ierr)
      if (stopped /= 0) then
        if (freeze_kind == 0) then
          write(*,*) 'Temperature is under the value established for stopping the simulation'
        else
          write(*,*) 'Energy density is under the value established for stopping the simulation'
        end if
        call frame(iout)
        exit
      end if
    end if
    if (s(7) /= tout_prev) then
      ! an output time was reached (echo.f90:243-248; the grid advanced tout
      ! itself): the frame and the restart dump
      call frame(iout)
      if (restart_type /= 0) then
        ierr = echo_dump(grid, trim(dump_dir)//'/dump.dat'//c_null_char, iout, thyp)
        if (ierr /= 0) call fail('echo_dump', ierr)
      end if
      iout = iout + 1
    end if
    if (hyp_compu == 1) then
      ! echo.f90:252-258
      if ((s(1) > thyp) .or. ((thyp - s(1)) < 1.d-10)) then
        thyp = thyp + freeze_dt
        if ((abs(thyp - cfg%tmax) < 1.d-10) .or. (thyp > cfg%tmax)) thyp = cfg%tmax
        ierr = echo_fo_slice(grid, nfo)
        if (ierr /= 0) call fail('echo_fo_slice', ierr)
        ierr = echo_fo_write(grid, trim(hyp_file)//c_null_char, 0_c_int32_t)
        if (ierr /= 0) call fail('echo_fo_write', ierr)
      end if
    end if
  end do
  write(*,'(a,es12.4,a,i8,a)') ' zone-updates/s', dble(cfg%nx) * cfg%ny * cfg%nz * iter / ttot, &
                               '  (', iter, ' steps, wall time of the device calls)'
  ierr = echo_download_state(grid, v, u)
  if (ierr /= 0) call fail('echo_download_state', ierr)
  if (len_trim(dump_file) > 0) then
    open(31, file=trim(dump_file), access='stream', form='unformatted', status='replace')
    write(31) v, u
    close(31)
    write(*,*) 'state written to ', trim(dump_file)
  end if
  ierr = echo_wait_frames(grid)
  if (ierr /= 0) call fail('echo_wait_frames', ierr)
  call echo_destroy(grid)

contains

  ! out_vars: a frame with the selected fields when any is selected
  subroutine frame(io)
    integer(c_int32_t), intent(in) :: io
    if (all(sel == 0)) return
    ierr = echo_write_frame(grid, sel, int(outp_prec, c_int32_t), io, trim(outdir)//c_null_char)
    if (ierr /= 0) call fail('echo_write_frame', ierr)
  end subroutine frame

  ! INIT_TYPE 3: the events of generate_events (echo_glauber_generate into
  ! GLAUB_DIR), event EV_START read back, its energy deposition on the device
  ! (echo_init_glauber; ybeam = log(RADS / m_N), glaubermc.f90:741); on MHD
  ! grids with the field of the event's protons (echo_init_glauber_mhd,
  ! init.f90:2796-2811 without a B-field file; gamma_col = RADS / (2 m_N),
  ! common.f90:2261)
  subroutine glauber_init()
    type(echo_glauber_sampler) :: smp
    type(echo_glauber_event) :: ev
    type(echo_nucleons) :: lists(4)
    integer(c_int32_t) :: nev
    integer, parameter :: cap = 600
    real(c_double), allocatable, target :: pxy(:), sxy(:), cxy(:)
    integer(c_int32_t), allocatable, target :: pty(:), sty(:)
    real(8), parameter :: am = 0.93891897d0   ! nucleon mass, common.f90:54
    select case (trim(nucleus))
    case ('Au'); smp = echo_glauber_sampler(197, 79, 0, 0, 0, 0, 6.38d0, 0.535d0, 0.1693d0, 0d0, 0d0, 20d0)
    case ('Pb'); smp = echo_glauber_sampler(208, 82, 0, 0, 0, 0, 6.62d0, 0.546d0, 0.1604d0, 0d0, 0d0, 20d0)
    case default
      write(*,*) 'NUCLEUS ', trim(nucleus), ' is not supported (Au, Pb of nuclear_data.dat)'
      stop 1
    end select
    smp%nconf = int(nconf, c_int32_t)
    smp%nbcoll = int(nbcoll, c_int32_t)
    smp%min_participants = int(min_part, c_int32_t)
    smp%fixed_b = int(fixed_b, c_int32_t)
    smp%sigma_nn = sigma_in
    smp%bimpact = bimpact
    ierr = echo_glauber_generate(smp, seed, trim(glauber_dir)//c_null_char, nev)
    if (ierr /= 0) call fail('echo_glauber_generate', ierr)
    if (ev_start < 1 .or. ev_start > nev) then
      write(*,*) 'EV_START ', ev_start, ' but ', nev, ' events were generated'
      stop 1
    end if
    allocate(pxy(2 * cap), sxy(2 * cap), cxy(2 * cap * cap / 4), pty(cap), sty(cap))
    ev%part_xy = c_loc(pxy)
    ev%part_type = c_loc(pty)
    ev%spec_xy = c_loc(sxy)
    ev%spec_type = c_loc(sty)
    ev%coll_xy = c_loc(cxy)
    ierr = echo_glauber_read_event(trim(glauber_dir)//c_null_char, int(ev_start, c_int32_t), ev, &
                                   int(cap, c_int32_t), int(cap, c_int32_t), int(cap * cap / 8, c_int32_t))
    if (ierr /= 0) call fail('echo_glauber_read_event', ierr)
    write(*,'(a,i6,a,f8.3,a,i5,a,i6)') ' Glauber event', ev_start, '  b', ev%b, '  participants', &
        ev%n_part_a + ev%n_part_b, '  collisions', ev%n_coll
    gpar%ybeam = log(rads / am)
    if (cfg%mhd == 0) then
      ierr = echo_init_glauber(grid, gpar, pxy, ev%n_part_a, pxy(2 * ev%n_part_a + 1:), ev%n_part_b, &
                               cxy, ev%n_coll)
      if (ierr /= 0) call fail('echo_init_glauber', ierr)
      return
    end if
    ! spectators A, participants A, spectators B, participants B
    lists(1) = echo_nucleons(c_loc(sxy), c_loc(sty), ev%n_spec_a)
    lists(2) = echo_nucleons(c_loc(pxy), c_loc(pty), ev%n_part_a)
    lists(3) = echo_nucleons(c_loc(sxy(2 * ev%n_spec_a + 1)), c_loc(sty(ev%n_spec_a + 1)), ev%n_spec_b)
    lists(4) = echo_nucleons(c_loc(pxy(2 * ev%n_part_a + 1)), c_loc(pty(ev%n_part_a + 1)), ev%n_part_b)
    bpar%gamma_col = rads / (2d0 * am)
    ierr = echo_init_glauber_mhd(grid, gpar, pxy, ev%n_part_a, pxy(2 * ev%n_part_a + 1:), ev%n_part_b, &
                                 cxy, ev%n_coll, bpar, lists)
    if (ierr /= 0) call fail('echo_init_glauber_mhd', ierr)
  end subroutine glauber_init

  subroutine fail(what, code)
    character(len=*), intent(in) :: what
    integer, intent(in) :: code
    write(*,'(a,a,i6,a,a)') what, ' failed: ', code, ' ', trim(echo_error_message())
    stop 3
  end subroutine fail

  ! EoS table rows (T, e/T^4, p/T^4, c_s^2) of eos_data/<name>
  subroutine read_table(name, r, nr)
    character(len=*), intent(in) :: name
    real(8), allocatable, intent(inout) :: r(:,:)
    integer(c_int32_t), intent(out) :: nr
    integer :: i, nn
    open(30, file='eos_data/'//name, status='old', iostat=ios)
    if (ios /= 0) then
      write(*,*) 'cannot open eos_data/', name
      stop 1
    end if
    read(30, *) nn
    deallocate(r)
    allocate(r(4, nn))
    do i = 1, nn
      read(30, *) r(:, i)
    end do
    close(30)
    nr = int(nn, c_int32_t)
  end subroutine read_table

  ! ideal Gubser flow at tau = TSTART, q = 1 fm^-1 (init.f90:2926-2935)
  subroutine gubser_init(w)
    real(8), intent(out) :: w(:,:,:,:)
    real(8) :: dx, dy, x, y, r, tau, temp, e, kap, pi
    integer :: ix, iy
    pi = acos(-1d0)
    tau = cfg%tmin
    dx = (cfg%x2 - cfg%x1) / cfg%nx
    dy = (cfg%y2 - cfg%y1) / cfg%ny
    w = 0
    do iy = 1, cfg%ny
      do ix = 1, cfg%nx
        x = cfg%x1 + (ix - 0.5d0) * dx
        y = cfg%y1 + (iy - 0.5d0) * dy
        r = sqrt(x * x + y * y)
        temp = hbar * t0gub / tau * (4 * tau**2)**(1d0/3d0) / &
               (1d0 + 2d0 * (tau**2 + r**2) + (tau**2 - r**2)**2)**(1d0/3d0)
        e = 47.5d0 * pi**2 / (30d0 * hbar**3) * temp**4
        kap = atanh(2d0 * tau * r / (1d0 + tau**2 + r**2))
        w(ix, iy, :, 1) = 1d0
        w(ix, iy, :, 2) = x / r * tanh(kap)
        w(ix, iy, :, 3) = y / r * tanh(kap)
        w(ix, iy, :, 5) = e / 3d0
      end do
    end do
  end subroutine gubser_init

end program echo_amd_driver
