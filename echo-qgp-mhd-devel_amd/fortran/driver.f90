! echo_amd_driver: a Fortran caller of the MI355X grid update, reading the
! reference's param.dat format (10-character flags, common.f90:874-1215) and
! running the time loop of echo.f90:185-262 through the C-ABI.
!
! Usage: echo_amd_driver [param_file]
! Supported reference flags: INIT_TYPE (4, ideal Gubser flow), COORD,
! VISCOUS, BULK, MHD, DIVCLEAN, GLM_PARAM, CUT_TEMP, NX, NY, NZ, XMIN..ZMAX,
! TSTART, TSTOP, DTOUT, MAXDT, CFL, REC_ALGO, INT_ALGO, MAXSPEED, ENEZERO,
! ETA_S, TAU_PI_C, TRACE_IMP, EOS, EOS_FILE, ALG_SOLV, EL_COND, CHIR_COND,
! bound_con; other reference flags are ignored.  Driver flags: GUB_T0...
! (T0 of the Gubser initialiser in fm^-1; the reference fixes 2.0 at
! common.f90:1468), NSTEPS...= (step limit), CHUNK....= (steps per device
! call), DUMP_FILE= (final v then u, FP64 stream), DEVICE...=.
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
  integer(c_int32_t) :: iout
  integer(c_int64_t) :: nfo
  real(8), allocatable :: v(:,:,:,:), u(:,:,:,:), rows(:,:)
  integer(8) :: c0, c1, crate
  real(8), parameter :: hbar = 0.197326d0

  pfile = 'param.dat'
  if (command_argument_count() >= 1) call get_command_argument(1, pfile)
  ! defaults of common.f90
  cfg = echo_config(0, 0, 0, 2, 0, 0, 0, 0, 1, 4, 3, 2, 5, 1, 4, 0, &
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
      case default
        write(*,*) 'INT_ALGO ', trim(val), ' is not supported (RK2, RK3)'
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
  if (cfg%init_type /= 4) then
    write(*,*) 'this driver initialises INIT_TYPE=4 (ideal Gubser flow) only; ', &
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
    call gubser_init(v)
    ierr = echo_init_from_prim(grid, v)
    if (ierr /= 0) call fail('echo_init_from_prim', ierr)
    ierr = echo_set_tout(grid, cfg%tmin + cfg%dtout)
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
    if (s(7) /= tout_prev) then
      ! an output time was reached (echo.f90:243-248; the grid advanced tout
      ! itself): the restart dump
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
  call echo_destroy(grid)

contains

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
