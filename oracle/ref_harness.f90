! Test-infrastructure harness: exposes the reference ECHO-QGP modules
! (compiled unmodified from the reference sources by oracle/build_ref.sh)
! through bind(C) entry points, so that tests can check the C oracle
! against the reference Fortran on identical inputs.
!
! Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
! the library this file builds (oracle/_ref/libechoref.so).  It is never
! part of the product path.
!
! The set-up follows the reference's own order:
!   echo.f90:73-109   system_setup -> parallel_grid_setup -> common_alloc ->
!                     set_eos_pointers -> table read -> find_przero ->
!                     evolve_alloc
!   common.f90:441-481 common_grid
!   init.f90:137-149  ibc filled with boundary_cond
!   init.f90:2926-2958 per-cell prim2cons followed by cons2prim
!   echo.f90:185-241  one time step (vnewstep/vold bookkeeping, RK stages,
!                     metric update after the step)
module ref_harness
  use iso_c_binding
  use parallel
  use common
  use eos
  use viscous
  use system_eqgp
  use work
  use evolve
  use holib
  implicit none

  real(8) :: h_dt = 0.d0
  real(8) :: h_dtout = 1.d30
  integer :: h_iter = 0

contains

  ! ic(1:3)  nx,ny,nz            ic(4)  coordinates (1 Minkowski, 2 Bjorken)
  ! ic(5)    viscosity 0/1       ic(6)  bulk viscosity 0/1
  ! ic(7)    mhd 0/1/2           ic(8)  divclean 0/1
  ! ic(9)    eqstate 1/2/3       ic(10) reconstruction code (see rec_name)
  ! ic(11)   nrk 2/3/4           ic(12) boundary_cond 0/2
  ! ic(13)   algo_solver         ic(14) obtained (0 'no', 1 'zz')
  ! ic(15)   init_type           ic(16) derivatives_out 0/1
  ! rc(1:6)  x1,x2,y1,y2,z1,z2   rc(7)  tmin      rc(8)  cfl
  ! rc(9)    maxdt               rc(10) maxspeed  rc(11) eta_over_s
  ! rc(12)   tau_pi_coeff        rc(13) enezero   rc(14) glm_alpha
  ! rc(15)   sigma_el_cond       rc(16) sigma_chiral_cond
  ! rc(17)   dtout               rc(18) smooth_temp (CUT_TEMP)
  ! rc(19)   tmax
  function rec_name(code) result(name)
    integer, intent(in) :: code
    character(5) :: name
    select case (code)
    case (1); name = 'TVD2 '
    case (2); name = 'CENO3'
    case (3); name = 'WENO3'
    case (4); name = 'WENO5'
    case (5); name = 'PPM4 '
    case (6); name = 'MPE3 '
    case (7); name = 'MPE5 '
    case (8); name = 'MPE7 '
    case default; name = 'WENO5'
    end select
  end function rec_name

  integer(c_int) function ref_setup(ic, rc, tabdir_len, tabname) bind(C, name="ref_setup")
    integer(c_int), intent(in) :: ic(16)
    real(c_double), intent(in) :: rc(19)
    integer(c_int), value :: tabdir_len
    character(kind=c_char), intent(in) :: tabname(*)
    integer :: i, error_flag

    error_flag = 0
    call release_all()
    nx = ic(1); ny = ic(2); nz = ic(3)
    coordinates = ic(4)
    viscosity = (ic(5) /= 0)
    bulkvis = (ic(6) /= 0) .and. viscosity
    mhd = (ic(7) /= 0)
    rmhd = (ic(7) == 2)
    divclean = (ic(8) /= 0) .and. mhd
    eqstate = ic(9)
    numerically = 0
    recal = rec_name(ic(10))
    nrk = ic(11)
    boundary_cond = ic(12)
    algo_solver = ic(13)
    if (ic(14) == 0) then
      obtained = 'no'
    else
      obtained = 'zz'
    end if
    init_type = ic(15)
    derivatives_out = (ic(16) /= 0)
    print_rho_comov = .false.
    out_sel%rc = 0

    x1 = rc(1); x2 = rc(2); y1 = rc(3); y2 = rc(4); z1 = rc(5); z2 = rc(6)
    tmin = rc(7); tauzero = tmin
    cfl = rc(8); maxdt = rc(9); maxspeed = rc(10)
    eta_over_s = rc(11); tau_pi_coeff = rc(12)
    enezero = rc(13); glm_alpha = rc(14)
    sigma_el_cond = rc(15); sigma_chiral_cond = rc(16)
    h_dtout = rc(17)
    smooth_temp = rc(18)
    enable_smooth_viscosity = (smooth_temp > 0.d0)
    tmax = rc(19)
    glm_ch_old = 0.5d0
    aflux = 0.d0

    if (eqstate == 3) then
      eos_tab_name = ''
      do i = 1, tabdir_len
        eos_tab_name(i:i) = tabname(i)
      end do
    end if

    call system_setup(coordinates, system_metric, viscosity, nv, nu, nov, nou)
    call parallel_grid_setup(nx, ny, nz, mx, my, mz, ix0, ix1, ix2, iy0, iy1, iy2, iz0, iz1, iz2)
    call common_alloc
    call common_grid(x1, x2, y1, y2, z1, z2)
    ibc = boundary_cond
    call set_eos_pointers(eqstate, eos_energy, eos_pressure, eos_sound)
    if (eqstate == 3) then
      call read_data_from_tabulated_file(entries, eos_tab_name, temperature_array, energy_density_array, &
           pressure_array, sound_speed_array, min_en_dens, max_en_dens)
    end if
    call find_przero(przero, enezero, error_flag)
    call evolve_alloc
    t = tmin
    timeold = t
    timeinterval = 0.d0
    g_cov_3_old = 0.d0
    call system_metric()
    h_dt = maxdt
    tout = t + h_dtout
    h_iter = 0
    run_crashed = .false.
    errcode = 0
    ref_setup = nv
  end function ref_setup

  ! Frees what common_alloc, evolve_alloc, work_alloc and the table reader
  ! allocated, so that ref_setup can be called again in one process.
  subroutine release_all()
    use out, only: arr_to_print_8, arr_to_print_8_bis, arr_to_print_8_ter, arr_to_print_8_quater, &
                   arr_to_print_8_quinquies
    call release_fo()
    if (allocated(arr_to_print_8)) deallocate(arr_to_print_8, arr_to_print_8_bis, arr_to_print_8_ter)
    if (allocated(arr_to_print_8_quater)) deallocate(arr_to_print_8_quater, arr_to_print_8_quinquies)
    if (allocated(prefix_dir)) deallocate(prefix_dir)
    if (allocated(recvcounts)) deallocate(recvcounts)
    if (allocated(displs)) deallocate(displs)
    if (allocated(kv)) deallocate(kv)
    if (allocated(ku)) deallocate(ku)
    if (allocated(v)) deallocate(v)
    if (allocated(u)) deallocate(u)
    if (allocated(deriv)) deallocate(deriv)
    if (allocated(vold)) deallocate(vold)
    if (allocated(uold)) deallocate(uold)
    if (allocated(vnewstep)) deallocate(vnewstep)
    if (allocated(unewstep)) deallocate(unewstep)
    if (allocated(x)) deallocate(x, ddx, x0, y, ddy, y0, z, ddz, z0)
    if (allocated(ibc)) deallocate(ibc)
    if (allocated(vall)) deallocate(vall)
    if (allocated(vderivatives)) deallocate(vderivatives)
    if (allocated(w)) deallocate(w)
    if (allocated(u0)) deallocate(u0)
    if (allocated(ax)) deallocate(ax, ay, az)
    if (allocated(du)) deallocate(du)
    if (allocated(du_stiff)) deallocate(du_stiff)
    if (allocated(w0)) deallocate(w0, w1, wl, wr)
    if (allocated(wd)) deallocate(wd, wdl, wdr)
    if (allocated(temperature_array)) deallocate(temperature_array)
    if (allocated(energy_density_array)) deallocate(energy_density_array)
    if (allocated(pressure_array)) deallocate(pressure_array)
    if (allocated(sound_speed_array)) deallocate(sound_speed_array)
  end subroutine release_all

  integer(c_int) function ref_nv() bind(C, name="ref_nv")
    ref_nv = nv
  end function ref_nv

  ! Per-cell prim2cons then cons2prim, as the reference initialisers do
  ! (init.f90:2937-2958).  vin holds (nx,ny,nz,nv) primitives, Fortran order.
  integer(c_int) function ref_init_from_prim(vin) bind(C, name="ref_init_from_prim")
    real(c_double), intent(in) :: vin(nx, ny, nz, nv)
    real(8), allocatable :: vv(:), uu(:)
    integer :: ix, iy, iz, error_flag, iter_error
    allocate(vv(nv), uu(nv))
    ref_init_from_prim = 0
    do iz = 1, nz
      do iy = 1, ny
        do ix = 1, nx
          vv(1:nv) = vin(ix, iy, iz, 1:nv)
          if (viscosity) then
            call system_prim2cons_0(vv, uu, error_flag)
          else
            call system_prim2cons(vv, uu, error_flag)
          end if
          if (error_flag > 0) then
            ref_init_from_prim = error_flag
            return
          end if
          if (.not. viscosity) then
            call system_cons2prim(uu, vv, iter_error, error_flag)
            if (error_flag > 0) then
              ref_init_from_prim = error_flag
              return
            end if
          end if
          v(ix, iy, iz, 1:nv) = vv(1:nv)
          u(ix, iy, iz, 1:nu) = uu(1:nu)
        end do
      end do
    end do
  end function ref_init_from_prim

  subroutine ref_set_state(vin, uin) bind(C, name="ref_set_state")
    real(c_double), intent(in) :: vin(nx, ny, nz, nv), uin(nx, ny, nz, nv)
    v(1:nx, 1:ny, 1:nz, 1:nv) = vin
    u(1:nx, 1:ny, 1:nz, 1:nu) = uin
  end subroutine ref_set_state

  subroutine ref_get_state(vout, uout) bind(C, name="ref_get_state")
    real(c_double), intent(out) :: vout(nx, ny, nz, nv), uout(nx, ny, nz, nv)
    vout = v(1:nx, 1:ny, 1:nz, 1:nv)
    uout = u(1:nx, 1:ny, 1:nz, 1:nu)
  end subroutine ref_get_state

  subroutine ref_get_deriv(dout) bind(C, name="ref_get_deriv")
    real(c_double), intent(out) :: dout(nx, ny, nz, nderivatives)
    dout = deriv(1:nx, 1:ny, 1:nz, 1:nderivatives)
  end subroutine ref_get_deriv

  ! scal(1) t, (2) dt, (3) timeold, (4) timeinterval, (5) g_cov_3_old,
  ! (6) glm_ch_old, (7) tout, (8) iterations, (9) run_crashed, (10) errcode
  subroutine ref_get_scalars(scal) bind(C, name="ref_get_scalars")
    real(c_double), intent(out) :: scal(10)
    scal(1) = t; scal(2) = h_dt; scal(3) = timeold; scal(4) = timeinterval
    scal(5) = g_cov_3_old; scal(6) = glm_ch_old; scal(7) = tout
    scal(8) = dble(h_iter)
    if (run_crashed) then
      scal(9) = 1.d0
    else
      scal(9) = 0.d0
    end if
    scal(10) = dble(errcode)
  end subroutine ref_get_scalars

  ! nsteps iterations of the main-loop body of echo.f90:185-262 (without
  ! file output).  Returns 0, or the error code of a crashed step.
  integer(c_int) function ref_steps(nsteps) bind(C, name="ref_steps")
    integer(c_int), value :: nsteps
    integer :: n, irk
    ref_steps = 0
    do n = 1, nsteps
      if (mhd) then
        vnewstep = v(:, :, :, krh:kbz)
        unewstep = u(:, :, :, krh:kbz)
      else
        vnewstep = v(:, :, :, krh:kpr)
        unewstep = u(:, :, :, krh:kpr)
      end if
      do irk = 1, nrk
        call evolve_main(irk, h_dt)
        if (run_crashed) exit
      end do
      if (run_crashed) then
        ref_steps = max(errcode, 1)
        return
      end if
      timeold = t
      t = t + h_dt
      g_cov_3_old = g_cov(3)
      call system_metric()
      h_iter = h_iter + 1
      timeinterval = h_dt
      vold = vnewstep
      uold = unewstep
      if ((t > tout) .or. ((tout - t) < 1.e-10)) then
        tout = tout + h_dtout
        if ((abs(tout - tmax) < 1.e-10) .or. (tout > tmax)) tout = tmax
      end if
    end do
  end function ref_steps

  ! evolve_u2v (evolve.f90:769-868) alone on the current v, u: pins the
  ! con2prim repair on host-forced states.  Returns errcode if the run
  ! crashed, else 0.
  integer(c_int) function ref_u2v() bind(C, name="ref_u2v")
    run_crashed = .false.
    errcode = 0
    call evolve_u2v
    ref_u2v = 0
    if (run_crashed) ref_u2v = max(errcode, 1)
  end function ref_u2v

  ! One RK stage, evolve_main(irk, dt) of evolve.f90:114-298, without the
  ! step bookkeeping of ref_steps (for stage-by-stage comparisons).  irk = 1
  ! also takes the step-start snapshot vnewstep / unewstep.
  integer(c_int) function ref_stage(irk) bind(C, name="ref_stage")
    integer(c_int), value :: irk
    if (irk == 1) then
      if (mhd) then
        vnewstep = v(:, :, :, krh:kbz)
        unewstep = u(:, :, :, krh:kbz)
      else
        vnewstep = v(:, :, :, krh:kpr)
        unewstep = u(:, :, :, krh:kpr)
      end if
    end if
    call evolve_main(irk, h_dt)
    ref_stage = 0
    if (run_crashed) ref_stage = max(errcode, 1)
  end function ref_stage

  ! ---------------------------------------------------------------------
  ! Function-level entry points (pointwise operators of system.f90,
  ! holib.f90, eos.f90, viscous.f90), all at the current metric.
  ! ---------------------------------------------------------------------

  subroutine ref_set_time(tt) bind(C, name="ref_set_time")
    real(c_double), value :: tt
    t = tt
    call system_metric()
  end subroutine ref_set_time

  ! One reconstruction of a pencil f(-ngc_in+1 : n+ngc_in) given as a flat
  ! array of n+2*ngc_in values, left/right states at interfaces 0..n.
  subroutine ref_rec(code, n, ngc_in, f, gl, gr) bind(C, name="ref_rec")
    integer(c_int), value :: code, n, ngc_in
    real(c_double), intent(in) :: f(n + 2*ngc_in)
    real(c_double), intent(out) :: gl(n + 1), gr(n + 1)
    real(8), allocatable :: ff(:), l(:), r(:)
    allocate(ff(1-ngc_in:n+ngc_in), l(0:n), r(0:n))
    ff = f
    select case (code)
    case (1); call holib_rectvd2(ff(-1:n+2), l, r, n)
    case (2); call holib_recceno3(ff(-2:n+3), l, r, n)
    case (3); call holib_recweno3(ff(-1:n+2), l, r, n)
    case (4); call holib_recweno5(ff(-2:n+3), l, r, n)
    case (5); call holib_recppm4(ff(-1:n+2), l, r, n)
    case (6); call holib_recmpe3(ff(-2:n+3), l, r, n)
    case (7); call holib_recmpe5(ff(-2:n+3), l, r, n)
    case (8); call holib_recmpe7(ff(-3:n+4), l, r, n)
    end select
    gl = l
    gr = r
  end subroutine ref_rec

  integer(c_int) function ref_prim2cons(vin, uout) bind(C, name="ref_prim2cons")
    real(c_double), intent(inout) :: vin(nv)
    real(c_double), intent(out) :: uout(nv)
    integer :: err
    call system_prim2cons(vin, uout, err)
    ref_prim2cons = err
  end function ref_prim2cons

  integer(c_int) function ref_cons2prim(uin, vinout) bind(C, name="ref_cons2prim")
    real(c_double), intent(inout) :: uin(nv), vinout(nv)
    integer :: err, iterr
    err = 0
    call system_cons2prim(uin, vinout, iterr, err)
    ref_cons2prim = err
  end function ref_cons2prim

  integer(c_int) function ref_flux(vin, dir, uout, fout, vf) bind(C, name="ref_flux")
    real(c_double), intent(inout) :: vin(nv)
    integer(c_int), value :: dir
    real(c_double), intent(out) :: uout(nv), fout(nv), vf(2)
    integer :: err
    call system_flux(vin, uout, fout, vf, dir, err)
    ref_flux = err
  end function ref_flux

  ! eos_pressure(rh,en) -> pr, dprdrh, dprden   (eos.f90 procedure pointer)
  integer(c_int) function ref_eos_pressure(rh, en, out) bind(C, name="ref_eos_pressure")
    real(c_double), value :: rh, en
    real(c_double), intent(out) :: out(3)
    integer :: err
    real(8) :: pr, d1, d2
    err = 0
    call eos_pressure(rh, en, pr, d1, d2, err)
    out(1) = pr; out(2) = d1; out(3) = d2
    ref_eos_pressure = err
  end function ref_eos_pressure

  integer(c_int) function ref_eos_energy(rh, pr, en) bind(C, name="ref_eos_energy")
    real(c_double), value :: rh, pr
    real(c_double), intent(out) :: en
    integer :: err
    err = 0
    call eos_energy(rh, en, pr, err)
    ref_eos_energy = err
  end function ref_eos_energy

  integer(c_int) function ref_eos_sound(rh, en, pr, cs2) bind(C, name="ref_eos_sound")
    real(c_double), value :: rh, en, pr
    real(c_double), intent(out) :: cs2
    integer :: err
    err = 0
    call eos_sound(rh, en, pr, cs2, err)
    ref_eos_sound = err
  end function ref_eos_sound

  ! get_derived_data(rh,pr) -> en, temp, entropy density
  integer(c_int) function ref_eos_derived(rh, pr, out) bind(C, name="ref_eos_derived")
    real(c_double), value :: rh, pr
    real(c_double), intent(out) :: out(3)
    integer :: err
    real(8) :: en, temp, s
    err = 0
    call get_derived_data(rh, pr, en, temp, s, err)
    out(1) = en; out(2) = temp; out(3) = s
    ref_eos_derived = err
  end function ref_eos_derived

  ! Table as converted by read_data_from_tabulated_file (eos.f90:265-334).
  integer(c_int) function ref_eos_table(n, tarr, earr, parr, carr) bind(C, name="ref_eos_table")
    integer(c_int), value :: n
    real(c_double), intent(out) :: tarr(n), earr(n), parr(n), carr(n)
    integer :: m
    m = min(n, entries)
    tarr(1:m) = temperature_array(1:m)
    earr(1:m) = energy_density_array(1:m)
    parr(1:m) = pressure_array(1:m)
    carr(1:m) = sound_speed_array(1:m)
    ref_eos_table = entries
  end function ref_eos_table

  real(c_double) function ref_przero() bind(C, name="ref_przero")
    ref_przero = przero
  end function ref_przero

  ! out_vars (out.f90:232-591) on the current state: sel(1:28) are the
  ! out_sel flags in the order of type out_type (common.f90:122-151);
  ! writes <dir>/outr0001/out<iout>.dat and appends <dir>/outr0001/time.dat
  subroutine ref_out_vars(sel, precision, iout, dirlen, dir) bind(C, name="ref_out_vars")
    use out
    integer(c_int), intent(in) :: sel(28)
    integer(c_int), value :: precision, iout, dirlen
    character(kind=c_char), intent(in) :: dir(*)
    integer :: i, io
    custom_output_directory = ' '
    do i = 1, dirlen
      custom_output_directory(i:i) = dir(i)
    end do
    out_sel%density = sel(1); out_sel%vx = sel(2); out_sel%vy = sel(3); out_sel%vz = sel(4)
    out_sel%pressure = sel(5); out_sel%energy_density = sel(6); out_sel%temperature = sel(7)
    out_sel%entropy_density = sel(8); out_sel%bulk = sel(9); out_sel%pitt = sel(10)
    out_sel%pitx = sel(11); out_sel%pity = sel(12); out_sel%pitz = sel(13); out_sel%pixy = sel(14)
    out_sel%pixz = sel(15); out_sel%piyz = sel(16); out_sel%pixx = sel(17); out_sel%piyy = sel(18)
    out_sel%pizz = sel(19); out_sel%v0 = sel(20); out_sel%bx = sel(21); out_sel%by = sel(22)
    out_sel%bz = sel(23); out_sel%ex = sel(24); out_sel%ey = sel(25); out_sel%ez = sel(26)
    out_sel%glm = sel(27); out_sel%rc = sel(28)
    output_precision = precision
    out_freeze = .false.
    derivatives_out = .false.
    flows_out = .false.
    id_of_the_run = 1
    if (.not. allocated(arr_to_print_8)) call out_alloc
    io = iout
    call out_vars(io, 0)
  end subroutine ref_out_vars

  ! Frees what allocate_hypersurface_arrays (hypersurface.f90:93-168) and
  ! out_alloc (out.f90:62-120) allocated, and clears the slice times, so that
  ! ref_fo_setup can be called again in one process.
  subroutine release_fo()
    use out, only: arr_to_print_8, arr_to_print_8_bis, arr_to_print_8_ter, arr_to_print_8_quater, &
                   arr_to_print_8_quinquies
    use hypersurface
    if (allocated(field)) deallocate(field)
    if (allocated(v_backup)) deallocate(v_backup)
    if (allocated(primitives)) deallocate(primitives)
    if (allocated(der_backup)) deallocate(der_backup)
    if (allocated(hy_ide)) deallocate(hy_ide)
    if (allocated(hy_vis)) deallocate(hy_vis)
    if (allocated(hy_deriv)) deallocate(hy_deriv)
    if (allocated(hy_mhd)) deallocate(hy_mhd)
    if (allocated(arr_to_print_8)) deallocate(arr_to_print_8, arr_to_print_8_bis, arr_to_print_8_ter)
    if (allocated(arr_to_print_8_quater)) deallocate(arr_to_print_8_quater, arr_to_print_8_quinquies)
    if (allocated(prefix_dir)) deallocate(prefix_dir)
    if (allocated(recvcounts)) deallocate(recvcounts)
    if (allocated(displs)) deallocate(displs)
    if (allocated(uall)) deallocate(uall, vallold, uallold)
    time_arr = 0.d0
  end subroutine release_fo

  ! Freeze-out set-up as echo.f90:80-148 does it with OUT_FREEZE on:
  ! out_alloc (which calls allocate_hypersurface_arrays, out.f90:81-84) and
  ! get_hysufile_ready (hypersurface.f90:1137-1200), which writes the header
  ! of <dir>/outr0001/hypersurface.txt.
  subroutine ref_fo_setup(ftype, fvalue, dirlen, dir) bind(C, name="ref_fo_setup")
    use out
    use hypersurface
    integer(c_int), value :: ftype, dirlen
    real(c_double), value :: fvalue
    character(kind=c_char), intent(in) :: dir(*)
    integer :: i
    call release_fo()
    custom_output_directory = ' '
    do i = 1, dirlen
      custom_output_directory(i:i) = dir(i)
    end do
    id_of_the_run = 1
    out_freeze = .true.
    derivatives_out = .false.
    print_rho_comov = .false.
    restart_type = 1        ! out_alloc also allocates the dump arrays
    restarted_run = .false.
    freeze_type = ftype
    freeze_value = fvalue
    call out_alloc
    call get_hysufile_ready()
  end subroutine ref_fo_setup

  ! out_hyp (out.f90:1568-1645) on the current state at the current time:
  ! the first call stores the slice, later calls find the crossings between
  ! the stored slice and this one and append them to hypersurface.txt.
  ! Returns 1 when out_hyp flagged run_crashed.
  integer(c_int) function ref_fo_slice() bind(C, name="ref_fo_slice")
    use out
    use hypersurface
    integer :: io
    t_collect_d_a = -1.d30
    run_crashed = .false.
    io = 1
    call out_hyp(io)
    ref_fo_slice = merge(1, 0, run_crashed)
  end function ref_fo_slice

  ! the stored elements: hy_ide(1:n, 1:13), then hy_vis(1:n, 1:7) (viscous)
  ! or hy_mhd(1:n, 1:3) (MHD), as rows of out(1:nw, 1:n)
  subroutine ref_fo_get(n, nw, out) bind(C, name="ref_fo_get")
    use hypersurface
    integer(c_int), value :: n, nw
    real(c_double), intent(out) :: out(nw, n)
    integer :: i
    do i = 1, n
      out(1:13, i) = hy_ide(i, 1:13)
      if (nw >= 20) out(14:20, i) = hy_vis(i, 1:7)
      if (nw == 16) out(14:16, i) = hy_mhd(i, 1:3)
    end do
  end subroutine ref_fo_get

  ! Output set-up for restart dumps without the freeze-out hypersurface:
  ! out_alloc with restart_type = 1 (out.f90:108-118), files under
  ! <dir>/outr0001/
  subroutine ref_dump_setup(dirlen, dir) bind(C, name="ref_dump_setup")
    use out
    use hypersurface
    integer(c_int), value :: dirlen
    character(kind=c_char), intent(in) :: dir(*)
    integer :: i
    call release_fo()
    custom_output_directory = ' '
    do i = 1, dirlen
      custom_output_directory(i:i) = dir(i)
    end do
    id_of_the_run = 1
    out_freeze = .false.
    derivatives_out = .false.
    restart_type = 1
    restarted_run = .false.
    call out_alloc
  end subroutine ref_dump_setup

  ! out_dump (out.f90:1313-1390) of the current state, after the collection
  ! that out_vars / out_hyp make before it; thyp is the freeze-out cadence
  ! time it records
  subroutine ref_out_dump(iout, thyp_in) bind(C, name="ref_out_dump")
    use out
    integer(c_int), value :: iout
    real(c_double), value :: thyp_in
    t_collect_d_a = -1.d30
    call collect_distributed_arrays
    thyp = thyp_in
    call out_dump(iout)
  end subroutine ref_out_dump

end module ref_harness
