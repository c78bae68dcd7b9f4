! cube_p3m_driver.f90 -- a one-image Fortran host that calls the step through
! the ISO_C_BINDING interface, the way kick.f90 / drift.f90 callers would.
!
! Usage: cube_p3m_driver <dir> <ng> <np> <a_mid> <dt> <dt_old> <dt_drift>
!   or   cube_p3m_driver --version | --sizes
!   or   cube_p3m_driver --write-header <file>   (a sim_header with known values,
!        written the way checkpoint.f90:22 writes it)
!
! Reads the buffered layout from raw stream files in <dir> (the format of the
! reference's checkpoint streams, checkpoint.f90:11-33): xp.bin (int16,
! 3*np), vp.bin (real32, 3*np), pid.bin (int64, np), rhoc.bin (int32), then
!   1. kick (cube_p3m_kick) and writes vp_kick.bin and kick.txt,
!   2. buffer_v (cube_p3m_buffer_v): the buffer copies take the kicked
!      velocities, as main.f90:37-38 orders kick and buffer_v,
!   3. drift with the kicked velocities (cube_p3m_drift) and writes
!      xp_drift.bin, vp_drift.bin, pid_drift.bin, rhoc_drift.bin,
!      spine.bin (idx_b_l, idx_b_r, ppl0, pplr, pprl, ppr0, ppl, ppr) and
!      drift.txt.
! Exit status 0 on success; 3 when no GPU is visible.
program cube_p3m_driver
  use iso_c_binding
  use cube_p3m_iface
  implicit none
  character(512) :: dir, arg
  integer :: ng, nc, nt, nte, nnt, ncb, u
  integer(c_int64_t) :: np, nrc, nsp, nsb
  real(c_float) :: a_mid, dt, dt_old, dt_drift
  integer(c_int16_t), allocatable :: xp(:)
  real(c_float), allocatable :: vp(:)
  integer(c_int64_t), allocatable :: pid(:), idx_b_l(:), idx_b_r(:), sp(:, :)
  integer(c_int32_t), allocatable :: rhoc(:)
  type(cube_p3m_config) :: cfg
  type(c_ptr) :: ctx
  type(cube_kick_in) :: kin
  type(cube_kick_out) :: kout
  type(cube_drift_in) :: din
  type(cube_drift_out) :: dout
  integer(c_int) :: st
  ! sim_header of parameters.f90:183-210 (not bind(C): written component by component)
  type sim_header
    integer(8) nplocal, npglobal, izipx, izipv, image, nn, nnt, nt, ncell, ncb, timestep, cur_checkpoint, &
        cur_halofind, cur_powerpoint, calculate_PK, cic_iapm
    real a, t, tau, dt_pp, dt_pm3, dt_pm2, dt_pm1, dt_vmax, mass_p_cdm, m_nu(3), Mass_nu, box, h0, omega_m, &
        omega_nu, omega_l, s8, vsim2phys, sigma_vres, sigma_vi, z_i, vz_max
  endtype
  type(sim_header) :: sim

  call get_command_argument(1, arg)
  if (trim(arg) == '--write-header') then
    call get_command_argument(2, arg)
    sim = sim_header(1234567_8, 1234567_8, 2, 2, 1, 1, 4, 8, 32, 4, 17, 3, 0, 0, 0, 0, &
                     0.25, 1.5, 2.5, 0.01, 0.02, 0.03, 0.04, 0.05, 1.0, [0.02, 0.02, 0.02], 0.06, 600.0, 0.6817, &
                     0.301, 0.0, 0.699, 0.8, 12.5, 0.0, 3.5, 49.0, 4.3)
    open(newunit=u, file=trim(arg), access='stream', form='unformatted', status='replace')
    write(u) sim
    close(u)
    stop
  endif
  if (trim(arg) == '--version') then
    print '(a)', c_string(cube_p3m_version())
    stop
  endif
  if (trim(arg) == '--sizes') then  ! bind(C) type sizes, to check against the C structs
    print '(6i6)', c_sizeof(cfg), c_sizeof(kin), c_sizeof(kout), c_sizeof(din), c_sizeof(dout), &
        c_sizeof(cube_ic_params(0, 0d0, 0d0, 0d0, 0d0, 0d0, 0d0, 0, 0.0))
    stop
  endif
  dir = arg
  call get_command_argument(2, arg); read(arg, *) ng
  call get_command_argument(3, arg); read(arg, *) np
  call get_command_argument(4, arg); read(arg, *) a_mid
  call get_command_argument(5, arg); read(arg, *) dt
  call get_command_argument(6, arg); read(arg, *) dt_old
  call get_command_argument(7, arg); read(arg, *) dt_drift

  cfg = cube_default_config(ng, 600.0 * ng / 512.0)
  nnt = cfg%nnt; ncb = cfg%ncb
  nc = ng / cfg%ratio_cs; nt = nc / nnt; nte = nt + 2 * ncb
  nrc = int(nte, c_int64_t)**3 * nnt**3
  nsb = int(nte, c_int64_t)**2 * nnt**3
  nsp = int(nt, c_int64_t)**2 * nnt**3

  st = cube_p3m_create(cfg, ctx)
  if (st == CUBE_ERR_NO_DEVICE) then
    print '(a)', 'no GPU visible'
    call exit(3)
  endif
  call cube_check(st, ctx, 'cube_p3m_create')

  allocate(xp(3 * np), vp(3 * np), pid(np), rhoc(nrc), idx_b_l(nsb), idx_b_r(nsb), sp(nsp, 6))
  open(newunit=u, file=trim(dir)//'/xp.bin', access='stream', form='unformatted', status='old'); read(u) xp; close(u)
  open(newunit=u, file=trim(dir)//'/vp.bin', access='stream', form='unformatted', status='old'); read(u) vp; close(u)
  open(newunit=u, file=trim(dir)//'/pid.bin', access='stream', form='unformatted', status='old'); read(u) pid; close(u)
  open(newunit=u, file=trim(dir)//'/rhoc.bin', access='stream', form='unformatted', status='old'); read(u) rhoc
  close(u)
  open(newunit=u, file=trim(dir)//'/idx_b_r.bin', access='stream', form='unformatted', status='old'); read(u) idx_b_r
  close(u)

  kin%a_mid = real(a_mid, c_double); kin%dt = real(dt, c_double); kin%mass_p_cdm = 1.0; kin%phase_mask = 0
  call cube_check(cube_p3m_kick(ctx, xp, vp, rhoc, idx_b_r, np, kin, kout), ctx, 'kick')
  open(newunit=u, file=trim(dir)//'/vp_kick.bin', access='stream', form='unformatted', status='replace')
  write(u) vp; close(u)
  open(newunit=u, file=trim(dir)//'/kick.txt', status='replace')
  write(u, '(6es24.16)') kout%dt_pm1, kout%dt_pm2, kout%dt_pm3, kout%dt_pp, kout%dt_vmax, kout%vz_max
  write(u, '(7i12)') kout%napm
  write(u, '(es24.16)') kout%testrho
  close(u)
  call cube_check(cube_p3m_buffer_v(ctx, vp, pid, rhoc, np), ctx, 'buffer_v')

  din%dt_old = real(dt_old, c_double); din%dt = real(dt_drift, c_double); din%vz_max = kout%vz_max
  call cube_check(cube_p3m_drift(ctx, xp, vp, pid, rhoc, np, idx_b_l, idx_b_r, sp(:, 1), sp(:, 2), sp(:, 3), &
                                 sp(:, 4), sp(:, 5), sp(:, 6), din, dout), ctx, 'drift')
  open(newunit=u, file=trim(dir)//'/xp_drift.bin', access='stream', form='unformatted', status='replace')
  write(u) xp(1:3 * dout%nplocal); close(u)
  open(newunit=u, file=trim(dir)//'/vp_drift.bin', access='stream', form='unformatted', status='replace')
  write(u) vp(1:3 * dout%nplocal); close(u)
  open(newunit=u, file=trim(dir)//'/pid_drift.bin', access='stream', form='unformatted', status='replace')
  write(u) pid(1:dout%nplocal); close(u)
  open(newunit=u, file=trim(dir)//'/rhoc_drift.bin', access='stream', form='unformatted', status='replace')
  write(u) rhoc; close(u)
  open(newunit=u, file=trim(dir)//'/spine.bin', access='stream', form='unformatted', status='replace')
  write(u) idx_b_l, idx_b_r, sp; close(u)
  open(newunit=u, file=trim(dir)//'/drift.txt', status='replace')
  write(u, '(i16, es24.16, i8)') dout%nplocal, dout%overhead_tile, dout%nlayer
  close(u)

  call cube_check(cube_p3m_destroy(ctx), c_null_ptr, 'cube_p3m_destroy')
end program
