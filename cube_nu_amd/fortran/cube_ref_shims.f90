! cube_ref_shims.f90 -- the reference-side binding a CUBE maintainer adds:
! drop-in `kick`, `drift`, `buffer_v` and `particle_mesh` subroutines that keep
! the reference's argument-less call surface (kick.f90:1, drift.f90:1,
! particle_mesh.f90:1; callers main.f90:34-38) and call libcube_p3m through
! the ISO_C_BINDING interface of cube_p3m_iface.f90.
!
! The `parameters` and `variables` modules here are stand-ins with the
! reference's names, kinds and array shapes (parameters.f90:8-132,
! variables.f90:21-113) at one compile-time size (ng = 64, one image), so that
! the shims compile and run as they would inside CUBE; in CUBE the real
! modules replace them.  The program at the end is a one-image host: it fills
! the module arrays from raw stream files, then calls kick, buffer_v and drift
! the way main.f90:34-38 does, and writes the results.
!
! Usage: cube_ref_host <dir> <np> <a_mid> <dt> <dt_old> <dt_drift>
! (dt_old is the value before the kick; the drift after it uses dt_old = dt
! and dt = dt_drift, as the next step's timestep sets them)
module parameters
  implicit none
  integer, parameter :: ng = 64, nnt = 4, nns = 4, ratio_cs = 4, ncb = 4, ngb = 16, izipx = 2
  integer, parameter :: nc = ng / ratio_cs, nt = nc / nnt, nte = nt + 2 * ncb, ngp = ng / nnt
  integer, parameter :: ngt = ngp + 2 * ngb
  integer(8), parameter :: np_image_max = int(ng, 8)**3 * 4
  real, parameter :: box = 600.0 * ng / 512.0, apm1c = 3.5, apm2 = 3.5, apm3f = 3.5, app = 0.06
  real, parameter :: Mass_nu = 0.0, f_nu = 0.0
end module

module variables
  use parameters
  implicit none
  ! the sim header fields the step reads and writes (parameters.f90:183-210)
  type sim_state
    integer(8) :: nplocal = 0
    real :: dt_pp = 0, dt_pm3 = 0, dt_pm2 = 0, dt_pm1 = 0, dt_vmax = 0, mass_p_cdm = 1, vz_max = 0
  end type
  type(sim_state) :: sim
  real :: a_mid, dt, dt_old                                   ! variables.f90:35 (real(4))
  real(8) :: testrho                                          ! particle_mesh.f90:106
  real :: overhead_tile                                       ! drift.f90:85
  integer :: napm(7)                                          ! kick.f90:109-133
  integer(2) :: xp(3, np_image_max)                           ! integer(izipx)
  real :: vp(3, np_image_max)
  integer(8) :: pid(np_image_max)
  integer(4) :: rhoc(1-ncb:nt+ncb, 1-ncb:nt+ncb, 1-ncb:nt+ncb, nnt, nnt, nnt)
  integer(8) :: idx_b_l(1-ncb:nt+ncb, 1-ncb:nt+ncb, nnt, nnt, nnt), idx_b_r(1-ncb:nt+ncb, 1-ncb:nt+ncb, nnt, nnt, nnt)
  integer(8), dimension(nt, nt, nnt, nnt, nnt) :: ppl0, pplr, pprl, ppr0, ppl, ppr
  real :: tf1(nc / 2 + 1, nc, nc), tf2(ngt / 2 + 1, ngt, ngt)
end module

! The context, created once after the Green's functions exist
! (initialize.f90:334) and destroyed in finalize (finalize.f90:11-12).
module cube_gpu
  use iso_c_binding
  use cube_p3m_iface
  implicit none
  type(c_ptr) :: cube_ctx = c_null_ptr
contains
  subroutine cube_gpu_init()            ! call at the end of initialize
    use parameters
    type(cube_p3m_config) :: cfg
    cfg = cube_default_config(int(ng), box)
    cfg%nnt = nnt; cfg%nns = nns; cfg%ratio_cs = ratio_cs; cfg%ncb = ncb; cfg%ngb = ngb
    cfg%izipx = izipx; cfg%apm1c = apm1c; cfg%apm2 = apm2; cfg%apm3f = apm3f; cfg%app = app
    cfg%f_nu = f_nu; cfg%mass_nu = Mass_nu
    call cube_check(cube_p3m_create(cfg, cube_ctx), cube_ctx, 'cube_p3m_create')
  end subroutine
  subroutine cube_gpu_tf()              ! after get_tf_cb2matter rewrote tf* (power_nu.f90:424-431)
    use variables
    call cube_check(cube_p3m_set_tf(cube_ctx, 1, tf1, size(tf1, kind=8)), cube_ctx, 'tf1')
    call cube_check(cube_p3m_set_tf(cube_ctx, 2, tf2, size(tf2, kind=8)), cube_ctx, 'tf2')
    ! levels 3..6: tf3_2, tf3_4, tf3_6, tf3_8
  end subroutine
  subroutine cube_gpu_finalize()
    integer(c_int) :: st
    st = cube_p3m_destroy(cube_ctx)
    cube_ctx = c_null_ptr
  end subroutine
end module

! kick.f90 (the sim fields and module scalars it used to write are set from
! cube_kick_out; kick.f90:39, 69, 239, 298, 314-315)
subroutine kick
  use variables
  use cube_gpu
  implicit none
  type(cube_kick_in) :: kin
  type(cube_kick_out) :: kout
  ! a_mid and dt widen exactly to the FP64 fields; the library forms the kick
  ! coefficient a_mid*dt/6/pi in FP64
  kin = cube_kick_in(real(a_mid, 8), real(dt, 8), sim%mass_p_cdm, 0)   ! phase_mask 0 = PM1+PM2+PM3+PP
  call cube_check(cube_p3m_kick(cube_ctx, xp, vp, rhoc, idx_b_r, int(size(vp, 2), 8), kin, kout), &
                  cube_ctx, 'kick')
  sim%dt_pm1 = kout%dt_pm1; sim%dt_pm2 = kout%dt_pm2; sim%dt_pm3 = kout%dt_pm3
  sim%dt_pp = kout%dt_pp; sim%dt_vmax = kout%dt_vmax; sim%vz_max = kout%vz_max
  napm = kout%napm; testrho = kout%testrho
endsubroutine

! drift.f90 (drift.f90:21-26, 85, 182, 202)
subroutine drift
  use variables
  use cube_gpu
  implicit none
  type(cube_drift_in) :: din
  type(cube_drift_out) :: dout
  din = cube_drift_in(real(dt_old, 8), real(dt, 8), sim%vz_max)   ! dt_mid and dt_mid*v in FP64
  call cube_check(cube_p3m_drift(cube_ctx, xp, vp, pid, rhoc, int(size(vp, 2), 8), idx_b_l, idx_b_r, &
                                 ppl0, pplr, pprl, ppr0, ppl, ppr, din, dout), cube_ctx, 'drift')
  sim%nplocal = dout%nplocal; overhead_tile = dout%overhead_tile
endsubroutine

! buffer_v after the kick (buffer_particle.f90:11-27, main.f90:38)
subroutine buffer_v
  use variables
  use cube_gpu
  implicit none
  call cube_check(cube_p3m_buffer_v(cube_ctx, vp, pid, rhoc, int(size(vp, 2), 8)), cube_ctx, 'buffer_v')
endsubroutine

! particle_mesh(rho1) for PM1 (particle_mesh.f90:1-222)
subroutine particle_mesh(rho1)
  use variables
  use cube_gpu
  implicit none
  real, intent(out) :: rho1(nc, nc, nc)
  type(cube_kick_in) :: kin
  type(cube_kick_out) :: kout
  kin = cube_kick_in(real(a_mid, 8), real(dt, 8), sim%mass_p_cdm, 1)
  call cube_check(cube_p3m_particle_mesh(cube_ctx, xp, vp, rhoc, idx_b_r, int(size(vp, 2), 8), kin, rho1, kout), &
                  cube_ctx, 'particle_mesh')
  testrho = kout%testrho
endsubroutine

program cube_ref_host
  use iso_c_binding
  use variables
  use cube_gpu
  implicit none
  character(512) :: dir, arg
  integer(8) :: np
  integer :: u
  real :: dt_drift
  call get_command_argument(1, dir)
  call get_command_argument(2, arg); read(arg, *) np
  call get_command_argument(3, arg); read(arg, *) a_mid
  call get_command_argument(4, arg); read(arg, *) dt
  call get_command_argument(5, arg); read(arg, *) dt_old
  call get_command_argument(6, arg); read(arg, *) dt_drift
  xp = 0; vp = 0; pid = 0
  open(newunit=u, file=trim(dir)//'/xp.bin', access='stream', form='unformatted', status='old')
  read(u) xp(:, 1:np); close(u)
  open(newunit=u, file=trim(dir)//'/vp.bin', access='stream', form='unformatted', status='old')
  read(u) vp(:, 1:np); close(u)
  open(newunit=u, file=trim(dir)//'/pid.bin', access='stream', form='unformatted', status='old')
  read(u) pid(1:np); close(u)
  open(newunit=u, file=trim(dir)//'/rhoc.bin', access='stream', form='unformatted', status='old')
  read(u) rhoc; close(u)
  open(newunit=u, file=trim(dir)//'/idx_b_r.bin', access='stream', form='unformatted', status='old')
  read(u) idx_b_r; close(u)
  call cube_gpu_init()
  ! main.f90:37-38 then the next step's drift (main.f90:34) with dt_old = dt
  call kick
  call buffer_v
  open(newunit=u, file=trim(dir)//'/ref_kick.bin', access='stream', form='unformatted', status='replace')
  write(u) sim%dt_pm1, sim%dt_pm2, sim%dt_pm3, sim%dt_pp, sim%dt_vmax, sim%vz_max, testrho, napm, vp(:, 1:np)
  close(u)
  dt_old = dt; dt = dt_drift
  call drift
  open(newunit=u, file=trim(dir)//'/ref_drift.bin', access='stream', form='unformatted', status='replace')
  write(u) sim%nplocal, xp(:, 1:sim%nplocal), vp(:, 1:sim%nplocal), pid(1:sim%nplocal), rhoc, idx_b_l, idx_b_r
  close(u)
  call cube_gpu_finalize()
end program
