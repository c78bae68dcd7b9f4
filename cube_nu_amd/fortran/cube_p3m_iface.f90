! cube_p3m_iface.f90 -- ISO_C_BINDING interface to libcube_p3m (include/cube_p3m.h).
!
! The derived types are bind(C) mirrors of the C structs; every entry point is
! an interface block with the C name.  Arrays are passed as assumed-size
! (by reference) and scalars by value, the convention of the reference's only
! C interop (fof.f90:304-310 <-> atomic_dsu.c:25-39).  The compressed
! coordinates xp are assumed-type (TYPE(*)) so integer(1) and integer(2)
! arrays pass unchanged.  Particle indices stay 1-based.
!
! cube_check turns a non-zero status into `error stop` with the library's
! message, which is how the reference reports the same conditions
! (drift.f90:64-73, 87-93, 124-129, 139-143).
module cube_p3m_iface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: CUBE_OK = 0, CUBE_ERR_ARG = 1, CUBE_ERR_HIP = 2, &
       CUBE_ERR_TILE_OVERFLOW = 3, CUBE_ERR_IMAGE_OVERFLOW = 4, CUBE_ERR_OUT_OF_RANGE = 5, &
       CUBE_ERR_COUNT_MISMATCH = 6, CUBE_ERR_PID = 7, CUBE_ERR_SPINE = 8, CUBE_ERR_NO_DEVICE = 9, &
       CUBE_ERR_STATE = 10

  type, bind(C) :: cube_p3m_config
    integer(c_int32_t) :: ng, nnt, nns, ratio_cs, ncb, ngb, izipx, precision
    real(c_float) :: box, apm1c, apm2, apm3f, app
    integer(c_int32_t) :: pm3, pp, pp_corr
    real(c_double) :: f_nu, mass_nu
    real(c_float) :: image_buffer, tile_buffer
    integer(c_int64_t) :: workspace_bytes
    integer(c_int32_t) :: ires_fixed   ! 0: adaptive sub-tile resolution; 2..5: fixed
  end type

  type, bind(C) :: cube_kick_in
    real(c_double) :: a_mid, dt        ! FP64 step coefficients (the module's real(4) values widen exactly)
    real(c_float) :: mass_p_cdm
    integer(c_int32_t) :: phase_mask
  end type

  type, bind(C) :: cube_kick_out
    real(c_float) :: dt_pm1, dt_pm2, dt_pm3, dt_pp, dt_vmax, vz_max
    real(c_float) :: f2max_pm1, f2max_pm2, f2max_pm3, f2max_pp
    real(c_float) :: vmax(3)
    integer(c_int32_t) :: napm(7)
    real(c_double) :: testrho
  end type

  type, bind(C) :: cube_drift_in
    real(c_double) :: dt_old, dt
    real(c_float) :: vz_max
  end type

  type, bind(C) :: cube_drift_out
    integer(c_int64_t) :: nplocal
    real(c_float) :: overhead_tile
    integer(c_int32_t) :: nlayer
  end type

  type, bind(C) :: cube_ic_params
    integer(c_int64_t) :: seed
    real(c_double) :: z, h, omega_m, omega_b, n_s, sigma8
    integer(c_int32_t) :: nic
    real(c_float) :: box
  end type

  interface
    integer(c_int) function cube_p3m_create(cfg, ctx) bind(C, name='cube_p3m_create')
      import :: c_int, c_ptr, cube_p3m_config
      type(cube_p3m_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function cube_p3m_destroy(ctx) bind(C, name='cube_p3m_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function cube_p3m_last_error(ctx) bind(C, name='cube_p3m_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function cube_p3m_version() bind(C, name='cube_p3m_version')
      import :: c_ptr
    end function

    integer(c_int) function cube_p3m_set_tf(ctx, level, tf, n) bind(C, name='cube_p3m_set_tf')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_float
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: level
      real(c_float), intent(in) :: tf(*)
      integer(c_int64_t), value :: n
    end function

    integer(c_int) function cube_p3m_get_tf(ctx, level, tf, n) bind(C, name='cube_p3m_get_tf')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_float
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: level
      real(c_float), intent(out) :: tf(*)
      integer(c_int64_t), value :: n
    end function

    integer(c_int) function cube_p3m_particle_mesh(ctx, xp, vp, rhoc, idx_b_r, np, kin, density, kout) &
        bind(C, name='cube_p3m_particle_mesh')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_float, cube_kick_in, cube_kick_out
      type(c_ptr), value :: ctx
      type(*), intent(in) :: xp(*)
      real(c_float), intent(inout) :: vp(*)
      integer(c_int32_t), intent(in) :: rhoc(*)
      integer(c_int64_t), intent(in) :: idx_b_r(*)
      integer(c_int64_t), value :: np
      type(cube_kick_in), intent(in) :: kin
      real(c_float), intent(out) :: density(*)
      type(cube_kick_out), intent(out) :: kout
    end function

    integer(c_int) function cube_p3m_kick(ctx, xp, vp, rhoc, idx_b_r, np, kin, kout) bind(C, name='cube_p3m_kick')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_float, cube_kick_in, cube_kick_out
      type(c_ptr), value :: ctx
      type(*), intent(in) :: xp(*)
      real(c_float), intent(inout) :: vp(*)
      integer(c_int32_t), intent(in) :: rhoc(*)
      integer(c_int64_t), intent(in) :: idx_b_r(*)
      integer(c_int64_t), value :: np
      type(cube_kick_in), intent(in) :: kin
      type(cube_kick_out), intent(out) :: kout
    end function

    integer(c_int) function cube_p3m_drift(ctx, xp, vp, pid, rhoc, np, idx_b_l, idx_b_r, ppl0, pplr, pprl, ppr0, &
                                           ppl, ppr, din, dout) bind(C, name='cube_p3m_drift')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_float, cube_drift_in, cube_drift_out
      type(c_ptr), value :: ctx
      type(*), intent(inout) :: xp(*)
      real(c_float), intent(inout) :: vp(*)
      integer(c_int64_t), intent(inout) :: pid(*)
      integer(c_int32_t), intent(inout) :: rhoc(*)
      integer(c_int64_t), value :: np
      integer(c_int64_t), intent(out) :: idx_b_l(*), idx_b_r(*), ppl0(*), pplr(*), pprl(*), ppr0(*), ppl(*), ppr(*)
      type(cube_drift_in), intent(in) :: din
      type(cube_drift_out), intent(out) :: dout
    end function

    integer(c_int) function cube_p3m_buffer(ctx, xp, vp, pid, rhoc, nplocal, np_cap, idx_b_l, idx_b_r, np_ext) &
        bind(C, name='cube_p3m_buffer')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_float
      type(c_ptr), value :: ctx
      type(*), intent(inout) :: xp(*)
      real(c_float), intent(inout) :: vp(*)
      integer(c_int64_t), intent(inout) :: pid(*)
      integer(c_int32_t), intent(inout) :: rhoc(*)
      integer(c_int64_t), value :: nplocal, np_cap
      integer(c_int64_t), intent(out) :: idx_b_l(*), idx_b_r(*)
      integer(c_int64_t), intent(out) :: np_ext
    end function

    ! buffer_v (buffer_particle.f90:11-27), after kick (main.f90:38)
    integer(c_int) function cube_p3m_buffer_v(ctx, vp, pid, rhoc, np) bind(C, name='cube_p3m_buffer_v')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(inout) :: vp(*)
      integer(c_int64_t), intent(inout) :: pid(*)
      integer(c_int32_t), intent(in) :: rhoc(*)
      integer(c_int64_t), value :: np
    end function

    integer(c_int) function cube_p3m_cicpower(ctx, nw, xi, nbin, delta, mean) bind(C, name='cube_p3m_cicpower')
      import :: c_int, c_ptr, c_int32_t, c_float, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nw, nbin
      real(c_float), intent(out) :: xi(*)
      type(c_ptr), value :: delta          ! real(4) delta(nw,nw,nw) via c_loc, or c_null_ptr
      real(c_double), intent(out) :: mean
    end function

    integer(c_int) function cube_p3m_synchronize(ctx) bind(C, name='cube_p3m_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
  end interface

contains

  ! NUL-terminated C string -> Fortran string
  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(:), allocatable :: s
    character(kind=c_char), pointer :: a(:)
    integer :: n
    interface
      integer(c_size_t) function strlen(q) bind(C, name='strlen')
        import :: c_size_t, c_ptr
        type(c_ptr), value :: q
      end function
    end interface
    if (.not. c_associated(p)) then
      s = ''
      return
    endif
    n = int(strlen(p))
    call c_f_pointer(p, a, [n])
    allocate(character(n) :: s)
    s = transfer(a(1:n), s)
  end function

  subroutine cube_check(st, ctx, where)
    integer(c_int), intent(in) :: st
    type(c_ptr), intent(in) :: ctx
    character(*), intent(in) :: where
    if (st /= CUBE_OK) then
      print*, where, ': cube_p3m status', st, ' ', c_string(cube_p3m_last_error(ctx))
      error stop
    endif
  end subroutine

  ! Defaults of parameters.f90:8-132 for a one-image run of ng^3 particles.
  function cube_default_config(ng, box) result(cfg)
    integer, intent(in) :: ng
    real, intent(in) :: box
    type(cube_p3m_config) :: cfg
    cfg%ng = ng; cfg%nnt = 4; cfg%nns = 4; cfg%ratio_cs = 4; cfg%ncb = 4; cfg%ngb = 16; cfg%izipx = 2
    cfg%precision = 32; cfg%box = box; cfg%apm1c = 3.5; cfg%apm2 = 3.5; cfg%apm3f = 3.5; cfg%app = 0.06
    cfg%pm3 = 1; cfg%pp = 1; cfg%pp_corr = 1; cfg%f_nu = 0; cfg%mass_nu = 0
    cfg%image_buffer = 2.0; cfg%tile_buffer = 6.0; cfg%workspace_bytes = 0; cfg%ires_fixed = 0
  end function
end module cube_p3m_iface
