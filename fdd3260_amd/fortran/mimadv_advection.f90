!
! mimadv_advection.f90 -- Fortran drop-in for MIMICA's advection interface.
!
! The reference module `advection` (src/TOOLS/advection.f90:79) exports
!   advect, advection_x, advection_y, advection_z, advection_ac, ...
! and module `windadv` (src/INTEG/windadv.f90:41) exports qnadv_uvw.  This
! module keeps advection_x/y/z with the reference's argument list and
! meaning (explicit-shape (ip_start:ip_end, jp_start:jp_end, 1:nz) REAL(8)
! arrays, xs accumulated with -div) and forwards them through iso_c_binding
! to libmimadv.so (include/mimadv.h).  The derived-type entry points of the
! reference (advect, advection_ac, qnadv_uvw on atm_state / atm_winds /
! hydrometeor) are thin loops over these array routines, given in
! INTEGRATION.md; the array forms below are what they call.
!
! mimadv_init replaces advection_init (advection.f90:1045-1067) and takes
! explicitly the module globals the reference reads (shared_data.f90:
! scal_adv, lim_tol, mom_ord, lavisc, bcl, dx, dy, dz, fdz0, dzw, dz1w,
! dz1; muscllimiter of advq_muscl.f90:49; ppmlimiter of advq_ppm.f90:49).
! A non-zero status from the library stops the run, as the reference stops
! in checkall (check.f90:59-101) through stop_mimica(1).
!
module mimadv_advection
  use iso_c_binding
  implicit none
  private

  public :: mimadv_init, mimadv_finalize
  public :: advection_x, advection_y, advection_z, advection_ac_arrays, qnadv_uvw_arrays
  public :: advection_ac_1d_arrays
  public :: mimadv_set_density, get_momentum_arrays, get_velocities_arrays, get_advective_arrays, get_conserved_array
  public :: field_addr, fields_tendency

  integer(c_int), parameter :: MIMADV_DIR_X = 1, MIMADV_DIR_Y = 2, MIMADV_DIR_Z = 4
  integer(c_int), parameter :: MIMADV_HOST_PTRS = 16
  integer(c_int), parameter :: MIMADV_UNLIMITED = 128, MIMADV_SORDER1 = 256, MIMADV_DREF_DEN0 = 512

  type, bind(c) :: mimadv_grid_t
    integer(c_int) :: nx, ny, nz
    real(c_double) :: dx, dy, dz
    type(c_ptr) :: fdz0, dzw, dz1w, dz1
    integer(c_int) :: bcl(6)
  end type

  type, bind(c) :: mimadv_opts_t
    integer(c_int) :: scheme, muscl_limiter, ppm_limiter
    real(c_double) :: lim_tol
    integer(c_int) :: weno_bottom, mom_ord, lavisc
  end type

  interface
    integer(c_int) function c_create(grid, opts, ctx) bind(c, name='mimadv_create')
      import :: c_int, c_ptr, mimadv_grid_t, mimadv_opts_t
      type(mimadv_grid_t), intent(in) :: grid
      type(mimadv_opts_t), intent(in) :: opts
      type(c_ptr), intent(out) :: ctx
    end function
    subroutine c_destroy(ctx) bind(c, name='mimadv_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function c_last_error(ctx) bind(c, name='mimadv_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function c_scalar_tendency(ctx, dt0, u, v, w, nf, q, xs, adv_form, flags) &
        bind(c, name='mimadv_scalar_tendency')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt0
      type(c_ptr), value :: u, v, w
      integer(c_int), value :: nf
      type(c_ptr), value :: q, xs, adv_form
      integer(c_int), value :: flags
    end function
    integer(c_int) function c_momentum_tendency(ctx, ntime, mu, mv, mw, ut, vt, wt, ua, va, wa, flags) &
        bind(c, name='mimadv_momentum_tendency')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ntime
      type(c_ptr), value :: mu, mv, mw, ut, vt, wt, ua, va, wa
      integer(c_int), value :: flags
    end function
    integer(c_int) function c_set_reference_density(ctx, den0, avden0) &
        bind(c, name='mimadv_set_reference_density')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, den0, avden0
    end function
    integer(c_int) function c_get_momentum(ctx, u, v, w, inverse, flags) bind(c, name='mimadv_get_momentum')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, u, v, w
      integer(c_int), value :: inverse, flags
    end function
    integer(c_int) function c_get_advective(ctx, u0shift, v0shift, u, v, w, ua, va, wa, flags) &
        bind(c, name='mimadv_get_advective')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: u0shift, v0shift
      type(c_ptr), value :: u, v, w, ua, va, wa
      integer(c_int), value :: flags
    end function
    integer(c_int) function c_advection_1d(ctx, w0, dens, nf, q, xs, flags) bind(c, name='mimadv_advection_1d')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, w0, dens
      integer(c_int), value :: nf
      type(c_ptr), value :: q, xs
      integer(c_int), value :: flags
    end function
    integer(c_int) function c_get_conserved(ctx, dens, nf, q, inverse, flags) bind(c, name='mimadv_get_conserved')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dens
      integer(c_int), value :: nf
      type(c_ptr), value :: q
      integer(c_int), value :: inverse, flags
    end function
  end interface

  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: g_nx = 0, g_ny = 0, g_nz = 0
  logical, save :: g_conservative = .false.
  logical, save :: g_lim_def = .true.
  integer, save :: g_sorder = 2
  real(c_double), allocatable, target, save :: m_fdz0(:), m_dzw(:), m_dz1w(:), m_dz1(:)

contains

  subroutine check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (status == 0) return
    call c_f_pointer(c_last_error(ctx), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    enddo
    write(*, '(a,a,a,i0,a,512a)') 'mimadv: ', what, ' failed (status ', status, '): ', msg(1:n)
    error stop 1
  end subroutine

  integer(c_int) function bc_code(s)
    character(len=*), intent(in) :: s
    select case (trim(s))
    case ('per')
      bc_code = 0
    case ('nnf')
      bc_code = 1
    case default
      bc_code = 2
    end select
  end function

  ! One-time set-up (advection_init plus the module globals it relies on).
  ! sorder (default 2) and limit (default .true., the lim_def that
  ! advection_ac uses when ilim is absent) are the namelist values of
  ! cm_num (mimica.f90:74).
  subroutine mimadv_init(nx, ny, nz, dx, dy, dz, fdz0, dzw, dz1w, dz1, bcl, scal_adv, muscllimiter, &
                         ppmlimiter, lim_tol, mom_ord, lavisc, conservative, sorder, limit)
    integer, intent(in) :: nx, ny, nz, mom_ord
    real(8), intent(in) :: dx, dy, dz, lim_tol
    real(8), intent(in) :: fdz0(nz), dzw(nz), dz1w(nz), dz1(nz)
    character(len=3), intent(in) :: bcl(6)
    character(len=*), intent(in) :: scal_adv, muscllimiter, ppmlimiter
    logical, intent(in) :: lavisc, conservative
    integer, optional, intent(in) :: sorder
    logical, optional, intent(in) :: limit
    type(mimadv_grid_t) :: g
    type(mimadv_opts_t) :: o
    integer :: b
    if (c_associated(ctx)) call mimadv_finalize()
    allocate(m_fdz0(nz), m_dzw(nz), m_dz1w(nz), m_dz1(nz))
    m_fdz0 = fdz0; m_dzw = dzw; m_dz1w = dz1w; m_dz1 = dz1
    g%nx = nx; g%ny = ny; g%nz = nz
    g%dx = dx; g%dy = dy; g%dz = dz
    g%fdz0 = c_loc(m_fdz0); g%dzw = c_loc(m_dzw); g%dz1w = c_loc(m_dz1w); g%dz1 = c_loc(m_dz1)
    do b = 1, 6
      g%bcl(b) = bc_code(bcl(b))
    enddo
    select case (trim(scal_adv))
    case ('weno')
      o%scheme = 0
    case ('ppm')
      o%scheme = 1
    case ('muscl')
      o%scheme = 2
    case ('quick')
      o%scheme = 4
    case default
      ! every other value runs Lax-Wendroff (advection.f90:614-616)
      o%scheme = 5
    end select
    select case (trim(muscllimiter))
    case ('MC')
      o%muscl_limiter = 1
    case ('VL')
      o%muscl_limiter = 2
    case ('minmod')
      o%muscl_limiter = 0
    case default
      o%muscl_limiter = 3
    end select
    o%ppm_limiter = merge(1, 0, trim(ppmlimiter) == 'PPM')
    ! advection_init: limit = .false. also sets ppmlimiter = 'no-limit'
    ! (advection.f90:1056)
    if (present(limit)) then
      if (.not. limit) o%ppm_limiter = 0
    endif
    o%lim_tol = lim_tol
    o%weno_bottom = 0
    o%mom_ord = mom_ord
    o%lavisc = merge(1, 0, lavisc)
    g_nx = nx; g_ny = ny; g_nz = nz
    g_conservative = conservative
    g_sorder = 2
    if (present(sorder)) g_sorder = sorder
    g_lim_def = .true.
    if (present(limit)) g_lim_def = limit
    call check(c_create(g, o, ctx), 'mimadv_create')
  end subroutine

  subroutine mimadv_finalize()
    if (c_associated(ctx)) call c_destroy(ctx)
    ctx = c_null_ptr
    if (allocated(m_fdz0)) deallocate(m_fdz0, m_dzw, m_dz1w, m_dz1)
  end subroutine

  ! the line-kernel selection of advection_{x,y,z} (advection.f90:580-620)
  integer(c_int) function line_flags(lim)
    logical, intent(in) :: lim
    line_flags = 0
    if (.not. lim) line_flags = MIMADV_UNLIMITED
    if (g_sorder == 1) line_flags = ior(line_flags, MIMADV_SORDER1)
  end function

  subroutine scalar_dir(dirflag, dtl, xtmp, xs, vel, lim, adv)
    integer(c_int), intent(in) :: dirflag
    real(8), intent(in) :: dtl
    real(8), target, intent(in) :: xtmp(g_nx, g_ny, g_nz), vel(g_nx, g_ny, g_nz)
    real(8), target, intent(inout) :: xs(g_nx, g_ny, g_nz)
    logical, intent(in) :: lim
    logical, optional, intent(in) :: adv
    type(c_ptr), target :: qp(1), xp(1)
    integer(c_int), target :: af(1)
    type(c_ptr) :: pu, pv, pw
    af(1) = merge(1, 0, .not. g_conservative)
    if (present(adv)) af(1) = merge(1, 0, adv)
    qp(1) = c_loc(xtmp)
    xp(1) = c_loc(xs)
    pu = c_null_ptr; pv = c_null_ptr; pw = c_null_ptr
    if (dirflag == MIMADV_DIR_X) pu = c_loc(vel)
    if (dirflag == MIMADV_DIR_Y) pv = c_loc(vel)
    if (dirflag == MIMADV_DIR_Z) pw = c_loc(vel)
    call check(c_scalar_tendency(ctx, dtl, pu, pv, pw, 1_c_int, c_loc(qp), c_loc(xp), c_loc(af), &
                                 ior(ior(dirflag, MIMADV_HOST_PTRS), line_flags(lim))), 'advection')
  end subroutine

  ! Address of a whole field (the actual argument must be contiguous: an
  ! allocatable array or an array component of one).
  type(c_ptr) function field_addr(x)
    real(8), target, intent(in) :: x(g_nx * g_ny * g_nz)
    field_addr = c_loc(x)
  end function

  ! nf fields in one call: xs_f = ((xs_f - X_f) - Y_f) - Z_f over the
  ! directions of dirmask (overwrite: xs_f = ((0 - X_f) ...)), host arrays
  ! given by address; adv(f) the advective form of field f.  u, v, w are
  ! the face velocities of the directions used (the others may be any
  ! field of the grid's shape).
  subroutine fields_tendency(dtl, u, v, w, nf, q, xs, adv, dirmask, lim, overwrite)
    real(8), intent(in) :: dtl
    real(8), target, intent(in) :: u(g_nx, g_ny, g_nz), v(g_nx, g_ny, g_nz), w(g_nx, g_ny, g_nz)
    integer, intent(in) :: nf, dirmask
    type(c_ptr), target, intent(in) :: q(nf), xs(nf)
    logical, intent(in) :: adv(nf), lim, overwrite
    integer(c_int), target :: af(max(nf, 1))
    integer(c_int) :: flags
    integer :: f
    if (nf <= 0) return
    do f = 1, nf
      af(f) = merge(1, 0, adv(f))
    enddo
    flags = ior(ior(int(dirmask, c_int), MIMADV_HOST_PTRS), line_flags(lim))
    if (overwrite) flags = ior(flags, 8_c_int)
    call check(c_scalar_tendency(ctx, dtl, c_loc(u), c_loc(v), c_loc(w), int(nf, c_int), c_loc(q), c_loc(xs), &
                                 c_loc(af), flags), 'advection')
  end subroutine

  ! advection_x (advection.f90:556): xs = xs - (x flux divergence of xtmp)
  subroutine advection_x(name, dtl, xrho, xtmp, xs, u, lim, adv)
    character(len=2), intent(in) :: name
    real(8), intent(in) :: dtl
    real(8), intent(inout) :: xrho(g_nx, g_ny, g_nz)
    real(8), intent(in) :: xtmp(g_nx, g_ny, g_nz), u(g_nx, g_ny, g_nz)
    real(8), intent(inout) :: xs(g_nx, g_ny, g_nz)
    logical, intent(in) :: lim
    logical, optional, intent(in) :: adv
    call scalar_dir(MIMADV_DIR_X, dtl, xtmp, xs, u, lim, adv)
  end subroutine

  ! advection_y (advection.f90:637)
  subroutine advection_y(name, dtl, xrho, xtmp, xs, v, lim, adv)
    character(len=2), intent(in) :: name
    real(8), intent(in) :: dtl
    real(8), intent(inout) :: xrho(g_nx, g_ny, g_nz)
    real(8), intent(in) :: xtmp(g_nx, g_ny, g_nz), v(g_nx, g_ny, g_nz)
    real(8), intent(inout) :: xs(g_nx, g_ny, g_nz)
    logical, intent(in) :: lim
    logical, optional, intent(in) :: adv
    call scalar_dir(MIMADV_DIR_Y, dtl, xtmp, xs, v, lim, adv)
  end subroutine

  ! advection_z (advection.f90:719)
  subroutine advection_z(name, dtl, xrho, xtmp, xs, w, lim, adv)
    character(len=2), intent(in) :: name
    real(8), intent(in) :: dtl
    real(8), intent(inout) :: xrho(g_nx, g_ny, g_nz)
    real(8), intent(in) :: xtmp(g_nx, g_ny, g_nz), w(g_nx, g_ny, g_nz)
    real(8), intent(inout) :: xs(g_nx, g_ny, g_nz)
    logical, intent(in) :: lim
    logical, optional, intent(in) :: adv
    call scalar_dir(MIMADV_DIR_Z, dtl, xtmp, xs, w, lim, adv)
  end subroutine

  ! advection_ac (advection.f90:801) on the wind components: xs = xs + xsx +
  ! xsy + xsz with dt0, all three directions in one call
  subroutine advection_ac_arrays(dt0, u, v, w, xx, xs, ilim)
    real(8), intent(in) :: dt0
    real(8), target, intent(in) :: u(g_nx, g_ny, g_nz), v(g_nx, g_ny, g_nz), w(g_nx, g_ny, g_nz)
    real(8), target, intent(in) :: xx(g_nx, g_ny, g_nz)
    real(8), target, intent(inout) :: xs(g_nx, g_ny, g_nz)
    logical, optional, intent(in) :: ilim
    type(c_ptr), target :: qp(1), xp(1)
    integer(c_int), target :: af(1)
    logical :: lim
    lim = g_lim_def
    if (present(ilim)) lim = ilim
    af(1) = merge(1, 0, .not. g_conservative)
    qp(1) = c_loc(xx)
    xp(1) = c_loc(xs)
    call check(c_scalar_tendency(ctx, dt0, c_loc(u), c_loc(v), c_loc(w), 1_c_int, c_loc(qp), c_loc(xp), &
                                 c_loc(af), ior(ior(7_c_int, MIMADV_HOST_PTRS), line_flags(lim))), 'advection_ac')
  end subroutine

  ! qnadv_uvw (windadv.f90:46) on arrays: (ua, va, wa) OVERWRITTEN with
  ! +div of the momenta (mu, mv, mw) advected by (ut, vt, wt)
  subroutine qnadv_uvw_arrays(ntime, ua, va, wa, mu, mv, mw, ut, vt, wt)
    integer, intent(in) :: ntime
    real(8), target, intent(inout) :: ua(g_nx, g_ny, g_nz), va(g_nx, g_ny, g_nz), wa(g_nx, g_ny, g_nz)
    real(8), target, intent(in) :: mu(g_nx, g_ny, g_nz), mv(g_nx, g_ny, g_nz), mw(g_nx, g_ny, g_nz)
    real(8), target, intent(in) :: ut(g_nx, g_ny, g_nz), vt(g_nx, g_ny, g_nz), wt(g_nx, g_ny, g_nz)
    call check(c_momentum_tendency(ctx, int(ntime, c_int), c_loc(mu), c_loc(mv), c_loc(mw), c_loc(ut), &
                                   c_loc(vt), c_loc(wt), c_loc(ua), c_loc(va), c_loc(wa), MIMADV_HOST_PTRS), &
               'qnadv_uvw')
  end subroutine

  ! advection_ac_1d (advection.f90:993-1042) on arrays: xs = xs -
  ! dref*w0*grad_1d(xx/dref) on the interior columns.  dref is the 3D
  ! density dens when present (pressure%dens of the compressible
  ! CONSERVATIVE build), den0 of mimadv_set_density in a conservative build
  ! (CONSERVATIVE + ANELASTIC), 1 otherwise.  advection_1d (:918-990) is this
  ! call on state%es (and state%qt with with_lssub).
  subroutine advection_ac_1d_arrays(w0, xx, xs, dens)
    real(8), target, intent(in) :: w0(g_nz)
    real(8), target, intent(in) :: xx(g_nx, g_ny, g_nz)
    real(8), target, intent(inout) :: xs(g_nx, g_ny, g_nz)
    real(8), target, optional, intent(in) :: dens(g_nx, g_ny, g_nz)
    type(c_ptr), target :: qp(1), xp(1)
    type(c_ptr) :: pd
    integer(c_int) :: flags
    qp(1) = c_loc(xx)
    xp(1) = c_loc(xs)
    pd = c_null_ptr
    flags = MIMADV_HOST_PTRS
    if (present(dens)) then
      pd = c_loc(dens)
    else if (g_conservative) then
      flags = ior(flags, MIMADV_DREF_DEN0)
    endif
    call check(c_advection_1d(ctx, c_loc(w0), pd, 1_c_int, c_loc(qp), c_loc(xp), flags), 'advection_ac_1d')
  end subroutine

  ! Reference density profiles den0, avden0 (shared_data globals set in
  ! initial.f90:175); call once after mimadv_init
  subroutine mimadv_set_density(den0, avden0)
    real(8), target, intent(in) :: den0(g_nz), avden0(g_nz)
    call check(c_set_reference_density(ctx, c_loc(den0), c_loc(avden0)), 'set_reference_density')
  end subroutine

  ! get_momentum / get_velocities (funcpack.f90:99-316), anelastic branch, in place
  subroutine get_momentum_arrays(u, v, w)
    real(8), target, intent(inout) :: u(g_nx, g_ny, g_nz), v(g_nx, g_ny, g_nz), w(g_nx, g_ny, g_nz)
    call check(c_get_momentum(ctx, c_loc(u), c_loc(v), c_loc(w), 0_c_int, MIMADV_HOST_PTRS), 'get_momentum')
  end subroutine

  subroutine get_velocities_arrays(u, v, w)
    real(8), target, intent(inout) :: u(g_nx, g_ny, g_nz), v(g_nx, g_ny, g_nz), w(g_nx, g_ny, g_nz)
    call check(c_get_momentum(ctx, c_loc(u), c_loc(v), c_loc(w), 1_c_int, MIMADV_HOST_PTRS), 'get_velocities')
  end subroutine

  ! get_advective (funcpack.f90:65-92): wind_adv from the momenta
  subroutine get_advective_arrays(u0shift, v0shift, u, v, w, ua, va, wa)
    real(8), intent(in) :: u0shift, v0shift
    real(8), target, intent(in) :: u(g_nx, g_ny, g_nz), v(g_nx, g_ny, g_nz), w(g_nx, g_ny, g_nz)
    real(8), target, intent(out) :: ua(g_nx, g_ny, g_nz), va(g_nx, g_ny, g_nz), wa(g_nx, g_ny, g_nz)
    call check(c_get_advective(ctx, u0shift, v0shift, c_loc(u), c_loc(v), c_loc(w), c_loc(ua), c_loc(va), &
                               c_loc(wa), MIMADV_HOST_PTRS), 'get_advective')
  end subroutine

  ! get_conserved / get_scalars (funcpack.f90:397-514) on one array
  subroutine get_conserved_array(dens, q, inverse)
    real(8), target, intent(in) :: dens(g_nx, g_ny, g_nz)
    real(8), target, intent(inout) :: q(g_nx, g_ny, g_nz)
    logical, intent(in) :: inverse
    type(c_ptr), target :: qp(1)
    qp(1) = c_loc(q)
    call check(c_get_conserved(ctx, c_loc(dens), 1_c_int, c_loc(qp), merge(1_c_int, 0_c_int, inverse), &
                               MIMADV_HOST_PTRS), merge('get_scalars  ', 'get_conserved', inverse))
  end subroutine

end module mimadv_advection
