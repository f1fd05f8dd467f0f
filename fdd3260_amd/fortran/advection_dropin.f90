!
! advection_dropin.f90 -- MIMICA's advection interface on the model's own
! types, forwarding to libmimadv.so.
!
! Module `advection` keeps the public routines of src/TOOLS/advection.f90:79
! with their argument lists and meanings:
!   advect(dtl, dens, statel, hydromtrl, densa, statea, hydromtra, limiter, full)
!                                                   (advection.f90:84-314)
!   advection_x/y/z(name, dtl, xrho, xtmp, xs, u, lim, adv)  (:556, :637, :719)
!   advection_ac(windl, xx, xs, ilim, iscal, iaero)          (:801-915)
! and module `windadv` keeps qnadv_uvw(winda, windl) (src/INTEG/windadv.f90:46),
! so that callers (scalars.f90:153, tracers.f90:106, momentum.f90:187, ...)
! recompile unchanged.  Both USE the model's modules (gridno, shared_data,
! shared_wind, shared_diag and the type modules) for the globals the
! reference reads; muscllimiter and ppmlimiter, module variables of the
! replaced advection_muscl / advection_ppm (advq_muscl.f90:49,
! advq_ppm.f90:49), live here with their defaults.
!
! Semantics kept from the reference:
!   * with_dens when densa is present or in a conservative build, density
!     first, in flux form (adv = .false.); es, qt (with full) and, for
!     hydrometeors, q, n (moments == 2) and w (lmicro > 3 and h > ice) in the
!     form conservative selects (advection.f90:127-147, 338-362);
!   * every tendency accumulates xs = ((xs - X) - Y) - Z;
!   * per-direction diagnostics: diag(d)%pt, %qt, %qc, %qr, %qi for the
!     fields named es, qt, q1, q2, q3 (advection.f90:628-632), diag(d)%sca
!     and diag(d)%aero for advection_ac (:846-864), each d = 1, 2, 3 the
!     contribution of direction x, y, z.
! One deviation: the reference also sets with_dens in a conservative build
! without densa and then writes through the absent argument; here the
! density is advected only when dens and densa are both present.
!
module advection
  use iso_c_binding
  use gridno
  use shared_data
  use shared_wind
  use shared_diag
  use typedef_state
  use typedef_hydrometeor
  use typedef_wind
  use mimadv_advection, only: mimadv_init, field_addr, fields_tendency
  implicit none
  private
  public :: advect, advection_x, advection_y, advection_z, advection_ac, advection_ready
  public :: muscllimiter, ppmlimiter

  character(len=10), save :: muscllimiter = 'minmod'
  character(len=10), save :: ppmlimiter = 'PPM'
  logical, save :: lfirst = .true., lim_def = .true.

  integer, parameter :: MAXF = 64

  ! fields of one advect call: address of input and tendency, form, and the
  ! diagnostics slot of the field (0 none, 1 pt, 2 qt, 3 qc, 4 qr, 5 qi)
  type :: field_list
    integer :: n = 0
    type(c_ptr) :: q(MAXF), xs(MAXF)
    logical :: adv(MAXF)
    integer :: dslot(MAXF)
  end type field_list

contains

  ! advection_init (advection.f90:1045-1067) plus the context set-up
  subroutine advection_ready()
    if (.not. lfirst) return
    lim_def = limit
    if (.not. limit) ppmlimiter = 'no-limit'
    call mimadv_init(ip_end - ip_start + 1, jp_end - jp_start + 1, nz, dx, dy, dz, fdz0, dzw, dz1w, dz1, bcl, &
                     scal_adv, muscllimiter, ppmlimiter, lim_tol, mom_ord, lavisc, conservative, sorder=sorder, &
                     limit=limit)
    lfirst = .false.
  end subroutine

  integer function diag_slot(name)
    character(len=*), intent(in) :: name
    diag_slot = 0
    select case (trim(name))
    case ('es')
      if (out_diagt) diag_slot = 1
    case ('qt')
      if (out_diagq) diag_slot = 2
    case ('q1')
      if (out_diagl) diag_slot = 3
    case ('q2')
      if (out_diagr) diag_slot = 4
    case ('q3')
      if (out_diagi) diag_slot = 5
    end select
  end function

  subroutine add_field(fl, x, xs, adv, name)
    type(field_list), intent(inout) :: fl
    real(8), intent(in) :: x(:,:,:)
    real(8), intent(inout) :: xs(:,:,:)
    logical, intent(in) :: adv
    character(len=*), intent(in) :: name
    if (fl%n >= MAXF) error stop 'advection: too many fields in one call'
    fl%n = fl%n + 1
    fl%q(fl%n) = field_addr(x)
    fl%xs(fl%n) = field_addr(xs)
    fl%adv(fl%n) = adv
    fl%dslot(fl%n) = diag_slot(name)
  end subroutine

  ! xs = xs + part, diag(d)%<slot> = diag(d)%<slot> + cdiag*part, with
  ! part = -(direction d's divergence): the reference's xs - xx and
  ! diag - cdiag*xx, bit for bit
  subroutine apply_part(xs, part, slot, d)
    real(8), intent(inout) :: xs(:,:,:)
    real(8), intent(in) :: part(:,:,:)
    integer, intent(in) :: slot, d
    xs = xs + part
    select case (slot)
    case (1)
      diag(d)%pt = diag(d)%pt + cdiag * part
    case (2)
      diag(d)%qt = diag(d)%qt + cdiag * part
    case (3)
      diag(d)%qc = diag(d)%qc + cdiag * part
    case (4)
      diag(d)%qr = diag(d)%qr + cdiag * part
    case (5)
      diag(d)%qi = diag(d)%qi + cdiag * part
    end select
  end subroutine

  ! The tendencies of every field of the list in one library call for all
  ! fields and directions (no field feeds a diagnostic).
  subroutine run_fields(dtl, fl, lim)
    real(8), intent(in) :: dtl
    type(field_list), intent(inout) :: fl
    logical, intent(in) :: lim
    if (fl%n == 0) return
    call fields_tendency(dtl, wind_adv%u, wind_adv%v, wind_adv%w, fl%n, fl%q(1:fl%n), fl%xs(1:fl%n), &
                         fl%adv(1:fl%n), 7, lim, .false.)
  end subroutine

  subroutine advect(dtl, dens, statel, hydromtrl, densa, statea, hydromtra, limiter, full)
    real(8), intent(in) :: dtl
    real(8), dimension(ip_start:, jp_start:, :), optional, intent(in) :: dens
    type(atm_state), optional, intent(in) :: statel
    type(hydrometeor), dimension(:), optional, intent(in) :: hydromtrl
    real(8), dimension(ip_start:, jp_start:, :), optional, intent(inout) :: densa
    type(atm_state), optional, intent(inout) :: statea
    type(hydrometeor), dimension(:), optional, intent(inout) :: hydromtra
    logical, optional, intent(in) :: limiter, full
    type(field_list) :: fl
    logical :: lim, ful, with_dens, with_es, with_qt, with_hydro, adv
    integer :: h, d, f
    character(len=1) :: car1
    ! which tendency array each list entry accumulates into, for the
    ! diagnostics path: 0 densa, 1 es, 2 qt, 10*h + 1..3 hydrometeor q, n, w
    integer :: owner(MAXF)
    real(8), allocatable, target :: part(:,:,:,:)
    type(c_ptr) :: pp(MAXF)

    call advection_ready()
    lim = lim_def
    if (present(limiter)) lim = limiter
    ful = .true.
    if (present(full)) ful = full
    with_dens = present(dens) .and. present(densa)
    with_es = present(statel) .and. present(statea)
    with_qt = with_es .and. ful
    with_hydro = present(hydromtrl) .and. present(hydromtra) .and. nhydro > 0
    adv = .not. conservative
    ! density first, for mass consistency (advection.f90:338-340)
    if (with_dens) then
      call add_field(fl, dens, densa, .false., 'de')
      owner(fl%n) = 0
    endif
    if (with_es) then
      call add_field(fl, statel%es, statea%es, adv, 'es')
      owner(fl%n) = 1
    endif
    if (with_qt) then
      call add_field(fl, statel%qt, statea%qt, adv, 'qt')
      owner(fl%n) = 2
    endif
    if (with_hydro) then
      do h = 1, nhydro
        write(car1, '(i1)') h
        call add_field(fl, hydromtrl(h)%q, hydromtra(h)%q, adv, 'q'//car1)
        owner(fl%n) = 10 * h + 1
        if (moments == 2) then
          call add_field(fl, hydromtrl(h)%n, hydromtra(h)%n, adv, 'n'//car1)
          owner(fl%n) = 10 * h + 2
        endif
        if (lmicro > 3 .and. h > ice) then
          call add_field(fl, hydromtrl(h)%w, hydromtra(h)%w, adv, 'w'//car1)
          owner(fl%n) = 10 * h + 3
        endif
      enddo
    endif
    if (all(fl%dslot(1:fl%n) == 0)) then
      call run_fields(dtl, fl, lim)
      return
    endif
    ! some field feeds a diagnostic: one call per direction into
    ! temporaries, added in the reference's order
    allocate(part(ip_start:ip_end, jp_start:jp_end, 1:nz, fl%n))
    ! the library writes interior cells only: the ghosts of each temporary
    ! must hold zero, which the accumulators then receive unchanged
    part = 0.d0
    do f = 1, fl%n
      pp(f) = field_addr(part(:, :, :, f))
    enddo
    do d = 1, 3
      if (d == 2 .and. jp_end == jp_start) cycle  ! 2D build: no y pass
      call fields_tendency(dtl, wind_adv%u, wind_adv%v, wind_adv%w, fl%n, fl%q(1:fl%n), pp(1:fl%n), &
                           fl%adv(1:fl%n), ishft(1, d - 1), lim, .true.)
      do f = 1, fl%n
        call add_to_owner(f, part(:, :, :, f), d)
      enddo
    enddo
    deallocate(part)

  contains

    subroutine add_to_owner(f, part, d)
      integer, intent(in) :: f, d
      real(8), intent(in) :: part(:,:,:)
      integer :: o, hh
      o = owner(f)
      select case (o)
      case (0)
        call apply_part(densa, part, fl%dslot(f), d)
      case (1)
        call apply_part(statea%es, part, fl%dslot(f), d)
      case (2)
        call apply_part(statea%qt, part, fl%dslot(f), d)
      case default
        hh = o / 10
        select case (mod(o, 10))
        case (1)
          call apply_part(hydromtra(hh)%q, part, fl%dslot(f), d)
        case (2)
          call apply_part(hydromtra(hh)%n, part, fl%dslot(f), d)
        case (3)
          call apply_part(hydromtra(hh)%w, part, fl%dslot(f), d)
        end select
      end select
    end subroutine

  end subroutine advect

  ! advection_x/y/z (advection.f90:556-798): xs = xs - (divergence of xtmp
  ! along one direction by the face velocity vel); name routes the
  ! diagnostics; xrho is used only by the split form, which is not built
  subroutine advection_dir(d, name, dtl, xtmp, xs, vel, lim, adv)
    integer, intent(in) :: d
    character(len=2), intent(in) :: name
    real(8), intent(in) :: dtl
    real(8), dimension(ip_start:, jp_start:, :), intent(in) :: xtmp, vel
    real(8), dimension(ip_start:, jp_start:, :), intent(inout) :: xs
    logical, intent(in) :: lim
    logical, optional, intent(in) :: adv
    type(c_ptr) :: qp(1), xp(1)
    logical :: af(1)
    integer :: slot
    real(8), allocatable, target :: part(:,:,:)
    call advection_ready()
    af(1) = .not. conservative
    if (present(adv)) af(1) = adv
    qp(1) = field_addr(xtmp)
    slot = diag_slot(name)
    if (slot == 0) then
      xp(1) = field_addr(xs)
      call fields_tendency(dtl, vel, vel, vel, 1, qp, xp, af, ishft(1, d - 1), lim, .false.)
    else
      allocate(part(ip_start:ip_end, jp_start:jp_end, 1:nz))
      part = 0.d0  ! ghosts are not written by the library
      xp(1) = field_addr(part)
      call fields_tendency(dtl, vel, vel, vel, 1, qp, xp, af, ishft(1, d - 1), lim, .true.)
      call apply_part(xs, part, slot, d)
    endif
  end subroutine

  subroutine advection_x(name, dtl, xrho, xtmp, xs, u, lim, adv)
    character(len=2) :: name
    real(8) :: dtl
    real(8), dimension(ip_start:ip_end, jp_start:jp_end, 1:nz) :: xtmp, xs, xrho, u
    logical :: lim
    logical, optional :: adv
    call advection_dir(1, name, dtl, xtmp, xs, u, lim, adv)
  end subroutine

  subroutine advection_y(name, dtl, xrho, xtmp, xs, v, lim, adv)
    character(len=2) :: name
    real(8) :: dtl
    real(8), dimension(ip_start:ip_end, jp_start:jp_end, 1:nz) :: xtmp, xs, xrho, v
    logical :: lim
    logical, optional :: adv
    call advection_dir(2, name, dtl, xtmp, xs, v, lim, adv)
  end subroutine

  subroutine advection_z(name, dtl, xrho, xtmp, xs, w, lim, adv)
    character(len=2) :: name
    real(8) :: dtl
    real(8), dimension(ip_start:ip_end, jp_start:jp_end, 1:nz) :: xtmp, xs, xrho, w
    logical :: lim
    logical, optional :: adv
    call advection_dir(3, name, dtl, xtmp, xs, w, lim, adv)
  end subroutine

  ! advection_ac (advection.f90:801-915): one tracer by the winds windl with
  ! the full step dt0, xs = xs + xsx + xsy + xsz; diagnostics of tracer
  ! iscal in diag(d)%sca and of aerosol iaero (101..199 number, 201.. mass)
  ! in diag(d)%aero
  subroutine advection_ac(windl, xx, xs, ilim, iscal, iaero)
    type(atm_winds), intent(in) :: windl
    real(8), dimension(ip_start:, jp_start:, :), intent(inout) :: xs
    real(8), dimension(ip_start:, jp_start:, :), intent(in) :: xx
    logical, optional :: ilim
    integer, optional :: iscal, iaero
    logical :: lim, aerod, diags
    type(c_ptr) :: qp(1), xp(1)
    logical :: af(1)
    integer :: d, ia
    real(8), allocatable, target :: part(:,:,:)
    call advection_ready()
    lim = lim_def
    if (present(ilim)) lim = ilim
    aerod = .false.
    ia = 0
    if (present(iaero)) then
      ia = iaero
      if (iaero > 0) aerod = .true.
    endif
    diags = (out_diaga .and. aerod .and. ia > 100) .or. (out_diags .and. present(iscal))
    af(1) = .not. conservative
    qp(1) = field_addr(xx)
    if (.not. diags) then
      xp(1) = field_addr(xs)
      call fields_tendency(dt0, windl%u, windl%v, windl%w, 1, qp, xp, af, 7, lim, .false.)
      return
    endif
    allocate(part(ip_start:ip_end, jp_start:jp_end, 1:nz))
    part = 0.d0  ! ghosts are not written by the library
    xp(1) = field_addr(part)
    do d = 1, 3
      if (d == 2 .and. jp_end == jp_start) cycle
      call fields_tendency(dt0, windl%u, windl%v, windl%w, 1, qp, xp, af, ishft(1, d - 1), lim, .true.)
      xs = xs + part
      if (out_diaga .and. aerod .and. ia < 200 .and. ia > 100) &
        diag(d)%aero(ia - 100)%n = diag(d)%aero(ia - 100)%n + cdiag * part
      if (out_diaga .and. aerod .and. ia > 200) diag(d)%aero(ia - 200)%m = diag(d)%aero(ia - 200)%m + cdiag * part
      if (out_diags .and. present(iscal)) diag(d)%sca(:, :, :, iscal) = diag(d)%sca(:, :, :, iscal) + cdiag * part
    enddo
  end subroutine

end module advection

module windadv
  use gridno
  use shared_data, only: ntime
  use shared_wind
  use typedef_wind
  use advection, only: advection_ready
  use mimadv_advection, only: qnadv_uvw_arrays
  implicit none
  private
  public :: qnadv_uvw

contains

  ! qnadv_uvw (windadv.f90:46-79): winda OVERWRITTEN with +div of the
  ! momenta windl advected by wind_adv, directions in ntime-parity order
  subroutine qnadv_uvw(winda, windl)
    type(atm_winds), intent(inout) :: winda
    type(atm_winds), intent(in) :: windl
    call advection_ready()
    call qnadv_uvw_arrays(ntime, winda%u, winda%v, winda%w, windl%u, windl%v, windl%w, wind_adv%u, wind_adv%v, &
                          wind_adv%w)
  end subroutine

end module windadv
