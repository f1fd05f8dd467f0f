!
! drive_dropin.f90 -- calls the drop-in modules `advection` and `windadv` on
! the model's types the way the model's callers do (scalars.f90:153,
! tracers.f90:106, momentum.f90:187), on fields read from a stream file,
! and writes the results.
!
! usage: drive_dropin <input> <output>
!   input (stream, native endian):
!     nx ny nz scheme limiter (int32), dx dy dz dtl cdiag (real64),
!     fdz0 dzw dz1w dz1 (nz each),
!     u v w dens es qt q1 n1 q2 n2 a_dens a_es a_qt a_q1 a_n1 a_q2 a_n2 mu mv mw xx
!     (nx*ny*nz each, x fastest; a_* the initial tendency accumulators)
!   output (stream), nx*ny*nz real64 each:
!     1-7   advect(dtl, dens, state, hydro, densa, statea, hydroa): densa, es,
!           qt, q1, n1, q2, n2 accumulators (nhydro = 2, moments = 2)
!     8-14  the same with per-direction diagnostics on (es, qt, q1, q2)
!     15-26 diag(1..3)%pt, diag(1..3)%qt, diag(1..3)%qc, diag(1..3)%qr
!     27    advect(dtl, statel=, statea=, full=.false.): es accumulator
!     28    its qt accumulator (untouched)
!     29-32 advection_ac(wind_adv, xx, xs, iscal=1): xs, diag(1..3)%sca(:,:,:,1)
!     33-35 qnadv_uvw(winda, windl) with ntime = 1: winda%u, %v, %w
!     36-38 xs after advection_x('es', ...), then _y, then _z
!
program drive_dropin
  use gridno
  use shared_data
  use shared_wind
  use shared_diag
  use typedef_state
  use typedef_hydrometeor
  use typedef_wind
  use advection
  use windadv
  implicit none
  integer :: nx, ny, nz_, scheme, limiter, ios, d, h
  real(8) :: dtl
  real(8), allocatable :: dens(:,:,:), densa(:,:,:), xx(:,:,:), xs(:,:,:), dummy(:,:,:)
  real(8), allocatable :: a0(:,:,:,:)
  type(atm_state) :: st, sta
  type(hydrometeor) :: hy(2), hya(2)
  type(atm_winds) :: windl, winda
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old', iostat=ios)
  if (ios /= 0) error stop 2
  read(10) nx, ny, nz_, scheme, limiter
  read(10) dx, dy, dz, dtl, cdiag
  nz = nz_
  ip_start = 1; ip_end = nx; jp_start = 1; jp_end = ny
  it_start = 4; it_end = nx - 2
  jt_start = merge(4, 1, ny > 1); jt_end = merge(ny - 2, 1, ny > 1)
  allocate(fdz0(nz), dzw(nz), dz1w(nz), dz1(nz))
  read(10) fdz0, dzw, dz1w, dz1
  allocate(wind_adv%u(nx,ny,nz), wind_adv%v(nx,ny,nz), wind_adv%w(nx,ny,nz))
  allocate(dens(nx,ny,nz), st%es(nx,ny,nz), st%qt(nx,ny,nz))
  do h = 1, 2
    allocate(hy(h)%q(nx,ny,nz), hy(h)%n(nx,ny,nz))
  enddo
  allocate(a0(nx,ny,nz,7))
  allocate(windl%u(nx,ny,nz), windl%v(nx,ny,nz), windl%w(nx,ny,nz), xx(nx,ny,nz))
  read(10) wind_adv%u, wind_adv%v, wind_adv%w, dens, st%es, st%qt, hy(1)%q, hy(1)%n, hy(2)%q, hy(2)%n
  read(10) a0
  read(10) windl%u, windl%v, windl%w, xx
  close(10)

  select case (scheme)
  case (0)
    scal_adv = 'weno'
  case (1)
    scal_adv = 'ppm'
  case (4)
    scal_adv = 'quick'
  case (5)
    scal_adv = 'lw'
  case default
    scal_adv = 'muscl'
  end select
  select case (limiter)
  case (1)
    muscllimiter = 'MC'
  case default
    muscllimiter = 'minmod'
  end select
  nhydro = 2
  moments = 2
  lmicro = 0
  conservative = .false.
  limit = .true.

  allocate(densa(nx,ny,nz), sta%es(nx,ny,nz), sta%qt(nx,ny,nz), xs(nx,ny,nz), dummy(nx,ny,nz))
  do h = 1, 2
    allocate(hya(h)%q(nx,ny,nz), hya(h)%n(nx,ny,nz))
  enddo
  allocate(diag(3))
  do d = 1, 3
    allocate(diag(d)%pt(nx,ny,nz), diag(d)%qt(nx,ny,nz), diag(d)%qc(nx,ny,nz), diag(d)%qr(nx,ny,nz), &
             diag(d)%qi(nx,ny,nz), diag(d)%sca(nx,ny,nz,1))
    diag(d)%pt = 0.d0; diag(d)%qt = 0.d0; diag(d)%qc = 0.d0; diag(d)%qr = 0.d0; diag(d)%qi = 0.d0
    diag(d)%sca = 0.d0
  enddo

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  ! 1-7: the scalar tendencies of one stage (scalars.f90:153)
  call reset()
  call advect(dtl, dens, st, hy, densa, sta, hya)
  call dump()
  ! 8-26: the same with the advection diagnostics of es, qt, q1, q2
  out_diagt = .true.; out_diagq = .true.; out_diagl = .true.; out_diagr = .true.
  call reset()
  call advect(dtl, dens, st, hy, densa, sta, hya)
  call dump()
  do d = 1, 3
    write(11) diag(d)%pt
  enddo
  do d = 1, 3
    write(11) diag(d)%qt
  enddo
  do d = 1, 3
    write(11) diag(d)%qc
  enddo
  do d = 1, 3
    write(11) diag(d)%qr
  enddo
  out_diagt = .false.; out_diagq = .false.; out_diagl = .false.; out_diagr = .false.
  ! 27-28: state only, full = .false. (prognostic_pressure.f90:141 style)
  call reset()
  call advect(dtl, statel=st, statea=sta, full=.false.)
  write(11) sta%es, sta%qt
  ! 29-32: one tracer with its diagnostics (tracers.f90:106)
  out_diags = .true.
  dt0 = dtl
  xs = a0(:, :, :, 1)
  call advection_ac(wind_adv, xx, xs, iscal=1)
  write(11) xs
  do d = 1, 3
    write(11) diag(d)%sca(:, :, :, 1)
  enddo
  out_diags = .false.
  ! 33-35: momentum (momentum.f90:187), odd step
  ntime = 1
  allocate(winda%u(nx,ny,nz), winda%v(nx,ny,nz), winda%w(nx,ny,nz))
  call qnadv_uvw(winda, windl)
  write(11) winda%u, winda%v, winda%w
  ! 36-38: the line routines in sequence on one tendency
  xs = a0(:, :, :, 2)
  call advection_x('es', dtl, dummy, st%es, xs, wind_adv%u, .true.)
  write(11) xs
  call advection_y('es', dtl, dummy, st%es, xs, wind_adv%v, .true.)
  write(11) xs
  call advection_z('es', dtl, dummy, st%es, xs, wind_adv%w, .true.)
  write(11) xs
  close(11)

contains

  subroutine reset()
    densa = a0(:, :, :, 1)
    sta%es = a0(:, :, :, 2)
    sta%qt = a0(:, :, :, 3)
    hya(1)%q = a0(:, :, :, 4)
    hya(1)%n = a0(:, :, :, 5)
    hya(2)%q = a0(:, :, :, 6)
    hya(2)%n = a0(:, :, :, 7)
  end subroutine

  subroutine dump()
    write(11) densa, sta%es, sta%qt, hya(1)%q, hya(1)%n, hya(2)%q, hya(2)%n
  end subroutine

end program drive_dropin
