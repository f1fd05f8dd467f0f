!
! drive_shim.f90 -- calls the drop-in advection routines as a model caller
! does, on fields read from a stream file, and writes the results.
!
! usage: drive_shim <input> <output>
!   input  (stream, native endian): nx ny nz scheme limiter limit sorder (int32),
!          dx dy dz dt0 (real64), fdz0 dzw dz1w dz1 (nz each),
!          u v w q mu mv mw (nx*ny*nz each, x fastest)
!   output (stream): xs after advection_x; xs after advection_x, _y, _z in
!          sequence; xs of advection_ac_arrays; ua va wa of qnadv_uvw_arrays;
!          xs of advection_ac_1d_arrays with w0(k) = -1e-3*k and dref =
!          den0(k) = 1.2 - 0.01*k (the conservative anelastic build);
!          with that density (avden0 = den0): the momenta of get_momentum_arrays
!          on u v w, get_advective_arrays of those momenta (u0shift 1.25,
!          v0shift -0.5), get_velocities_arrays of the momenta, and q after
!          get_conserved_array then after its inverse (get_scalars), with
!          dens(i,j,k) = den0(k) + 1e-3*mod(i+2j, 5)
!
program drive_shim
  use mimadv_advection
  implicit none
  integer :: nx, ny, nz, scheme, limiter, ilimit, sorder, ios
  logical :: lim
  real(8) :: dx, dy, dz, dt0
  real(8), allocatable :: fdz0(:), dzw(:), dz1w(:), dz1(:), w0(:), den0(:)
  integer :: i, j, k
  real(8), allocatable :: u(:,:,:), v(:,:,:), w(:,:,:), q(:,:,:), mu(:,:,:), mv(:,:,:), mw(:,:,:)
  real(8), allocatable :: xs(:,:,:), dummy(:,:,:), ua(:,:,:), va(:,:,:), wa(:,:,:), dens(:,:,:)
  character(len=512) :: fin, fout
  character(len=5) :: sname
  character(len=10) :: lname
  character(len=3) :: bcl(6)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(10, file=trim(fin), access='stream', form='unformatted', status='old', iostat=ios)
  if (ios /= 0) error stop 2
  read(10) nx, ny, nz, scheme, limiter, ilimit, sorder
  read(10) dx, dy, dz, dt0
  allocate(fdz0(nz), dzw(nz), dz1w(nz), dz1(nz))
  read(10) fdz0, dzw, dz1w, dz1
  allocate(u(nx,ny,nz), v(nx,ny,nz), w(nx,ny,nz), q(nx,ny,nz), mu(nx,ny,nz), mv(nx,ny,nz), mw(nx,ny,nz))
  read(10) u, v, w, q, mu, mv, mw
  close(10)

  select case (scheme)
  case (0)
    sname = 'weno'
  case (1)
    sname = 'ppm'
  case (4)
    sname = 'quick'
  case (5)
    sname = 'lw'
  case default
    sname = 'muscl'
  end select
  lim = ilimit /= 0
  select case (limiter)
  case (1)
    lname = 'MC'
  case (2)
    lname = 'VL'
  case (3)
    lname = 'Fromm'
  case default
    lname = 'minmod'
  end select
  bcl = (/ 'per', 'per', 'per', 'per', 'nnf', 'nnf' /)
  call mimadv_init(nx, ny, nz, dx, dy, dz, fdz0, dzw, dz1w, dz1, bcl, sname, lname, 'PPM', 1.d-9, 4, &
                   .true., .true., sorder=sorder, limit=lim)

  allocate(xs(nx,ny,nz), dummy(nx,ny,nz), ua(nx,ny,nz), va(nx,ny,nz), wa(nx,ny,nz))
  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  xs = 0.d0
  call advection_x('q1', dt0, dummy, q, xs, u, lim)
  write(11) xs
  call advection_y('q1', dt0, dummy, q, xs, v, lim)
  call advection_z('q1', dt0, dummy, q, xs, w, lim)
  write(11) xs
  xs = 0.d0
  call advection_ac_arrays(dt0, u, v, w, q, xs)
  write(11) xs
  call qnadv_uvw_arrays(2, ua, va, wa, mu, mv, mw, u, v, w)
  write(11) ua, va, wa
  allocate(w0(nz), den0(nz))
  do k = 1, nz
    w0(k) = -1.d-3 * k
    den0(k) = 1.2d0 - 0.01d0 * k
  enddo
  call mimadv_set_density(den0, den0)
  xs = 0.d0
  call advection_ac_1d_arrays(w0, q, xs)
  write(11) xs
  ua = u
  va = v
  wa = w
  call get_momentum_arrays(ua, va, wa)
  write(11) ua, va, wa
  call get_advective_arrays(1.25d0, -0.5d0, ua, va, wa, mu, mv, mw)
  write(11) mu, mv, mw
  call get_velocities_arrays(ua, va, wa)
  write(11) ua, va, wa
  allocate(dens(nx,ny,nz))
  do k = 1, nz
    do j = 1, ny
      do i = 1, nx
        dens(i,j,k) = den0(k) + 1.d-3 * mod(i + 2 * j, 5)
      enddo
    enddo
  enddo
  xs = q
  call get_conserved_array(dens, xs, .false.)
  write(11) xs
  call get_conserved_array(dens, xs, .true.)
  write(11) xs
  close(11)
  call mimadv_finalize()
end program drive_shim
