!
! model_standins.f90 -- the MIMICA modules the drop-in advection interface
! USEs, reduced to the names it reads, so that the drop-in builds and is
! tested outside the model.  In the model build these modules come from the
! model itself and this file is not compiled:
!   gridno              INCLUDE/wgrid.csh:4-68 (generated; ip/jp/it/jt bounds, nz, ice)
!   shared_data         src/TYPES/shared_data.f90 (namelist and run globals)
!   typedef_state       src/TYPES/typedef_state.f90:41-46 (atm_state)
!   typedef_wind        src/TYPES/typedef_wind.f90:40-45 (atm_winds)
!   typedef_hydrometeor src/TYPES/typedef_hydrometeor.f90:42-51 (hydrometeor)
!   shared_wind         src/TYPES/shared_wind.f90:44 (wind_adv)
!   typedef_diag, shared_diag  src/TYPES/typedef_diag.f90:53-83, shared_diag.f90:61 (diag)
! Component names and meanings are the model's; the generated constants of
! gridno are variables here so a driver can size them at run time.
!
module gridno
  implicit none
  integer :: ip_start = 1, ip_end = 1, jp_start = 1, jp_end = 1
  integer :: it_start = 4, it_end = 1, jt_start = 1, jt_end = 1
  integer :: nz = 1
  integer, parameter :: ice = 3  ! index of the ice hydrometeor (wgrid.csh:59-60)
end module gridno

module shared_data
  implicit none
  logical :: limit = .true.          ! flux limiter for scalar advection (:251)
  logical :: conservative = .false.  ! flux form (:265)
  logical :: lavisc = .true.         ! momentum artificial viscosity (:256)
  integer :: ntime = 0               ! time step counter (:285)
  integer :: lmicro = 0              ! microphysics level (:291)
  integer :: sorder = 2              ! integration order (:299)
  integer :: mom_ord = 4             ! momentum advection order (:300)
  integer :: moments = 1             ! moments of the bulk scheme (:317)
  integer :: nhydro = 0              ! number of hydrometeors (:52)
  character(len=5) :: scal_adv = 'weno'  ! scalar scheme (:320)
  character(len=3) :: bcl(6) = (/ 'per', 'per', 'per', 'per', 'nnf', 'nnf' /)  ! (:325)
  real(8) :: dt0 = 1.d0              ! basic time step (:334)
  real(8) :: lim_tol = 1.d-9         ! limiter tolerance (:348)
  real(8) :: dx = 1.d0, dy = 1.d0, dz = 1.d0  ! (:349)
  real(8) :: cdiag = 1.d0            ! fraction of tendencies in diag (:400)
  logical :: out_diagu = .false., out_diagv = .false., out_diagw = .false.
  logical :: out_diagt = .false., out_diagq = .false., out_diagl = .false.
  logical :: out_diagr = .false., out_diagi = .false., out_diaga = .false., out_diags = .false.
  real(8), allocatable :: fdz0(:), dzw(:), dz1w(:), dz1(:)  ! vertical metrics (:525-526)
  real(8), allocatable :: den0(:), avden0(:)                ! anelastic reference density
end module shared_data

module typedef_state
  implicit none
  type :: atm_state
    real(8), dimension(:,:,:), allocatable :: es     ! potential temperature
    real(8), dimension(:,:,:), allocatable :: qt     ! total water mixing ratio
    real(8), dimension(:,:,:,:), allocatable :: scal ! additional tracers
  end type atm_state
end module typedef_state

module typedef_wind
  implicit none
  type :: atm_winds
    real(8), dimension(:,:,:), allocatable :: u, v, w
  end type atm_winds
end module typedef_wind

module typedef_hydrometeor
  implicit none
  type :: hydrometeor
    real(8), dimension(:,:,:), allocatable :: q   ! mixing ratio
    real(8), dimension(:,:,:), allocatable :: n   ! number concentration
    real(8), dimension(:,:,:), allocatable :: w   ! liquid water mass
    real(8), dimension(:,:,:), allocatable :: d   ! mean size
    real(8), dimension(:,:,:), allocatable :: vp  ! mass-weighted fall speed
  end type hydrometeor
end module typedef_hydrometeor

module shared_wind
  use typedef_wind
  implicit none
  type(atm_winds) :: wind_adv   ! advecting velocity of the stage
end module shared_wind

module typedef_diag
  implicit none
  type :: aerodiag
    real(8), dimension(:,:,:), allocatable :: n, m
  end type aerodiag
  type :: tendencies
    real(8), dimension(:,:,:), allocatable :: u, v, w, qt, pt, qc, qr, qi
    real(8), dimension(:,:,:,:), allocatable :: sca
    type(aerodiag), dimension(:), allocatable :: aero
  end type tendencies
end module typedef_diag

module shared_diag
  use typedef_diag
  implicit none
  type(tendencies), dimension(:), allocatable :: diag  ! diag(1..3): x, y, z contributions
end module shared_diag
