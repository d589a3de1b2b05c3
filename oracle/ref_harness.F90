!> ref_harness.F90 -- TEST INFRASTRUCTURE ONLY.
!!
!! Drives the reference time steppers, compiled unmodified from
!! src/utilities/solver_library.F90 of the reference, with the oracle's
!! right-hand side (oc_rhs in sedcol_oracle.c).  The driver type extends the
!! reference's type_rhs_driver (solver_library.F90:37-49) and overrides
!! get_rhs (:48) the way fabm_sediment_driver.F90:575 does, so ode_solver
!! (:80-189) runs exactly as in the reference.  Built by oracle/Makefile into
!! oracle/_ref/ only; used by tests to pin the oracle's steppers.
module ref_harness
  use iso_c_binding
  use solver_library
  implicit none
  private

  integer, parameter :: rk = selected_real_kind(13)
  integer, parameter :: nv = 11

  type, extends(type_rhs_driver) :: type_oracle_driver
    type(c_ptr) :: params = c_null_ptr
    type(c_ptr) :: domain = c_null_ptr
    integer :: ncell = 0
    real(c_double), allocatable :: cpk(:), rpk(:)
  contains
    procedure :: get_rhs => oracle_get_rhs
  end type

  interface
    integer(c_int) function oc_rhs(p, d, conc, rhs, flux) bind(C, name='oc_rhs')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: p, d
      real(c_double) :: conc(*), rhs(*)
      type(c_ptr), value :: flux
    end function
  end interface

  public :: ref_ode_solver

contains

  ! conc(i,1,k,n) (Fortran layout, i = column) <-> packed x(n*ncell + i*K + k)
  subroutine oracle_get_rhs(rhs_driver, rhs)
    class(type_oracle_driver), intent(inout) :: rhs_driver
    real(rk), intent(inout), dimension(:,:,:,:), pointer :: rhs
    integer :: i, k, n, rc, nk
    nk = rhs_driver%knum
    do n = 1, rhs_driver%nvar
      do k = 1, nk
        do i = 1, rhs_driver%inum
          rhs_driver%cpk((n-1)*rhs_driver%ncell + (i-1)*nk + k) = rhs_driver%conc(i,1,k,n)
        end do
      end do
    end do
    rc = oc_rhs(rhs_driver%params, rhs_driver%domain, rhs_driver%cpk, rhs_driver%rpk, c_null_ptr)
    do n = 1, rhs_driver%nvar
      do k = 1, nk
        do i = 1, rhs_driver%inum
          rhs(i,1,k,n) = rhs_driver%rpk((n-1)*rhs_driver%ncell + (i-1)*nk + k)
        end do
      end do
    end do
  end subroutine

  !> nsteps calls of the reference ode_solver on ncol uniform columns of K
  !! layers.  conc is the packed oracle layout, updated in place.
  integer(c_int) function ref_ode_solver(params, domain, ncol, nk, conc, dt, method, &
      nsteps, dt_min, relative_change_min, last_min_dt) bind(C, name='ref_ode_solver')
    type(c_ptr), value :: params, domain
    integer(c_int), value :: ncol, nk, method, nsteps
    real(c_double) :: conc(*)
    real(c_double), value :: dt, dt_min, relative_change_min
    real(c_double) :: last_min_dt
    type(type_oracle_driver) :: drv
    real(rk), dimension(:,:,:,:), allocatable, target :: c4
    integer :: i, k, n, s, ncell

    ncell = ncol*nk
    allocate(c4(ncol,1,nk,nv))
    do n = 1, nv
      do k = 1, nk
        do i = 1, ncol
          c4(i,1,k,n) = conc((n-1)*ncell + (i-1)*nk + k)
        end do
      end do
    end do
    drv%inum = ncol
    drv%jnum = 1
    drv%knum = nk
    drv%nvar = nv
    drv%ncell = ncell
    drv%params = params
    drv%domain = domain
    drv%dt_min = dt_min
    drv%relative_change_min = relative_change_min
    drv%adaptive_solver_diagnostics = .true.
    drv%last_min_dt = last_min_dt
    allocate(drv%cpk(nv*ncell), drv%rpk(nv*ncell))
    drv%conc => c4
    do s = 1, nsteps
      call ode_solver(drv, dt, method)
    end do
    last_min_dt = drv%last_min_dt
    do n = 1, nv
      do k = 1, nk
        do i = 1, ncol
          conc((n-1)*ncell + (i-1)*nk + k) = c4(i,1,k,n)
        end do
      end do
    end do
    ref_ode_solver = 0
  end function

end module ref_harness
