!> ref_gpu_driver.F90 -- TEST INFRASTRUCTURE ONLY.
!!
!! The reference's own time steppers (ode_solver, src/utilities/
!! solver_library.F90:80-189, compiled unmodified into oracle/_ref) driving
!! the GPU right-hand side.  type_sed_gpu is the extension a maintainer adds
!! inside the reference tree: it extends type_rhs_driver (:37-49) and
!! overrides get_rhs (:48) with the library's get_rhs, reached through the
!! ISO_C_BINDING module sedcol_hip, the way fabm_sediment_driver.F90:575
!! overrides it with the FABM one.  ode_solver then runs unchanged, one GPU
!! right-hand side per stage.
!!
!! Usage: ref_gpu_driver <input> <output>
!! input (stream access, native byte order): int32 ncol, nk, method, nsteps;
!! double dt; the sedcol_params struct; dz, dzc (distance to the layer above
!! at each layer, top unused), phi, phi_intf, bioturbation factor, each
!! (ncol, nk); bioturbation coefficient, temperature (ncol); bottom water and
!! top fluxes (ncol, nvar); conc (ncol, nk, nvar).
!! output: conc (ncol, nk, nvar) after nsteps ode_solver calls, last_min_dt,
!! fluxes (ncol, nvar) as left by get_rhs.
module ref_gpu
  use iso_c_binding
  use solver_library
  use sedcol_hip
  implicit none
  private

  type, extends(type_rhs_driver), public :: type_sed_gpu
    type(type_sed_hip) :: gpu
  contains
    procedure :: get_rhs => sed_gpu_get_rhs
  end type

contains

  subroutine sed_gpu_get_rhs(rhs_driver, rhs)
    class(type_sed_gpu), intent(inout) :: rhs_driver
    real(selected_real_kind(13)), dimension(:,:,:,:), pointer, intent(inout) :: rhs
    integer :: rc
    rhs_driver%gpu%conc => rhs_driver%conc
    call rhs_driver%gpu%get_rhs(rhs, rc)
    if (rc /= 0) then
      write(0,'(A,I0,A,A)') 'get_rhs failed with status ', rc, ': ', trim(sedcol_error_message())
      error stop 2
    end if
  end subroutine

end module ref_gpu

program ref_gpu_driver
  use iso_c_binding
  use solver_library
  use sedcol_hip
  use ref_gpu
  implicit none

  integer, parameter :: nv = 11
  type(type_sed_gpu) :: drv
  type(sedcol_params) :: p
  integer(c_int32_t) :: ncol, nk, method, nsteps
  real(c_double) :: dt
  real(c_double), allocatable :: dz(:,:), dzc(:,:), phi(:,:), phii(:,:), bfac(:,:), biot(:), temp(:)
  real(c_double), allocatable :: cbw(:,:), fup(:,:), conc(:,:,:)
  real(rk), dimension(:,:,:,:), allocatable, target :: c4
  real(rk), dimension(:,:,:), allocatable, target :: bdys, fluxes
  integer :: u, s, k, n, rc
  character(len=512) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) ncol, nk, method, nsteps
  read(u) dt
  read(u) p
  allocate(dz(ncol,nk), dzc(ncol,nk), phi(ncol,nk), phii(ncol,nk), bfac(ncol,nk), biot(ncol), temp(ncol))
  allocate(cbw(ncol,nv), fup(ncol,nv), conc(ncol,nk,nv))
  read(u) dz, dzc, phi, phii, bfac, biot, temp, cbw, fup, conc
  close(u)

  drv%gpu%params = p
  drv%gpu%dt_min = p%dt_min
  drv%gpu%relative_change_min = p%relative_change_min
  call drv%gpu%initialize(int(ncol), 1, int(nk), rc=rc)
  if (rc /= 0) call die('initialize', rc)
  rc = sedcol_set_geometry(drv%gpu%h, dz, dzc, phi, phii, bfac, biot)
  if (rc /= 0) call die('set_geometry', rc)

  allocate(c4(ncol,1,nk,nv), bdys(ncol,1,nv+1), fluxes(ncol,1,nv))
  do n = 1, nv
    do k = 1, nk
      c4(:,1,k,n) = conc(:,k,n)
    end do
    bdys(:,1,n+1) = cbw(:,n)
    fluxes(:,1,n) = fup(:,n)
  end do
  bdys(:,1,1) = temp
  drv%gpu%bdys => bdys
  drv%gpu%fluxes => fluxes

  ! the type_rhs_driver fields ode_solver reads (solver_library.F90:37-49)
  drv%inum = ncol
  drv%jnum = 1
  drv%knum = nk
  drv%nvar = nv
  drv%dt_min = p%dt_min
  drv%relative_change_min = p%relative_change_min
  drv%adaptive_solver_diagnostics = .true.
  drv%conc => c4
  do s = 1, nsteps
    call ode_solver(drv, real(dt, rk), int(method))
  end do

  do n = 1, nv
    do k = 1, nk
      conc(:,k,n) = c4(:,1,k,n)
    end do
  end do
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) conc
  write(u) real(drv%last_min_dt, c_double)
  write(u) fluxes
  close(u)
  call drv%gpu%finalize()
  write(*,'(A,I0,A,I0)') 'ref_gpu_driver: method ', method, ', steps ', nsteps

contains

  subroutine die(what, code)
    character(len=*), intent(in) :: what
    integer, intent(in) :: code
    write(0,'(A,A,I0,A,A)') what, ' failed with status ', code, ': ', trim(sedcol_error_message())
    error stop 1
  end subroutine

end program ref_gpu_driver
