!> sedcol_hip.F90 -- ISO_C_BINDING interface of libsedcol.so for Fortran hosts.
!!
!! The C functions are declared in include/sedcol.h.  type_sed_hip carries the
!! same fields a caller of the reference fills in type_rhs_driver / type_sed
!! (src/utilities/solver_library.F90:37-49, src/drivers/
!! fabm_sediment_driver.F90:69-113): inum/jnum/knum/nvar, conc(i,j,k,n),
!! bdys(i,j,1:nvar+1), fluxes(i,j,1:nvar), dt_min, relative_change_min,
!! last_min_dt.  ode_solver_hip has the argument list of ode_solver
!! (solver_library.F90:80) plus an optional status code, so a call
!!     call ode_solver(sed, dt, ode_method)
!! becomes
!!     call ode_solver_hip(sed_hip, dt, ode_method)
!! with conc, bdys and fluxes read from and written back to the same arrays.
module sedcol_hip

  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: rk = selected_real_kind(13)

  integer(c_int32_t), parameter, public :: SEDCOL_NVAR = 11
  integer(c_int32_t), parameter, public :: SEDCOL_EULER = 0, SEDCOL_RUNGE_KUTTA_4 = 1, &
    SEDCOL_ADAPTIVE_EULER = 2, SEDCOL_RUNGE_KUTTA_4_38 = 3, SEDCOL_EMP1 = 4, SEDCOL_EMP2 = 5, &
    SEDCOL_EMP_ADAPTIVE = 6, SEDCOL_ROS2_ADAPTIVE = 7
  integer(c_int32_t), parameter, public :: SEDCOL_LAYOUT_FORTRAN = 0, SEDCOL_LAYOUT_PACKED = 1
  integer(c_int32_t), parameter, public :: SEDCOL_ARITH_REFERENCE = 0, SEDCOL_ARITH_FAST = 1

  type, bind(C), public :: sedcol_desc
    integer(c_int32_t) :: ncol, kmax, nvar, layout
    type(c_ptr) :: nlayers = c_null_ptr
    type(c_ptr) :: mask = c_null_ptr
  end type

  type, bind(C), public :: sedcol_params
    real(c_double) :: D0 = 0.9d0, dt_min = 1.d-9, relative_change_min = -0.9d0
    real(c_double) :: rL = 0.043d0/86400.d0, rS = 0.001d0/86400.d0, CprodMax = 9600.d0/86400.d0
    real(c_double) :: rnit = 200.d0/86400.d0, ksO2nitri = 20.d0, rODUox = 20.d0/86400.d0
    real(c_double) :: ksO2oduox = 1.d0, ksO2oxic = 3.d0, ksNO3denit = 1.d0, kinO2denit = 70.d0
    real(c_double) :: kinNO3anox = 1.d0, kinO2anox = 1.d0
    real(c_double) :: rPads = (0.01d0*20.d0)/86400.d0, PAdsODU = 70.d0
    real(c_double) :: emp_rtol = 1.d-6, emp_atol = 1.d-6, emp_dt_min = 1.d0, emp_dt_max = 3600.d0
    integer(c_int32_t) :: bcup_dis = 2, bcup_part = 1, bcdown = 3, clamp = 1
    integer(c_int32_t) :: arith = 0, emp_weights = 0, adaptive_method = 0
  end type

  interface
    integer(c_int) function sedcol_create(desc, h) bind(C, name='sedcol_create')
      import :: c_int, c_ptr, sedcol_desc
      type(sedcol_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function sedcol_destroy(h) bind(C, name='sedcol_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function sedcol_last_error() bind(C, name='sedcol_last_error')
      import :: c_ptr
    end function
    integer(c_int) function sedcol_set_params(h, p) bind(C, name='sedcol_set_params')
      import :: c_int, c_ptr, sedcol_params
      type(c_ptr), value :: h
      type(sedcol_params), intent(in) :: p
    end function
    integer(c_int) function sedcol_set_geometry(h, dz, dzc, phi, phi_intf, bfac, biot) &
        bind(C, name='sedcol_set_geometry')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: dz(*), dzc(*), phi(*), phi_intf(*), bfac(*), biot(*)
    end function
    integer(c_int) function sedcol_set_forcing(h, temp, c_bw, flux_top) bind(C, name='sedcol_set_forcing')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: temp(*), c_bw(*), flux_top(*)
    end function
    ! bioturbation profile 3 (fabm_sediment_driver.F90:618-645); zw = (k_l, L1, L2, beta, b, dry density)
    integer(c_int) function sedcol_set_bioturbation_zw(h, zw) bind(C, name='sedcol_set_bioturbation_zw')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: zw(6)
    end function
    ! flux_cap of fabm_sediment_driver.F90:284-291 (bcup_part = 4)
    integer(c_int) function sedcol_set_flux_cap(h, flux_cap) bind(C, name='sedcol_set_flux_cap')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: flux_cap(*)
    end function
    ! get_boundary_conditions (fabm_sediment_component.F90:1865-2030) on the device
    integer(c_int) function sedcol_assemble_forcing(h, temp_surface, c_surface, w_surface, present, flux_out) &
        bind(C, name='sedcol_assemble_forcing')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: temp_surface(*), c_surface(*), w_surface(*)
      integer(c_int32_t), intent(in) :: present(*)
      real(c_double), intent(out) :: flux_out(*)
    end function
    integer(c_int) function sedcol_set_forcing_series(h, nrec, t0, dt_rec, temp, c_bw, flux_top, nrec_flux) &
        bind(C, name='sedcol_set_forcing_series')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: nrec, nrec_flux
      real(c_double), value :: t0, dt_rec
      real(c_double), intent(in) :: temp(*), c_bw(*), flux_top(*)
    end function
    integer(c_int) function sedcol_set_state(h, conc) bind(C, name='sedcol_set_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: conc(*)
    end function
    integer(c_int) function sedcol_get_state(h, conc) bind(C, name='sedcol_get_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: conc(*)
    end function
    integer(c_int) function sedcol_rhs(h, conc, rhs, flux_top_out) bind(C, name='sedcol_rhs')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: conc(*)
      real(c_double), intent(out) :: rhs(*), flux_top_out(*)
    end function
    integer(c_int) function sedcol_step(h, dt, method, conc_inout, flux_top_out) bind(C, name='sedcol_step')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt
      integer(c_int32_t), value :: method
      real(c_double), intent(inout) :: conc_inout(*)
      real(c_double), intent(out) :: flux_top_out(*)
    end function
    integer(c_int) function sedcol_run(h, dt, nsteps, method, steps_per_launch) bind(C, name='sedcol_run')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: dt
      integer(c_int32_t), value :: nsteps, method, steps_per_launch
    end function
    integer(c_int) function sedcol_integrate(h, t_start, t_end) bind(C, name='sedcol_integrate')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: t_start, t_end
    end function
    integer(c_int) function sedcol_get_fluxes(h, flux_top_out) bind(C, name='sedcol_get_fluxes')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: flux_top_out(*)
    end function
    integer(c_int) function sedcol_get_budget(h, budget) bind(C, name='sedcol_get_budget')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: budget(*)
    end function
    integer(c_int) function sedcol_get_last_min_dt(h, v) bind(C, name='sedcol_get_last_min_dt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: v
    end function
    !> clears last_min_dt and the per-column step counters
    integer(c_int) function sedcol_reset_diagnostics(h) bind(C, name='sedcol_reset_diagnostics')
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    !> per-column spin-up (replaces the presimulation of
    !> fabm_sediment_component.F90:557-632); cfloor(nvar); flux_ref and
    !> flux_floor may be c_null_ptr (residual criterion only)
    integer(c_int) function sedcol_spinup(h, t_start, chunk, t_max, tol, cfloor, flux_ref, flux_rtol, &
        flux_floor, t_conv, resid, n_unconverged) bind(C, name='sedcol_spinup')
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: t_start, chunk, t_max, tol, flux_rtol
      real(c_double), intent(in) :: cfloor(*)
      type(c_ptr), value :: flux_ref, flux_floor
      real(c_double), intent(out) :: t_conv(*), resid(*)
      integer(c_int32_t), intent(out) :: n_unconverged
    end function
    !> per-cell DIC and N2-N release of the resident state (the
    !> hzg_omexdia_p_denit diagnostic of the standalone output)
    integer(c_int) function sedcol_rates(h, dic_release, n2_release) bind(C, name='sedcol_rates')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: dic_release(*), n2_release(*)
    end function
  end interface

  public :: sedcol_create, sedcol_destroy, sedcol_set_params, sedcol_set_geometry, sedcol_set_forcing, &
    sedcol_set_forcing_series, sedcol_set_state, sedcol_get_state, sedcol_rhs, sedcol_step, sedcol_run, &
    sedcol_integrate, sedcol_get_fluxes, sedcol_get_budget, sedcol_get_last_min_dt, sedcol_spinup, &
    sedcol_rates, sedcol_reset_diagnostics, sedcol_set_bioturbation_zw, sedcol_set_flux_cap, &
    sedcol_assemble_forcing, sedcol_error_message

  !> GPU-backed counterpart of type_sed + type_rhs_driver
  type, public :: type_sed_hip
    type(c_ptr) :: h = c_null_ptr
    integer :: inum = 0, jnum = 0, knum = 0, nvar = 11
    real(rk) :: dt_min = 1.d-9
    real(rk) :: relative_change_min = -0.9d0
    real(rk) :: last_min_dt = 1.e20
    type(sedcol_params) :: params
    real(rk), dimension(:,:,:,:), pointer :: conc => null()
    real(rk), dimension(:,:,:), pointer :: bdys => null()
    real(rk), dimension(:,:,:), pointer :: fluxes => null()
  contains
    procedure :: initialize => hip_initialize
    procedure :: set_geometry => hip_set_geometry
    procedure :: get_rhs => hip_get_rhs
    procedure :: finalize => hip_finalize
  end type

  public :: ode_solver_hip

contains

  !> error text of the last failing call
  function sedcol_error_message() result(msg)
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    call c_f_pointer(sedcol_last_error(), s, [256])
    msg = ''
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function

  !> create the device domain (inum x jnum columns of knum layers, mask(i,j)
  !! > 0 = land, as fabm_sed_grid%mask_hz) and load the parameters
  subroutine hip_initialize(self, inum, jnum, knum, mask, rc)
    class(type_sed_hip), intent(inout) :: self
    integer, intent(in) :: inum, jnum, knum
    integer(c_int32_t), dimension(inum*jnum), target, intent(in), optional :: mask
    integer, intent(out), optional :: rc
    type(sedcol_desc) :: d
    integer :: rc_
    self%inum = inum
    self%jnum = jnum
    self%knum = knum
    d%ncol = inum*jnum
    d%kmax = knum
    d%nvar = SEDCOL_NVAR
    d%layout = SEDCOL_LAYOUT_FORTRAN
    d%nlayers = c_null_ptr
    d%mask = c_null_ptr
    if (present(mask)) d%mask = c_loc(mask)
    rc_ = sedcol_create(d, self%h)
    self%params%dt_min = self%dt_min
    self%params%relative_change_min = self%relative_change_min
    if (rc_ == 0) rc_ = sedcol_set_params(self%h, self%params)
    if (present(rc)) rc = rc_
  end subroutine

  !> grid and porosity: dz(i,j,k), dzc(i,j,k-1) as fabm_sed_grid
  !! (fabm_sediment_driver.F90:41-56), porosity and interface porosity,
  !! bioturbation factor (:96-98) and coefficient
  subroutine hip_set_geometry(self, dz, dzc, porosity, intf_porosity, bioturbation_factor, bioturbation, rc)
    class(type_sed_hip), intent(inout) :: self
    real(rk), intent(in) :: dz(:,:,:), dzc(:,:,:), porosity(:,:,:), intf_porosity(:,:,:)
    real(rk), intent(in) :: bioturbation_factor(:,:,:), bioturbation
    integer, intent(out), optional :: rc
    real(rk), allocatable :: dzc_k(:,:,:), biot(:)
    integer :: rc_
    allocate(dzc_k(self%inum, self%jnum, self%knum), biot(self%inum*self%jnum))
    dzc_k(:,:,1) = 0.0_rk
    if (self%knum > 1) dzc_k(:,:,2:self%knum) = dzc(:,:,1:self%knum-1)
    biot = bioturbation
    rc_ = sedcol_set_geometry(self%h, dz, dzc_k, porosity, intf_porosity, bioturbation_factor, biot)
    if (present(rc)) rc = rc_
  end subroutine

  !> get_rhs (fabm_sediment_driver.F90:575-717) for the current conc; the top
  !! fluxes go to fluxes(:,:,n) as at :692
  subroutine hip_get_rhs(self, rhs, rc)
    class(type_sed_hip), intent(inout) :: self
    real(rk), dimension(:,:,:,:), intent(out) :: rhs
    integer, intent(out), optional :: rc
    real(rk), allocatable :: f(:)
    integer :: rc_
    rc_ = sedcol_set_forcing(self%h, self%bdys(:,:,1), self%bdys(:,:,2:self%nvar+1), self%fluxes)
    allocate(f(self%inum*self%jnum*self%nvar))
    if (rc_ == 0) rc_ = sedcol_rhs(self%h, self%conc, rhs, f)
    if (rc_ == 0) call keep_solute_fluxes(self, f)
    if (present(rc)) rc = rc_
  end subroutine

  !> one ode_solver call on conc (solver_library.F90:80), forcing from
  !! bdys/fluxes, the caller's minimum clamp included when params%clamp
  subroutine ode_solver_hip(self, dt, method, rc)
    class(type_sed_hip), intent(inout) :: self
    real(rk), intent(in) :: dt
    integer, intent(in) :: method
    integer, intent(out), optional :: rc
    real(rk), allocatable :: f(:)
    real(c_double) :: lm
    integer :: rc_
    rc_ = sedcol_set_forcing(self%h, self%bdys(:,:,1), self%bdys(:,:,2:self%nvar+1), self%fluxes)
    allocate(f(self%inum*self%jnum*self%nvar))
    if (rc_ == 0) rc_ = sedcol_step(self%h, real(dt, c_double), int(method, c_int32_t), self%conc, f)
    if (rc_ == 0) then
      call keep_solute_fluxes(self, f)
      rc_ = sedcol_get_last_min_dt(self%h, lm)
      self%last_min_dt = lm
    end if
    if (present(rc)) rc = rc_
  end subroutine

  !> solutes report the computed top flux, particulates keep the prescribed one
  subroutine keep_solute_fluxes(self, f)
    class(type_sed_hip), intent(inout) :: self
    real(rk), intent(in) :: f(:)
    integer :: n, nc
    nc = self%inum*self%jnum
    do n = 6, 10
      self%fluxes(:,:,n) = reshape(f((n-1)*nc+1:n*nc), [self%inum, self%jnum])
    end do
  end subroutine

  subroutine hip_finalize(self)
    class(type_sed_hip), intent(inout) :: self
    integer :: rc_
    if (c_associated(self%h)) rc_ = sedcol_destroy(self%h)
    self%h = c_null_ptr
  end subroutine

end module sedcol_hip
