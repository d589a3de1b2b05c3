!> fabmsed_hip.F90 -- Fortran host driving the GPU library through sedcol_hip.
!!
!! Follows the standalone harness of the reference
!! (examples/standalone/omexdia_p/main.F90:41-158): reference grid
!! (init_grid, fabm_sediment_driver.F90:127-177), linear porosity (:280) and
!! interface porosity (:434-435), bioturbation profile 0, initial
!! concentrations = namelist initial values / porosity (:456), bottom water
!! and detritus fluxes as main.F90:82-107 (including ODU_bw = O2_bw at :96),
!! then nsteps ode_solver calls with RK4 and the minimum clamp.
!! Usage: fabmsed_hip <output file> [nsteps] [dt]
!! The output (stream access, native doubles after four int32: inum, jnum,
!! knum, nsteps) holds dz, dzc(:,:,k>=2 at k), porosity, interface porosity,
!! bioturbation factor, initial conc, bdys, fluxes, final conc, final fluxes.
program fabmsed_hip

  use iso_c_binding
  use sedcol_hip
  implicit none

  integer, parameter :: inum = 2, jnum = 2, knum = 15, nvar = 11
  real(rk), parameter :: dzmin = 0.004_rk, porosity_max = 0.7_rk, porosity_fac = 0.9_rk
  real(rk), parameter :: bioturbation = 4.5_rk
  type(type_sed_hip) :: sed
  real(rk), dimension(inum, jnum, knum) :: dz, zc, porosity, intf_porosity, bfac, dzc_k
  real(rk), dimension(inum, jnum, knum+1) :: zi
  real(rk), dimension(inum, jnum, knum-1) :: dzc
  real(rk), dimension(inum, jnum, knum, nvar), target :: conc, conc0
  real(rk), dimension(inum, jnum, nvar+1), target :: bdys
  real(rk), dimension(inum, jnum, nvar), target :: fluxes, fluxes0
  real(rk), dimension(nvar) :: init
  real(rk) :: fac, dt
  integer :: i, j, k, n, t, nsteps, rc, u
  character(len=512) :: out, arg

  call get_command_argument(1, out)
  if (len_trim(out) == 0) out = 'fabmsed_hip.bin'
  nsteps = 240
  dt = 360.0_rk
  call get_command_argument(2, arg)
  if (len_trim(arg) > 0) read(arg, *) nsteps
  call get_command_argument(3, arg)
  if (len_trim(arg) > 0) read(arg, *) dt

  ! grid (fabm_sediment_driver.F90:147-168)
  fac = 0.18_rk/((knum+1)/2.0_rk * dzmin) - 1.0_rk
  zi(:,:,1) = 0.0_rk
  do k = 1, knum; do j = 1, jnum; do i = 1, inum
    dz(i,j,k) = (1.0_rk + (fac-1.0_rk)*(k-1)/(knum-1)) * dzmin
    zc(i,j,k) = zi(i,j,k) + 0.5_rk*dz(i,j,k)
    zi(i,j,k+1) = zi(i,j,k) + dz(i,j,k)
  end do; end do; end do
  dzc = zc(:,:,2:knum) - zc(:,:,1:knum-1)
  ! porosity (:280) with a small per-column offset, interface porosity (:434-435)
  do k = 1, knum
    do j = 1, jnum
      do i = 1, inum
        porosity(i,j,k) = (porosity_max - 0.02_rk*(i-1) - 0.01_rk*(j-1)) * (1.0_rk - porosity_fac*zc(i,j,k))
      end do
    end do
  end do
  intf_porosity(:,:,1) = porosity(:,:,1)
  intf_porosity(:,:,2:knum) = 0.5_rk*(porosity(:,:,1:knum-1) + porosity(:,:,2:knum))
  bfac = 1.0_rk

  ! initial values (examples/standalone/omexdia_p/fabm_sed.nml:36-45), per pore water
  init = [4.0e3_rk, 4.0e3_rk, 0.92e3_rk, 0.04e3_rk, 40.0_rk, 100.0_rk, 20.0_rk, 40.0_rk, 100.0_rk, 10.0_rk, 0.0_rk]
  do n = 1, nvar
    conc(:,:,:,n) = init(n)/porosity
  end do
  conc0 = conc

  ! boundary conditions (main.F90:82-107, run_sed.nml)
  bdys = 0.0_rk
  bdys(:,:,1) = 5.0_rk
  bdys(:,:,1+6) = 250.0_rk   ! O2
  bdys(:,:,1+7) = 14.0_rk    ! NO3
  bdys(:,:,1+8) = 4.0_rk     ! NH4
  bdys(:,:,1+9) = 250.0_rk   ! ODU = O2 as in main.F90:96
  bdys(:,:,1+10) = 0.6_rk    ! PO4
  fluxes = 0.0_rk
  fluxes(:,:,1) = 2.0_rk/86400.0_rk
  fluxes(:,:,2) = 24.0_rk/86400.0_rk
  fluxes(:,:,3) = 0.44_rk/86400.0_rk
  fluxes(:,:,4) = 0.12_rk/86400.0_rk
  fluxes(:,:,5) = 0.08_rk/86400.0_rk
  fluxes0 = fluxes

  sed%params%clamp = 1
  sed%params%bcup_dis = 2
  call sed%initialize(inum, jnum, knum, rc=rc)
  if (rc /= 0) call die('initialize', rc)
  call sed%set_geometry(dz, dzc, porosity, intf_porosity, bfac, bioturbation, rc=rc)
  if (rc /= 0) call die('set_geometry', rc)
  sed%conc => conc
  sed%bdys => bdys
  sed%fluxes => fluxes

  do t = 1, nsteps
    call ode_solver_hip(sed, dt, int(SEDCOL_RUNGE_KUTTA_4), rc)
    if (rc /= 0) call die('ode_solver_hip', rc)
  end do
  call sed%finalize()

  dzc_k(:,:,1) = 0.0_rk
  dzc_k(:,:,2:knum) = dzc
  open(newunit=u, file=trim(out), access='stream', form='unformatted', status='replace')
  write(u) int(inum, c_int32_t), int(jnum, c_int32_t), int(knum, c_int32_t), int(nsteps, c_int32_t)
  write(u) dz, dzc_k, porosity, intf_porosity, bfac, conc0, bdys, fluxes0, conc, fluxes
  close(u)
  write(*,'(A,I0,A)') 'fabmsed_hip: ', nsteps, ' RK4 steps done'

contains

  subroutine die(what, code)
    character(len=*), intent(in) :: what
    integer, intent(in) :: code
    write(0,'(A,A,I0,A,A)') what, ' failed with status ', code, ': ', trim(sedcol_error_message())
    error stop 1
  end subroutine

end program fabmsed_hip
