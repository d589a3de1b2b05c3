!> fabmsed1d_hip.F90 -- the standalone sediment column, read from the
!! reference's namelists, with the time loop resident on the GPU.
!!
!! Takes the place of the program fabmsed1d of
!! examples/standalone/omexdia_p/main.F90:10-166.  Run it in a directory
!! holding run_sed.nml (&run_nml, then &sed_nml) and fabm_sed.nml
!! (&fabm_nml, &hzg_omexdia_p); it writes output.dat in the layout of
!! main.F90:109-157, which plotsed1d.py reads.
!!
!! Set-up, as the reference does it:
!!   &run_nml defaults and reading          main.F90:41-54 (output = 86400/60
!!                                          when the namelist leaves it out:
!!                                          the default is taken with dt = 60)
!!   grid                                   fabm_sediment_driver.F90:127-177
!!   &sed_nml defaults, porosity, flux cap,
!!   bioturbation profile 0/1/2             :211-303
!!   bioturbation profile 3 (Zhang-Wirtz)   :618-645, sedcol_set_bioturbation_zw
!!   interface porosity                     :434-435
!!   initial value / porosity               :449-456
!!   bottom water, detritus fluxes          main.F90:82-107 (ODU = pel_O2, :96)
!!   tnum steps of ode_method + clamp       main.F90:120-131
!!   output every `output` steps            main.F90:133-157
!!
!! Between outputs the state stays on the GPU: each block of steps is one
!! sedcol_run call (ode_method 0-5) or one sedcol_integrate call over the
!! same time (6 adaptive EMP, 7 ROS2; the reference's ode_solver has neither).
!! The forcing is constant over the run.  The reference writes the computed
!! solute top fluxes back into `fluxes` after every step; with
!! bcup_dissolved_variables = 1 the computed top flux is the given one, so
!! the next step sees the same input and a resident block is the same run.
!!
!! Keys of &hzg_omexdia_p that the library reads: rLabile, rSemilabile,
!! CprodMax, rnit, ksO2nitri, rODUox, ksO2oduox, ksO2oxic, ksNO3denit,
!! kinO2denit, kinNO3anox, kinO2anox, PAds, PAdsODU and the *_init values
!! (fabm_sed.nml:51-77).  NH3Ads must be 0 (ammonium adsorption is not in the
!! library's kinetics).  The library carries detritus nitrogen as two state
!! variables; their initial values are ldetN_init and sdetN_init when given,
!! else NCrLdet*ldetC_init and NCrSdet*sdetC_init, the ratio that FABM's
!! hzg_omexdia_p keeps fixed.  Adsorbed phosphate starts at adsP_init (0).
!!
!! Usage: fabmsed1d_hip [--setup FILE] [--dump FILE] [--steps N]
!!   --setup FILE  write the set-up (stream file, layout below) and stop
!!                 before the first GPU call
!!   --dump FILE   after the run, write the final state (stream file)
!!   --steps N     run N steps instead of numyears*365*86400/int(dt)
!! Setup file: int32 knum, nvar, tnum, output, ode_method, bcup_dis,
!! bcup_part, bioturbation_profile; then doubles dt, bioturbation, zw(6),
!! the 17 real fields of sedcol_params from D0 to PAdsODU, dz, zc, zi(knum+1),
!! dzc (0 at k = 1), porosity, interface porosity, bioturbation factor,
!! flux cap, conc(knum, nvar), bdys(nvar+1), fluxes(nvar).
!! Dump file: int32 knum, nvar, steps; doubles conc(knum, nvar), fluxes(nvar),
!! denit(knum) (mmol N m-3 d-1).
program fabmsed1d_hip

  use iso_c_binding
  use sedcol_hip
  implicit none

  integer, parameter :: nvar = 11
  ! reference state ids of the library order (include/sedcol.h enum sedcol_var)
  integer, parameter :: v_ldetC = 1, v_sdetC = 2, v_ldetN = 3, v_sdetN = 4, v_detP = 5, v_O2 = 6, &
    v_NO3 = 7, v_NH4 = 8, v_ODU = 9, v_PO4 = 10, v_adsP = 11
  character(len=40), parameter :: state_names(nvar) = [character(len=40) :: &
    'detritus_labile_carbon', 'detritus_semilabile_carbon', 'detritus_labile_nitrogen', &
    'detritus_semilabile_nitrogen', 'detritus_labile_phosphorus', 'dissolved_oxygen', 'dissolved_nitrate', &
    'dissolved_ammonium', 'dissolved_reduced_substances', 'dissolved_phosphate', 'phosphate_adsorbed']
  character(len=*), parameter :: diag_name = 'hzg_omexdia_p_denit'

  ! &run_nml (main.F90:18-39)
  real(rk) :: dt, dzmin, dt_min, relative_change_min, bcup_dissolved_variables
  integer :: numyears, output, numlayers, ode_method, presimulation_years
  real(rk) :: pel_Temp, pel_NO3, pel_NH4, pel_PO4, pel_O2
  real(rk) :: pflux_lDetC, pflux_sDetC, pflux_lDetN, pflux_sDetN, pflux_lDetP
  real(rk) :: pflux_sDet, pflux_fDet, pflux_detP
  character(len=256) :: ugrid_name
  namelist /run_nml/ numyears, dt, output, numlayers, dzmin, ode_method, dt_min, relative_change_min, &
    ugrid_name, bcup_dissolved_variables, presimulation_years, pel_Temp, pel_NO3, pel_NH4, pel_PO4, &
    pel_O2, pflux_lDetC, pflux_sDetC, pflux_lDetN, pflux_sDetN, pflux_lDetP, pflux_sDet, pflux_fDet, pflux_detP

  ! &sed_nml (fabm_sediment_driver.F90:204-233)
  integer :: bioturbation_profile
  logical :: distributed_pom_flux
  real(rk) :: diffusivity, bioturbation, porosity_max, porosity_fac, k_par, bioturbation_depth
  real(rk) :: bioturbation_min, pom_flux_max
  real(rk) :: bioturb_k_l, bioturb_L1, bioturb_L2, bioturb_beta, bioturb_b, bioturb_dry_density
  namelist /sed_nml/ diffusivity, bioturbation_profile, bioturbation, porosity_max, porosity_fac, k_par, &
    distributed_pom_flux, pom_flux_max, bioturbation_depth, bioturbation_min, bioturb_k_l, bioturb_L1, &
    bioturb_L2, bioturb_beta, bioturb_b, bioturb_dry_density

  ! &fabm_nml and &hzg_omexdia_p (examples/standalone/omexdia_p/fabm_sed.nml)
  character(len=64) :: models(8)
  namelist /fabm_nml/ models
  real(rk) :: rLabile, rSemilabile, NCrLdet, NCrSdet, PAds, PAdsODU, NH3Ads, CprodMax, rnit, ksO2nitri
  real(rk) :: rODUox, ksO2oduox, ksO2oxic, ksNO3denit, kinO2denit, kinNO3anox, kinO2anox
  real(rk) :: ldetC_init, sdetC_init, ldetN_init, sdetN_init, ldetP_init, detP_init, oxy_init, odu_init
  real(rk) :: no3_init, nh3_init, po4_init, adsP_init
  namelist /hzg_omexdia_p/ rLabile, rSemilabile, NCrLdet, NCrSdet, PAds, PAdsODU, NH3Ads, CprodMax, rnit, &
    ksO2nitri, rODUox, ksO2oduox, ksO2oxic, ksNO3denit, kinO2denit, kinNO3anox, kinO2anox, ldetC_init, &
    sdetC_init, ldetN_init, sdetN_init, ldetP_init, detP_init, oxy_init, odu_init, no3_init, nh3_init, &
    po4_init, adsP_init

  type(type_sed_hip) :: sed
  real(rk), allocatable :: dz(:,:,:), zc(:,:,:), zi(:,:,:), dzc(:,:,:), dzc_k(:,:,:)
  real(rk), allocatable :: porosity(:,:,:), intf_porosity(:,:,:), bfac(:,:,:), fcap(:,:,:)
  real(rk), allocatable, target :: conc(:,:,:,:), bdys(:,:,:), fluxes(:,:,:)
  real(rk), allocatable :: ftop(:), dic(:), denit(:)
  real(rk) :: fac, init(nvar), zw(6), lm
  integer :: k, n, t, tnum, steps_arg, rc, u, nml_unit, iostat, chunk
  character(len=512) :: arg, setup_file, dump_file, iomsg
  integer :: iarg

  ! command line
  setup_file = ''
  dump_file = ''
  steps_arg = -1
  iarg = 1
  do while (iarg <= command_argument_count())
    call get_command_argument(iarg, arg)
    select case (trim(arg))
    case ('--setup')
      iarg = iarg + 1
      call get_command_argument(iarg, setup_file)
    case ('--dump')
      iarg = iarg + 1
      call get_command_argument(iarg, dump_file)
    case ('--steps')
      iarg = iarg + 1
      call get_command_argument(iarg, arg)
      read(arg, *, iostat=iostat) steps_arg
      if (iostat /= 0 .or. steps_arg < 0) call fail('--steps takes a count >= 0')
    case default
      call fail('unknown argument '//trim(arg)//' (usage: fabmsed1d_hip [--setup FILE] [--dump FILE] [--steps N])')
    end select
    iarg = iarg + 1
  end do

  ! &run_nml defaults (main.F90:27-50); output is taken with dt = 60
  dt = 60.0_rk
  numyears = 1
  output = 86400/int(dt)
  numlayers = 10
  dzmin = 0.005_rk
  ode_method = SEDCOL_RUNGE_KUTTA_4
  dt_min = 1.0_rk
  relative_change_min = -0.9_rk
  bcup_dissolved_variables = 2
  presimulation_years = -1
  ugrid_name = ''
  pel_Temp = 5.0_rk; pel_NO3 = 5.0_rk; pel_NH4 = 5.0_rk; pel_PO4 = 0.5_rk; pel_O2 = 250.0_rk
  pflux_lDetC = 10.0_rk; pflux_sDetC = 10.0_rk; pflux_lDetN = 1.5_rk; pflux_sDetN = 1.5_rk
  pflux_lDetP = 0.2_rk
  pflux_sDet = 10.0_rk; pflux_fDet = 10.0_rk; pflux_detP = 0.2_rk

  ! &sed_nml defaults (fabm_sediment_driver.F90:215-231)
  bioturbation_profile = 1
  distributed_pom_flux = .false.
  diffusivity = 0.9_rk
  bioturbation = 0.9_rk
  bioturbation_depth = 5.0_rk
  bioturbation_min = 0.2_rk
  porosity_max = 0.7_rk
  porosity_fac = 0.9_rk
  k_par = 2.0e-3_rk
  pom_flux_max = 2.0e4_rk
  bioturb_k_l = 0.11_rk; bioturb_L1 = 0.2_rk; bioturb_L2 = 0.6_rk; bioturb_beta = 0.22_rk
  bioturb_b = 1.334_rk; bioturb_dry_density = 1000.0_rk

  ! &hzg_omexdia_p: the library's default parameters; unset initial values
  ! are marked negative and resolved below
  models = ''
  rLabile = 0.043_rk; rSemilabile = 0.001_rk; NCrLdet = 0.22_rk; NCrSdet = 0.005_rk
  PAds = 0.01_rk; PAdsODU = 70.0_rk; NH3Ads = 0.0_rk; CprodMax = 9600.0_rk; rnit = 200.0_rk
  ksO2nitri = 20.0_rk; rODUox = 20.0_rk; ksO2oduox = 1.0_rk; ksO2oxic = 3.0_rk; ksNO3denit = 1.0_rk
  kinO2denit = 70.0_rk; kinNO3anox = 1.0_rk; kinO2anox = 1.0_rk
  ldetC_init = 0.0_rk; sdetC_init = 0.0_rk; ldetN_init = -1.0_rk; sdetN_init = -1.0_rk
  ldetP_init = -1.0_rk; detP_init = -1.0_rk; oxy_init = 0.0_rk; odu_init = 0.0_rk; no3_init = 0.0_rk
  nh3_init = 0.0_rk; po4_init = 0.0_rk; adsP_init = 0.0_rk

  ! run_sed.nml: &run_nml, then &sed_nml from the same unit (main.F90:52-71,
  ! fabm_sediment_driver.F90:233)
  open(newunit=nml_unit, file='run_sed.nml', action='read', status='old', iostat=iostat, iomsg=iomsg)
  if (iostat /= 0) call fail('cannot open run_sed.nml: '//trim(iomsg))
  read(nml_unit, nml=run_nml, iostat=iostat, iomsg=iomsg)
  if (iostat /= 0) call fail('reading &run_nml of run_sed.nml: '//trim(iomsg))
  read(nml_unit, nml=sed_nml, iostat=iostat, iomsg=iomsg)
  if (iostat /= 0) call fail('reading &sed_nml of run_sed.nml: '//trim(iomsg))
  close(nml_unit)

  open(newunit=nml_unit, file='fabm_sed.nml', action='read', status='old', iostat=iostat, iomsg=iomsg)
  if (iostat /= 0) call fail('cannot open fabm_sed.nml: '//trim(iomsg))
  read(nml_unit, nml=fabm_nml, iostat=iostat, iomsg=iomsg)
  if (iostat /= 0) call fail('reading &fabm_nml of fabm_sed.nml: '//trim(iomsg))
  if (trim(models(1)) /= 'hzg_omexdia_p' .or. any(len_trim(models(2:)) > 0)) &
    call fail('&fabm_nml models must be ''hzg_omexdia_p'' alone, the model the library implements')
  read(nml_unit, nml=hzg_omexdia_p, iostat=iostat, iomsg=iomsg)
  if (iostat /= 0) call fail('reading &hzg_omexdia_p of fabm_sed.nml: '//trim(iomsg))
  close(nml_unit)
  if (NH3Ads /= 0.0_rk) call fail('NH3Ads must be 0: ammonium adsorption is not modelled')
  if (ldetN_init < 0.0_rk) ldetN_init = NCrLdet*ldetC_init
  if (sdetN_init < 0.0_rk) sdetN_init = NCrSdet*sdetC_init
  if (detP_init < 0.0_rk) detP_init = max(ldetP_init, 0.0_rk)

  if (numlayers < 2) call fail('numlayers must be at least 2')
  if (int(dt) < 1) call fail('dt must be at least 1 s')
  if (output < 1) call fail('output must be at least 1')
  if (ode_method < SEDCOL_EULER .or. ode_method > SEDCOL_ROS2_ADAPTIVE) call fail('ode_method must be 0..7')
  if (bioturbation_profile < 0 .or. bioturbation_profile > 3) call fail('bioturbation_profile must be 0..3')
  tnum = numyears*365*86400/int(dt)
  if (steps_arg >= 0) tnum = steps_arg

  ! grid (fabm_sediment_driver.F90:147-168), one column
  allocate(dz(1,1,numlayers), zc(1,1,numlayers), zi(1,1,numlayers+1), dzc(1,1,numlayers-1))
  allocate(dzc_k(1,1,numlayers), porosity(1,1,numlayers), intf_porosity(1,1,numlayers))
  allocate(bfac(1,1,numlayers), fcap(1,1,numlayers))
  fac = 0.18_rk/((numlayers+1)/2.0_rk * dzmin) - 1.0_rk
  zi(:,:,1) = 0.0_rk
  do k = 1, numlayers
    dz(:,:,k) = (1.0_rk + (fac-1.0_rk)*(k-1)/(numlayers-1)) * dzmin
    zc(:,:,k) = zi(:,:,k) + 0.5_rk*dz(:,:,k)
    zi(:,:,k+1) = zi(:,:,k) + dz(:,:,k)
  end do
  dzc = zc(:,:,2:numlayers) - zc(:,:,1:numlayers-1)
  dzc_k(:,:,1) = 0.0_rk
  dzc_k(:,:,2:numlayers) = dzc

  ! porosity, flux cap and bioturbation factor (:273-303)
  bfac = 1.0_rk
  do k = 1, numlayers
    porosity(:,:,k) = porosity_max * (1.0_rk - porosity_fac * zc(:,:,k))
    fcap(:,:,k) = pom_flux_max/86400.0_rk * (1.0_rk - porosity(:,:,k)) * dz(:,:,k)
    if (k > 2) fcap(:,:,k) = min(fcap(:,:,k), fcap(:,:,k-1))
    select case (bioturbation_profile)
    case (1)
      bfac(:,:,k) = max(bioturbation_min/bioturbation, max(bioturbation_depth-100.0_rk*zi(:,:,k), 0.0_rk) &
                        /bioturbation_depth)
    case (2)
      bfac(:,:,k) = exp(-100.0_rk*zi(:,:,k)/bioturbation_depth)
    end select
  end do
  intf_porosity(:,:,1) = porosity(:,:,1)
  intf_porosity(:,:,2:numlayers) = 0.5_rk*(porosity(:,:,1:numlayers-1) + porosity(:,:,2:numlayers))
  zw = [bioturb_k_l, bioturb_L1, bioturb_L2, bioturb_beta, bioturb_b, bioturb_dry_density]

  ! initial concentrations per pore water (:449-456)
  init = 0.0_rk
  init(v_ldetC) = ldetC_init
  init(v_sdetC) = sdetC_init
  init(v_ldetN) = ldetN_init
  init(v_sdetN) = sdetN_init
  init(v_detP) = detP_init
  init(v_O2) = oxy_init
  init(v_NO3) = no3_init
  init(v_NH4) = nh3_init
  init(v_ODU) = odu_init
  init(v_PO4) = po4_init
  init(v_adsP) = adsP_init
  allocate(conc(1,1,numlayers,nvar), bdys(1,1,nvar+1), fluxes(1,1,nvar))
  do n = 1, nvar
    conc(:,:,:,n) = init(n)/porosity
  end do

  ! boundary conditions (main.F90:82-107)
  bdys = 0.0_rk
  bdys(:,:,1) = pel_Temp
  bdys(:,:,1+v_NO3) = pel_NO3
  bdys(:,:,1+v_NH4) = pel_NH4
  bdys(:,:,1+v_PO4) = pel_PO4
  bdys(:,:,1+v_O2) = pel_O2
  bdys(:,:,1+v_ODU) = pel_O2
  fluxes = 0.0_rk
  fluxes(:,:,v_ldetC) = pflux_lDetC/86400.0_rk
  fluxes(:,:,v_sdetC) = pflux_sDetC/86400.0_rk
  fluxes(:,:,v_ldetN) = pflux_lDetN/86400.0_rk
  fluxes(:,:,v_sdetN) = pflux_sDetN/86400.0_rk
  fluxes(:,:,v_detP) = pflux_lDetP/86400.0_rk

  ! library parameters: namelist rates per day to per second
  sed%dt_min = dt_min
  sed%relative_change_min = relative_change_min
  sed%params%D0 = diffusivity
  sed%params%rL = rLabile/86400.0_rk
  sed%params%rS = rSemilabile/86400.0_rk
  sed%params%CprodMax = CprodMax/86400.0_rk
  sed%params%rnit = rnit/86400.0_rk
  sed%params%ksO2nitri = ksO2nitri
  sed%params%rODUox = rODUox/86400.0_rk
  sed%params%ksO2oduox = ksO2oduox
  sed%params%ksO2oxic = ksO2oxic
  sed%params%ksNO3denit = ksNO3denit
  sed%params%kinO2denit = kinO2denit
  sed%params%kinNO3anox = kinNO3anox
  sed%params%kinO2anox = kinO2anox
  sed%params%rPads = (PAds*rODUox)/86400.0_rk
  sed%params%PAdsODU = PAdsODU
  sed%params%bcup_dis = int(bcup_dissolved_variables)
  sed%params%bcup_part = 1
  if (distributed_pom_flux) sed%params%bcup_part = 4
  sed%params%clamp = 1
  sed%params%adaptive_method = 0
  if (ode_method == SEDCOL_ROS2_ADAPTIVE) sed%params%adaptive_method = SEDCOL_ROS2_ADAPTIVE
  sed%params%dt_min = dt_min
  sed%params%relative_change_min = relative_change_min

  if (len_trim(setup_file) > 0) then
    call write_setup()
  else
    call run_column()
  end if

contains

  !> the set-up file of --setup; nothing touches the GPU before it
  subroutine write_setup()
    open(newunit=u, file=trim(setup_file), access='stream', form='unformatted', status='replace')
    write(u) int(numlayers, c_int32_t), int(nvar, c_int32_t), int(tnum, c_int32_t), int(output, c_int32_t), &
      int(ode_method, c_int32_t), int(sed%params%bcup_dis, c_int32_t), int(sed%params%bcup_part, c_int32_t), &
      int(bioturbation_profile, c_int32_t)
    write(u) dt, bioturbation, zw
    write(u) sed%params%D0, sed%params%dt_min, sed%params%relative_change_min, sed%params%rL, sed%params%rS, &
      sed%params%CprodMax, sed%params%rnit, sed%params%ksO2nitri, sed%params%rODUox, sed%params%ksO2oduox, &
      sed%params%ksO2oxic, sed%params%ksNO3denit, sed%params%kinO2denit, sed%params%kinNO3anox, &
      sed%params%kinO2anox, sed%params%rPads, sed%params%PAdsODU
    write(u) dz, zc, zi, dzc_k, porosity, intf_porosity, bfac, fcap, conc, bdys, fluxes
    close(u)
    write(*,'(A,A)') 'fabmsed1d_hip: set-up written to ', trim(setup_file)
  end subroutine

  !> the run of main.F90:66-161 on the GPU
  subroutine run_column()
    write(0,*) '  Initialise grid with', numlayers, 'vertical layers'
    call sed%initialize(1, 1, numlayers, rc=rc)
    if (rc /= 0) call die('initialize', rc)
    call sed%set_geometry(dz, dzc, porosity, intf_porosity, bfac, bioturbation, rc=rc)
    if (rc /= 0) call die('set_geometry', rc)
    if (bioturbation_profile == 3) then
      rc = sedcol_set_bioturbation_zw(sed%h, zw)
      if (rc /= 0) call die('set_bioturbation_zw', rc)
    end if
    if (distributed_pom_flux) then
      rc = sedcol_set_flux_cap(sed%h, fcap)
      if (rc /= 0) call die('set_flux_cap', rc)
    end if
    rc = sedcol_set_forcing(sed%h, bdys(:,:,1), bdys(:,:,2:nvar+1), fluxes)
    if (rc /= 0) call die('set_forcing', rc)
    rc = sedcol_set_state(sed%h, conc)
    if (rc /= 0) call die('set_state', rc)
    allocate(ftop(nvar), dic(numlayers), denit(numlayers))

    ! output.dat header (main.F90:109-117)
    ! one record per line, however many variables (list-directed output
    ! otherwise wraps at the runtime's default record length)
    open(newunit=u, file='output.dat', status='replace', action='write', recl=65536)
    write(u, fmt='(A,A,A,A)', advance='no') 'time(s) ', 'depth(m) ', 'layer-height(m) ', 'porosity() '
    do n = 1, nvar
      write(u, fmt='(A,A)', advance='no') ' ', trim(state_names(n))
    end do
    write(u, fmt='(A,A)', advance='no') ' ', diag_name
    write(u, *)

    ! time loop (main.F90:120-158): blocks of steps up to the next output
    t = 0
    do while (t < tnum)
      chunk = min(output - mod(t, output), tnum - t)
      if (ode_method >= SEDCOL_EMP_ADAPTIVE) then
        rc = sedcol_integrate(sed%h, real(t, c_double)*dt, real(t + chunk, c_double)*dt)
        if (rc /= 0) call die('integrate', rc)
      else
        rc = sedcol_run(sed%h, dt, int(chunk, c_int32_t), int(ode_method, c_int32_t), 0_c_int32_t)
        if (rc /= 0) call die('run', rc)
      end if
      t = t + chunk
      if (mod(t, output) == 0) then
        call fetch()
        write(0,*) ' elapsed ', t*dt/86400, ' days'
        if (ode_method == SEDCOL_ADAPTIVE_EULER) then
          rc = sedcol_get_last_min_dt(sed%h, lm)
          if (rc /= 0) call die('get_last_min_dt', rc)
          write(0, fmt='(A,F6.2,A)') '   last minimum timestep: ', min(lm, dt), 's'
          rc = sedcol_reset_diagnostics(sed%h)
          if (rc /= 0) call die('reset_diagnostics', rc)
        end if
        write(u, *) t*dt, 'fluxes', fluxes(1,1,:)
        do k = 1, numlayers
          write(u, fmt='(E15.3,A,E15.4E3,A,E15.4E3,A,E15.4E3)', advance='no') t*dt, ' ', zc(1,1,k), ' ', &
            dz(1,1,k), ' ', porosity(1,1,k)
          do n = 1, nvar
            write(u, fmt='(A,E15.4E3)', advance='no') ' ', conc(1,1,k,n)
          end do
          write(u, fmt='(A,E15.4E3)', advance='no') ' ', denit(k)
          write(u, *)
        end do
      end if
    end do
    close(u)

    if (len_trim(dump_file) > 0) then
      call fetch()
      open(newunit=u, file=trim(dump_file), access='stream', form='unformatted', status='replace')
      write(u) int(numlayers, c_int32_t), int(nvar, c_int32_t), int(tnum, c_int32_t)
      write(u) conc, fluxes, denit
      close(u)
    end if
    call sed%finalize()
    write(*,'(A,I0,A)') 'fabmsed1d_hip: ', tnum, ' steps done'
  end subroutine

  !> state, top fluxes (solutes computed, particulates prescribed) and the
  !! denitrification diagnostic of the resident state
  subroutine fetch()
    integer :: rc_
    rc_ = sedcol_get_state(sed%h, conc)
    if (rc_ /= 0) call die('get_state', rc_)
    rc_ = sedcol_get_fluxes(sed%h, ftop)
    if (rc_ /= 0) call die('get_fluxes', rc_)
    fluxes(1,1,v_O2:v_PO4) = ftop(v_O2:v_PO4)
    rc_ = sedcol_rates(sed%h, dic, denit)
    if (rc_ /= 0) call die('rates', rc_)
    denit = denit*86400.0_rk
  end subroutine

  subroutine fail(msg)
    character(len=*), intent(in) :: msg
    write(0,'(A,A)') 'fabmsed1d_hip: ', trim(msg)
    error stop 1
  end subroutine

  subroutine die(what, code)
    character(len=*), intent(in) :: what
    integer, intent(in) :: code
    write(0,'(A,A,A,I0,A,A)') 'fabmsed1d_hip: ', what, ' failed with status ', code, ': ', &
      trim(sedcol_error_message())
    error stop 1
  end subroutine

end program fabmsed1d_hip
