/*
 * sedcol.h -- C ABI of the MI355X sediment-column library (libsedcol.so).
 *
 * The library replaces the reference's sediment-column hot path:
 *   - get_rhs   src/drivers/fabm_sediment_driver.F90:575-717 (with diff3d
 *               :739-825 and the local reaction rates called at :700),
 *   - ode_solver src/utilities/solver_library.F90:80-189 (methods 0..3 keep
 *               their meaning; 4..6 add positivity-preserving EMP steppers),
 * driven per horizontal column on one GPU.  Fortran callers reach it through
 * the ISO_C_BINDING module mossco-code_amd/fortran/sedcol_hip.F90.
 *
 * Conventions
 *   - every function returns int status: 0 = SEDCOL_OK, otherwise a
 *     sedcol_status code; sedcol_last_error() describes the last failure of
 *     the calling thread.  Nothing calls exit/stop/abort.
 *   - all pointers are caller-owned host memory, read or written during the
 *     call only (the library copies), except the profiles given to
 *     sedcol_set_state_profiles (device memory); FP64 everywhere.
 *   - array layout is chosen at create time (sedcol_desc.layout):
 *       SEDCOL_LAYOUT_FORTRAN  uniform K; cell arrays (ncol, K) and state
 *                              (ncol, K, nvar) with the column index fastest,
 *                              i.e. Fortran conc(i,j,k,n) with ncol = I*J;
 *       SEDCOL_LAYOUT_PACKED   column c owns cells off[c]..off[c+1]-1 (top
 *                              first, off from nlayers); cell arrays (ncell)
 *                              and state (nvar, ncell), cell index fastest.
 *     Per-column arrays are (ncol); per-column-per-variable arrays are
 *     (nvar, ncol) with the column index fastest, like bdys/fluxes
 *     (fabm_sediment_component.F90:538-541).
 *   - a handle owns its device buffers and one HIP stream; it is not
 *     thread-safe (one handle per host thread, as type_sed is per PET).
 *   - a handle belongs to the device that is current when sedcol_create is
 *     called (hipSetDevice beforehand selects it); every call runs on that
 *     device and leaves the caller's current device as it found it.
 */
#ifndef SEDCOL_H
#define SEDCOL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SEDCOL_NVAR 11 /* state variables, order below */
#define SEDCOL_KMAX 64 /* layers per column */
#define SEDCOL_NBUD 16 /* budget entries per column */

/* state variable order (all per m3 pore water, mmol) */
enum sedcol_var {
  SEDCOL_LDETC = 0, SEDCOL_SDETC, SEDCOL_LDETN, SEDCOL_SDETN, SEDCOL_DETP,
  SEDCOL_O2, SEDCOL_NO3, SEDCOL_NH4, SEDCOL_ODU, SEDCOL_PO4, SEDCOL_ADSP
};

/* time-stepping methods.  0..3 = solver_library.F90:32-35. */
enum sedcol_method {
  SEDCOL_EULER = 0,
  SEDCOL_RUNGE_KUTTA_4 = 1,
  SEDCOL_ADAPTIVE_EULER = 2, /* global x0.25 subcycling, :104-140 */
  SEDCOL_RUNGE_KUTTA_4_38 = 3,
  SEDCOL_EMP1 = 4,           /* extended modified Patankar, 1st order */
  SEDCOL_EMP2 = 5,           /* extended modified Patankar, 2nd order */
  SEDCOL_EMP_ADAPTIVE = 6,   /* EMP-1/EMP-2 pair, per-column step size */
  SEDCOL_ROS2_ADAPTIVE = 7   /* linearly implicit ROS2 W-method, per-column
                                step size (sedcol_params.adaptive_method) */
};

enum sedcol_status {
  SEDCOL_OK = 0,
  SEDCOL_EINVAL = 1,   /* bad argument */
  SEDCOL_EDOMAIN = 2,  /* porosity/grid outside the valid domain
                          (fabm_sediment_driver.F90:503-530 'stop's) */
  SEDCOL_ENAN = 3,     /* NaN in the state (fabm_sediment_component.F90:1718-1723) */
  SEDCOL_EHIP = 4,     /* HIP runtime error */
  SEDCOL_ELIMIT = 5,   /* adaptive attempt limit reached */
  SEDCOL_ENOMEM = 6,
  SEDCOL_EUNSUPPORTED = 7
};

enum sedcol_layout { SEDCOL_LAYOUT_FORTRAN = 0, SEDCOL_LAYOUT_PACKED = 1 };

/* Arithmetic form of the right-hand side (sedcol_params.arith).
 * REFERENCE: the operation order of fabm_sediment_driver.F90:652-692,776-820
 *            (divisions by dzc, (1-phi)dz and phi as written there).
 * FAST:      the same mathematics with the geometry divisions folded into
 *            per-layer coefficients D/dzc and 1/(phi dz) computed once per
 *            forcing record, and merged Monod quotients; results differ from
 *            REFERENCE by rounding only.  Requires bcdown != 2 (REFERENCE is
 *            used otherwise). */
enum sedcol_arith { SEDCOL_ARITH_REFERENCE = 0, SEDCOL_ARITH_FAST = 1 };

typedef struct sedcol_handle sedcol_handle;

/* Domain description, the counterpart of type_rhs_driver's
 * inum/jnum/knum/nvar/mask (solver_library.F90:37-49) and fabm_sed_grid
 * (fabm_sediment_driver.F90:41-56). */
typedef struct {
  int32_t ncol;            /* columns (I*J) */
  int32_t kmax;            /* K for a uniform grid, max K otherwise (<= SEDCOL_KMAX) */
  int32_t nvar;            /* must equal SEDCOL_NVAR */
  int32_t layout;          /* enum sedcol_layout */
  const int32_t *nlayers;  /* PACKED: layers per column (ncol); NULL = all kmax */
  const int32_t *mask;     /* ncol, nonzero = land (never integrated); NULL = all wet */
} sedcol_desc;

/* Parameters.  Transport keys follow &sed_nml and type_sed
 * (fabm_sediment_driver.F90:73-89,211-233); the solver keys type_rhs_driver
 * (solver_library.F90:40-41); kinetic rates are SI (s-1, mmol m-3 s-1). */
typedef struct {
  double D0;                  /* molecular diffusivity at 0 degC, cm2 d-1 ('diffusivity') */
  double dt_min;              /* ADAPTIVE_EULER floor (solver_library.F90:40) */
  double relative_change_min; /* ADAPTIVE_EULER criterion (:41) */
  double rL, rS, CprodMax;    /* rLabile, rSemilabile (s-1), CprodMax (mmol m-3 s-1) */
  double rnit, ksO2nitri, rODUox, ksO2oduox, ksO2oxic;
  double ksNO3denit, kinO2denit, kinNO3anox, kinO2anox;
  double rPads, PAdsODU;      /* P adsorption rate (s-1), release half-saturation */
  double emp_rtol, emp_atol;  /* adaptive EMP error tolerances */
  double emp_dt_min, emp_dt_max;
  int32_t bcup_dis;           /* bcup_dissolved_variables: 1 flux, 2 Dirichlet, 3 closed */
  int32_t bcup_part;          /* bcup_particulate_variables: 1 flux, 2 Dirichlet, 3 closed,
                                 4 capped flux (sedcol_set_flux_cap) */
  int32_t bcdown;             /* 1, 2 or 3 (reference get_rhs uses 3) */
  int32_t clamp;              /* clamp to the minimum 0 after each fixed step, as
                                 fabm_sediment_component.F90:1726-1731 does */
  int32_t arith;              /* enum sedcol_arith */
  int32_t emp_weights;        /* Patankar factor of the EMP steppers: 0 the
                                 product over the destroyed components
                                 (Bruggeman et al. 2007), 1 their geometric
                                 mean (the same properties; long columns are
                                 not throttled by their component count) */
  int32_t adaptive_method;    /* stepper of sedcol_integrate / sedcol_spinup:
                                 0 or SEDCOL_EMP_ADAPTIVE the adaptive EMP pair,
                                 SEDCOL_ROS2_ADAPTIVE the linearly implicit
                                 ROS2 W-method (reaction Jacobian per cell times
                                 the per-species transport factor, one
                                 correction; positivity by step rejection) */
} sedcol_params;

/* Per-column diagnostics (sedcol_get_diagnostics), the per-column form of
 * last_min_dt / last_min_dt_grid_cell (solver_library.F90:42-43,130-135). */
typedef struct {
  int64_t accepted;   /* accepted steps */
  int64_t rejected;   /* rejected adaptive attempts */
  int64_t forced;     /* accepted at emp_dt_min with error above tolerance */
  int64_t zero_sink;  /* destruction terms of exhausted components (EMP) */
  int64_t newton;     /* Newton iterations of the Patankar factor (EMP) */
  double dt_last;     /* last accepted step */
  double dt_next;     /* step size carried into the next call */
  double dt_minacc;   /* smallest accepted step */
  double p_min;       /* smallest Patankar factor */
  int32_t status;     /* 0 ok, 1 NaN, 2 attempt limit */
  int32_t pad;
} sedcol_coldiag;

/* Schedule of adaptive integrations (sedcol_integrate).  No reference
 * counterpart: the reference steps the whole domain with one dt per
 * ode_solver call (solver_library.F90:80-189).  The default {0, 0, 1, 0}
 * integrates each layer bin in one persistent launch.  Results are identical
 * for every schedule. */
typedef struct {
  int32_t window;          /* attempts per column per launch; 0 = no limit */
  int32_t graph_windows;   /* window > 0: windows captured in one hipGraph and
                              replayed until every column is done; 0 = launch
                              windows directly (one launch per step with
                              window = 1) */
  int32_t check_every;     /* direct windows launched between completion checks (>= 1) */
  int32_t concurrent_bins; /* 1 = the layer bins run as concurrent branches */
} sedcol_schedule;

/* Kernel shape of the adaptive kernel per layer bin (bins hold the wet
 * columns with K <= 16, K <= 32 and K <= 64).  layers_per_lane: 1, or 2 for
 * adaptive EMP with product weights in the reference or fast form (other
 * parameter sets run such a bin with one layer per lane); 0 reads as 1.
 * lanes_per_column: 16, 32 or 64, at least the bin's largest K divided by
 * layers_per_lane; workgroup = lanes_per_column * columns_per_group threads,
 * 64, 128 or 256.  Results are identical for every shape. */
#define SEDCOL_NBIN 3
typedef struct {
  int32_t lanes_per_column[SEDCOL_NBIN];
  int32_t columns_per_group[SEDCOL_NBIN];
  int32_t layers_per_lane[SEDCOL_NBIN];
} sedcol_tuning;

/* --- lifecycle ---------------------------------------------------------- */

/* Replaces type_sed allocation in initialize (fabm_sediment_driver.F90:191-368)
 * and the conc allocation of its callers (fabm_sediment_component.F90:521,
 * examples/standalone/omexdia_p/main.F90:75-80).  The handle belongs to the
 * device current at this call (hipSetDevice beforehand selects it). */
int sedcol_create(const sedcol_desc *desc, sedcol_handle **out);
int sedcol_destroy(sedcol_handle *h);
const char *sedcol_last_error(void);

/* --- set-up ------------------------------------------------------------- */

/* Replaces the &sed_nml read and type_sed scalars (fabm_sediment_driver.F90:
 * 211-250) and the FABM parameter set (fabm_sed.nml:51-77). */
int sedcol_set_params(sedcol_handle *h, const sedcol_params *p);

/* Replaces init_grid (:127-177), the porosity / bioturbation set-up of
 * initialize (:278-313) and update_porosity (:393-442): the caller supplies
 * dz (layer heights), dzc (distance to the layer above, value at the top
 * layer ignored; the reference's dzc(k-1) is stored at layer k), phi
 * (porosity), phi_intf (interface porosity), bioturb_factor (at the upper
 * interface of each layer) and the per-column bioturbation coefficient
 * (cm2 d-1).  Validates as fabm_sed_check_domain (:488-545) and returns
 * SEDCOL_EDOMAIN instead of stopping. */
int sedcol_set_geometry(sedcol_handle *h, const double *dz, const double *dzc,
                        const double *phi, const double *phi_intf,
                        const double *bioturb_factor, const double *bioturbation);

/* Replaces bdys(:,:,1) (temperature), bdys(:,:,n+1) (bottom-water
 * concentrations) and fluxes(:,:,n) (prescribed top fluxes, positive
 * downward) as assembled by get_boundary_conditions
 * (fabm_sediment_component.F90:1865-2030).  Forcing stays constant until the
 * next call. */
int sedcol_set_forcing(sedcol_handle *h, const double *temp, const double *c_bw,
                       const double *flux_top);

/* A piecewise-constant forcing series of nrec records, record r applying on
 * [t0 + r*dt_rec, t0 + (r+1)*dt_rec); temp is (nrec, ncol), c_bw
 * (nrec, nvar, ncol), flux_top (nrec_flux, nvar, ncol) with nrec_flux = nrec
 * or 1 (one flux record for the whole series).  Used by sedcol_integrate.
 * No reference counterpart: the reference re-reads forcing every coupling
 * step (fabm_sediment_component.F90:1665); this lets one launch span many
 * records. */
int sedcol_set_forcing_series(sedcol_handle *h, int32_t nrec, double t0, double dt_rec,
                              const double *temp, const double *c_bw,
                              const double *flux_top, int32_t nrec_flux);

/* State in / out (conc pointer association, fabm_sediment_component.F90:521,
 * :1666-1667; init_concentrations :449-481 is the caller's). */
int sedcol_set_state(sedcol_handle *h, const double *conc);
int sedcol_get_state(sedcol_handle *h, double *conc);

/* Warm start on the device (no reference counterpart: it replaces the state
 * that the presimulation of fabm_sediment_component.F90:557-632 would
 * produce, with profiles predicted by a model such as the sedsurr network,
 * without a round trip through host memory).  prof: device memory of the
 * handle's device, one row per column (stride ld >= nvar*nz floats); row c
 * holds u = prof[c*ld + n*nz + j] for variable n at node j, and the profile
 * is y = u*scale[n*nz+j] + offset[n*nz+j], the log10 of (concentration +
 * eps[n]).  The nodes are the layer centres of an nz-layer geometric grid
 * (top layer dz1 m, constant growth) of the column's total depth: depth[c]
 * (m), or the column's own sum of dz when depth is NULL.  Every layer of a
 * wet column gets max(10^y(z) - eps[n], 0) at its own centre z, with y
 * linear in depth between nodes and constant beyond the end nodes (NaN
 * stays NaN).  Land columns keep their state.  2 <= nz <= 64; scale,
 * offset (nvar*nz), eps (nvar) and depth (ncol) are host memory. */
int sedcol_set_state_profiles(sedcol_handle *h, const float *prof, int64_t ld, int32_t nz, const double *scale,
                              const double *offset, const double *eps, double dz1, const double *depth);

/* --- hot path ----------------------------------------------------------- */

/* get_rhs (fabm_sediment_driver.F90:575-717): rhs of conc under the current
 * forcing; flux_top_out (nvar, ncol, may be NULL) = intFlux(:,:,1) as stored
 * into fluxes(:,:,n) at :692.  Masked columns return rhs = 0.  With a
 * forcing series (nrec > 1) active there is no single current forcing:
 * SEDCOL_EINVAL. */
int sedcol_rhs(sedcol_handle *h, const double *conc, double *rhs, double *flux_top_out);

/* sedcol_rhs without allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device or host memory during the call (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as sedcol_rhs, bitwise; it
 * calls no hipMalloc/hipFree and no allocating routine.  Its staging lives in
 * a fixed 64 MiB workspace per device (held from library load until the
 * process ends), so the domain must fit: (2 NV ncell + NV ncol) doubles in
 * the packed layout, (3 NV ncell + NV ncol) in the Fortran layout, i.e. up
 * to about 380 000 (packed) or 250 000 (Fortran) cells; larger domains return
 * SEDCOL_EINVAL.  Every call on one GPU shares that workspace and it keeps
 * the last call's contents (each call fills what it reads): the library runs
 * one such call per device at a time (a lock), so calls from several host
 * threads are serialised; nothing else may touch the workspace. */
int sedcol_rhs_noalloc(sedcol_handle *h, const double *conc, double *rhs, double *flux_top_out);

/* ode_solver(rhs_driver, dt, method) (solver_library.F90:80-189) on the
 * caller's conc, updated in place (the drop-in call used at
 * fabm_sediment_component.F90:1715 and main.F90:123).  flux_top_out = top
 * fluxes of the last right-hand side evaluated (NULL allowed); the clamp of
 * sedcol_params.clamp is applied after the step, as the callers do. */
int sedcol_step(sedcol_handle *h, double dt, int32_t method, double *conc_inout,
                double *flux_top_out);

/* nsteps fixed steps (methods 0..5) of dt on the device-resident state
 * (sedcol_set_state), steps_per_launch steps per kernel launch (0 = all in
 * one persistent launch).  Equivalent to nsteps ode_solver calls plus the
 * caller's clamp, the Run loop of fabm_sediment_component.F90:1700-1769.
 * ADAPTIVE_EULER runs in one workgroup when the wet columns fit one (at
 * most 32, 16 or 8 columns of up to 16, 32 or 64 layers), in launches of
 * 4096 steps of at most 2^22 attempts each (SEDCOL_ELIMIT beyond), and
 * otherwise as separate launches per attempt; steps_per_launch does not
 * apply to it.  Both give the same results. */
int sedcol_run(sedcol_handle *h, double dt, int32_t nsteps, int32_t method,
               int32_t steps_per_launch);

/* Adaptive EMP (method 6) on the device-resident state from t_start to
 * t_end under the forcing series (or the current forcing), each column with
 * its own step size; persistent launches by the handle's schedule
 * (sedcol_set_schedule). */
int sedcol_integrate(sedcol_handle *h, double t_start, double t_end);

/* Schedule and kernel shape of sedcol_integrate (no reference counterpart). */
int sedcol_set_schedule(sedcol_handle *h, const sedcol_schedule *s);
int sedcol_get_schedule(sedcol_handle *h, sedcol_schedule *s);
int sedcol_set_tuning(sedcol_handle *h, const sedcol_tuning *t);
int sedcol_get_tuning(sedcol_handle *h, sedcol_tuning *t);
/* Autotuner (no reference counterpart): times sedcol_integrate over
 * [t_start, t_start + t_span) for every supported shape of every bin, best
 * of `repeats` runs, from the current state, which is restored after every
 * trial and at the end.  The fastest shape per bin is installed and returned
 * in *best; ms_table (optional, SEDCOL_NBIN x 2 x 3 x 3 doubles: bin,
 * layers per lane 1/2, lanes 16/32/64, workgroup 64/128/256) receives every
 * time, -1 where a shape does not apply. */
int sedcol_autotune(sedcol_handle *h, double t_start, double t_span, int32_t repeats,
                    sedcol_tuning *best, double *ms_table);

/* Per-column spin-up to steady state, the replacement of the single-column
 * presimulation plus broadcast of fabm_sediment_component.F90:557-632 (the
 * reference integrates one column for presimulation_years with a fixed dt of
 * 3600 s and copies it to every wet column).  Every wet column is integrated
 * by the adaptive stepper that sedcol_params.adaptive_method selects
 * (adaptive EMP, or the linearly implicit method 7, the faster choice near
 * steady state) from its resident state under the current forcing, in
 * chunks of `chunk` seconds from t_start, until its steady-state residual
 * (sedcol_residual with cfloor) is below tol or t_max seconds have passed.
 * A column is retired from the integration as soon as it is converged.
 * flux_ref (nvar x ncol, optional): with it, a column also needs its top
 * fluxes F within flux_rtol * max(|flux_ref|, flux_floor[c]) of flux_ref
 * for every variable to count as converged.
 * t_conv (ncol, optional): time after t_start at which the column converged,
 * -1 if it did not (and for land columns).  resid (ncol, optional): final
 * residual.  *n_unconverged (optional): wet columns still above tol.
 * Runs under one forcing record: SEDCOL_EINVAL while a series (nrec > 1)
 * is active.  No reference counterpart for the per-column form. */
int sedcol_spinup(sedcol_handle *h, double t_start, double chunk, double t_max, double tol,
                  const double *cfloor, const double *flux_ref, double flux_rtol,
                  const double *flux_floor, double *t_conv, double *resid,
                  int32_t *n_unconverged);

/* Bioturbation profile 3 (Zhang & Wirtz 2017), fabm_sediment_driver.F90:
 * 618-645 with the &sed_nml keys of :204-231: on every right-hand side the
 * bioturbation factor follows the infauna biomass implied by the weighted TOC
 * profile of the labile and semi-labile detritus carbon (:380-386); the
 * bioturbation coefficient and its temperature factor are 1 (:622-623).  It
 * replaces the factor of sedcol_set_geometry until called with NULL, and
 * runs in the reference arithmetic form.  Detritus carbon must be positive
 * (the reference divides by the TOC; zero gives NaN, SEDCOL_ENAN). */
typedef struct {
  double k_l;         /* bioturb_k_l (0.11) */
  double L1, L2;      /* bioturb_L1, bioturb_L2 (0.2, 0.6) */
  double beta, b;     /* bioturb_beta, bioturb_b (0.22, 1.334) */
  double dry_density; /* bioturb_dry_density, kg m-3 (1000) */
} sedcol_bioturbation_zw;
int sedcol_set_bioturbation_zw(sedcol_handle *h, const sedcol_bioturbation_zw *zw);

/* Distributed-POM flux cap per cell (mmol m-2 s-1), the flux_cap of
 * fabm_sediment_driver.F90:284-291, needed when bcup_part = 4
 * (distributed_pom_flux, :239-243): the particulate top flux beyond the
 * first layer's cap passes to the interfaces below (diff3d :791-802).
 * Cell layout as sedcol_set_geometry. */
int sedcol_set_flux_cap(sedcol_handle *h, const double *flux_cap);

/* Boundary assembly on the device, the counterpart of get_boundary_conditions
 * (fabm_sediment_component.F90:1865-2030) for forcing record 0:
 *   temperature <- temp_surface (ncol; NULL = unchanged);
 *   for each variable n with present[n] != 0 (the `<name>_at_soil_surface`
 *   field exists):
 *     particulate: top flux = -c_surface * w_surface (w: z-velocity, m s-1,
 *                  negative downward; flux positive downward, :1978-1986);
 *     dissolved:   bottom-water value = c_surface; top flux =
 *                  -(c_top - c_surface) / dz_top * (B + D0 + 0.035 T)
 *                  * phi_top / 86400 / 10000 with bcup_dis = 1 (:2012-2014,
 *                  from the resident state's top layer), else 0 (:2021).
 * Variables without a field keep their forcing.  c_surface, w_surface are
 * (nvar, ncol); flux_out (nvar, ncol, optional) receives the top fluxes.
 * Not allowed while a forcing series (nrec > 1) is set. */
int sedcol_assemble_forcing(sedcol_handle *h, const double *temp_surface, const double *c_surface,
                            const double *w_surface, const int32_t *present, double *flux_out);

/* Per-cell release rates of the resident state under the current forcing
 * (cell layout as sedcol_set_geometry; mmol m-3 pore water s-1): DIC from
 * mineralisation (Cprod) and N2-N from denitrification (0.8 Denit), the
 * counterpart of the diagnostic variables the reference reads from FABM
 * after get_rhs (fabm_get_bulk_diagnostic_data, fabm_sediment_driver.F90:
 * 562; hzg_omexdia_p_denit in the output of examples/standalone/omexdia_p/
 * main.F90:109-157).  Land cells are 0.  SEDCOL_EINVAL while a forcing
 * series is active. */
int sedcol_rates(sedcol_handle *h, double *dic_release, double *n2_release);

/* --- restart ------------------------------------------------------------ */

/* Checkpoint of the device-resident integration state: state, budget
 * accumulators, top fluxes and per-column diagnostics (including the step
 * size carried into the next call), as one host byte blob of
 * sedcol_checkpoint_size bytes.  Loading it into a handle of the same
 * domain restores the state exactly, so an integration continued from it
 * gives the same results as one never interrupted.  The counterpart of the
 * restart read of fabm_sediment_component.F90:1328-1489, which restores the
 * state only. */
int sedcol_checkpoint_size(sedcol_handle *h, int64_t *bytes);
int sedcol_save_checkpoint(sedcol_handle *h, void *buf, int64_t bytes);
int sedcol_load_checkpoint(sedcol_handle *h, const void *buf, int64_t bytes);

/* --- outputs ------------------------------------------------------------ */

int sedcol_get_fluxes(sedcol_handle *h, double *flux_top_out); /* (nvar, ncol) */
/* budget (SEDCOL_NBUD, ncol) with entries fastest: 0..10 time-integrated top
 * flux per variable, 11 integrated DIC release, 12 integrated N2 release,
 * 13..15 mass added by the minimum clamp for C, N, P (mmol m-2). */
int sedcol_get_budget(sedcol_handle *h, double *budget);
int sedcol_reset_budget(sedcol_handle *h);
/* inventory (nvar, ncol): sum_k phi_k dz_k c_k (mmol m-2) of the resident state */
int sedcol_get_inventory(sedcol_handle *h, double *inventory);
int sedcol_get_diagnostics(sedcol_handle *h, sedcol_coldiag *diag); /* ncol entries */
int sedcol_reset_diagnostics(sedcol_handle *h);
/* steady-state residual per column: max |rhs| / max(|c|, cfloor[n]) (s-1);
 * NaN when any term is NaN.  SEDCOL_EINVAL while a forcing series is active. */
int sedcol_residual(sedcol_handle *h, const double *cfloor, double *resid);
/* last_min_dt of the reference ADAPTIVE_EULER (solver_library.F90:42,130-135) */
int sedcol_get_last_min_dt(sedcol_handle *h, double *last_min_dt);
int sedcol_synchronize(sedcol_handle *h);

/* timing of the last sedcol_run / sedcol_integrate measured with HIP events
 * on the handle's stream: total milliseconds and number of kernel launches */
int sedcol_last_timing(sedcol_handle *h, double *ms, int32_t *launches);

#ifdef __cplusplus
}
#endif
#endif
