// sedcol_kernels.hpp -- device data view and kernel launchers.
#pragma once

#include "sedcol_device.hpp"

namespace sc {

// Device-resident domain.  Cells are packed column by column (top layer
// first, column c owns off[c]..off[c+1]-1); cell arrays have ncell entries,
// state and cell accumulators are variable-major (x[n*ncell + cell]);
// per-column-per-variable arrays are (n*ncol + c).  Forcing is a series of
// nrec piecewise-constant records: temp (nrec, ncol), cbw and fup
// (nrec, NV, ncol).
struct Dev {
  int ncol;
  long long ncell;
  const int *off;
  const double *dz, *dzc, *phi, *phiI, *bfac, *biot;
  const double *fcap;  // ncell or null: distributed-POM flux cap (bcup_part = 4)
  ZwDev zw;            // bioturbation profile 3
  int nrec;
  double t0, dt_rec;
  const double *temp, *cbw, *fup;
  long long fup_rec_stride;  // NV*ncol, or 0 when one flux record serves all
  double *conc;     // NV*ncell
  // Budget sums, each followed by its Neumaier compensation (total = sum + comp):
  double *cellacc;  // 10*ncell: DIC, N2, clamp C, clamp N, clamp P (per m3 pore water), then comps
  double *colacc;   // 2*NV*ncol: time-integrated top flux per variable, then comps
  double *flux;     // NV*ncol: top flux of the last right-hand side
  sedcol_coldiag *diag;
  // adaptive integration position of each column between windows
  int *wrec;        // ncol: current forcing record (past the last one: done)
  double *wrem;     // ncol: time left in that record
  long long *watt;  // ncol: attempts since the start of the span
};

// Span of one adaptive integration, read by the kernel through a pointer
struct EmpCtl {
  double t_start, t_end;
  long long max_attempts;
  int r0, r1;   // first and last forcing record touched by the span
  int window;   // attempts per column per launch; <= 0: no limit
  int resume;   // 1: continue every column from d.wrec / d.wrem / d.watt
};

// rhs (NV*ncell), top fluxes (NV*ncol) and the per-cell DIC and N2 release
// rates (2*ncell: sinks[cell], sinks[ncell + cell]) of conc; any output may
// be null
void launch_rhs(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, int rec,
                const double *conc, double *rhs, double *flux, double *sinks = nullptr);
void launch_fixed(int lpc, int method, hipStream_t st, const Dev &d, const int *list, int n,
                  const sedcol_params &p, int rec, double dt, int nsteps, int *errflag);
// Adaptive EMP over the span in *d_ctl for the n columns of list, lpc lanes
// per column and blk threads per workgroup.  next: a device int set to 0
// before the launch (work-queue counter); pending: incremented once per
// column stopped by the window limit; work[c] receives the attempts column c
// has taken since the start of the span.  d_dev / d_params: device copies of
// the domain view and the parameters (p is the host copy, used to pick the
// kernel instance).  lpl: layers per lane, 1 (k_emp_adaptive / the
// linearly implicit kernel) or 2 (k_emp_adaptive2: adaptive EMP with product
// weights in the reference or fast form, lpc >= 16).  Returns false for an
// unsupported (lpc, blk, lpl).
bool launch_emp_adaptive(int lpc, int blk, int lpl, hipStream_t st, const Dev *d_dev, const int *list, int n,
                         const sedcol_params &p, const sedcol_params *d_params, const EmpCtl *d_ctl, int *next,
                         int *pending, int *work, int *errflag);
// blocks of that kernel resident at once on the current device (0: shape
// not supported)
int emp_resident_blocks(int lpc, int blk, int lpl, const sedcol_params &p);
// whether parameters p can run with two layers per lane
bool emp2_supported(const sedcol_params &p);
// start position of every listed column: record r0 - 1, nothing left in it,
// no attempts
void launch_emp_begin(hipStream_t st, const int *list, int n, int r0, int *wrec, double *wrem, long long *watt);
// depth terms of bioturbation profile 3 for the listed columns (zw must
// point to the device arrays e1, e2 (ncell) and cum (ncol) to fill)
void launch_zw_geometry(hipStream_t st, const Dev &d, const int *list, int n, double k_l, double *e1, double *e2,
                        double *cum);
// warm start of the listed columns (sedcol_set_state_profiles): prof row c
// (stride ld floats) holds the normalised log10 profile u of variable n at
// node j in prof[c ld + n nz + j]; y = u scale[n nz + j] + offset[n nz + j]
// on the centres of an nz-layer geometric grid from dz1 to the column's
// depth (depth[c], or the column's sum of dz when null); every layer gets
// max(10^y(z) - eps[n], 0) at its centre, y linear in depth between nodes
// and constant beyond the end nodes.  scale, offset, eps, depth: device.
constexpr int kWarmMaxNodes = 64;
struct WarmProfiles {
  const float *prof;
  long long ld;
  int nz;
  double dz1;
  const double *scale, *offset, *eps, *depth;
};
void launch_warm_profiles(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const WarmProfiles &w);
// v into p[0..n-1] (a kernel, so that it can be captured in a graph)
void launch_fill_int(hipStream_t st, int *p, int n, int v);
void launch_fill_dbl(hipStream_t st, double *p, long long n, double v);
// state of a uniform grid between the Fortran layout (ncol, K, NV) and the
// packed one (NV, ncol*K): src Fortran -> dst packed when to_packed, else
// src packed -> dst Fortran
void launch_layout(hipStream_t st, const double *src, double *dst, int ncol, int K, bool to_packed);
// boundary conditions of the listed columns from the fields at the soil
// surface into forcing record 0 (present: bit n set when variable n's
// field exists)
void launch_assemble(hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, unsigned present,
                     const double *ts, const double *cs, const double *ws, double *temp, double *cbw, double *fup);
void launch_rhs_store(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, int rec,
                      double *r_out, double *s_out);
void launch_ae_check(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double dt_red, double rcm,
                     const double *r_in, int *viol);
void launch_ae_commit(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double dt_red,
                      const double *r_in, const double *s_in);
// ADAPTIVE_EULER steps of a domain whose wet columns all sit in one bin of
// at most kAeFusedBlock / lpc columns, in one workgroup (no host round trip
// per attempt); errflag bit 1 NaN, bit 2 attempt limit; *min_dt (device) the
// smallest step taken
constexpr int kAeFusedBlock = 512;
constexpr int kAeFusedMaxAttempts = 1 << 22;
// (d_dev, d_params: device copies of the domain view and of p)
void launch_ae_fused(int lpc, hipStream_t st, const Dev *d_dev, const int *list, int n, const sedcol_params &p,
                     const sedcol_params *d_params, int rec, double dt, int nsteps, int *errflag, double *min_dt);
void launch_clamp(int lpc, hipStream_t st, const Dev &d, const int *list, int n, int *errflag);
void launch_budget(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double *budget);
void launch_inventory(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double *inv);
void launch_residual(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, int rec,
                     const double *cfloor, double *resid, double *flux = nullptr);

}  // namespace sc
