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
  int nrec;
  double t0, dt_rec;
  const double *temp, *cbw, *fup;
  long long fup_rec_stride;  // NV*ncol, or 0 when one flux record serves all
  double *conc;     // NV*ncell
  double *cellacc;  // 5*ncell: DIC, N2, clamp C, clamp N, clamp P (per m3 pore water)
  double *colacc;   // NV*ncol: time-integrated top flux
  double *flux;     // NV*ncol: top flux of the last right-hand side
  sedcol_coldiag *diag;
};

void launch_rhs(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, int rec,
                const double *conc, double *rhs, double *flux);
void launch_fixed(int lpc, int method, hipStream_t st, const Dev &d, const int *list, int n,
                  const sedcol_params &p, int rec, double dt, int nsteps, int *errflag);
// next: a device int set to 0 before the launch (work-queue counter);
// work[c] receives the attempts column c took in this launch
// d_dev / d_params: device copies of the domain view and the parameters
// (p is the host copy, used to pick the kernel instance)
void launch_emp_adaptive(int lpc, hipStream_t st, const Dev *d_dev, const int *list, int n, const sedcol_params &p,
                         const sedcol_params *d_params, double t_start, double t_end, int r0, int r1,
                         long long max_attempts, int *next, int *work, int *errflag);
void launch_rhs_store(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, int rec,
                      double *r_out, double *s_out);
void launch_ae_check(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double dt_red, double rcm,
                     const double *r_in, int *viol);
void launch_ae_commit(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double dt_red,
                      const double *r_in, const double *s_in);
void launch_clamp(int lpc, hipStream_t st, const Dev &d, const int *list, int n, int *errflag);
void launch_budget(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double *budget);
void launch_inventory(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double *inv);
void launch_residual(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, int rec,
                     const double *cfloor, double *resid);

}  // namespace sc
