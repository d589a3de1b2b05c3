// sedcol_kernels.hip -- gfx950 kernels of the sediment-column hot path.
//
// One column per lane group (LPC lanes, lane = layer, all 11 state
// variables in VGPRs); a 256-thread block holds 256/LPC columns.  Wet columns
// are binned by K into LPC = 16/32/64 lists so a group never wastes more than
// half its lanes.  Fixed-step methods keep the state in registers for a
// whole window of steps; the adaptive EMP kernel keeps it for the whole
// integration span and every forcing record inside it.
//
// Top fluxes are produced on lane 0 (one value per variable and stage); they
// are staged in LDS and accumulated by lanes 0..NV-1, lane n owning
// variable n, so no lane carries per-variable flux registers.
#include <atomic>

#include "sedcol_kernels.hpp"

namespace sc {

constexpr int BLOCK = 256;

// ---------------------------------------------------------------------------
// per-lane set-up
// ---------------------------------------------------------------------------

struct Geo {
  double phiI, bfac, B;
};

__device__ __forceinline__ void setup_lane(const Dev &d, int c, int l, Lane &L, Geo &G, long long &cell)
{
  const int o = d.off[c];
  L.l = l;
  L.K = d.off[c + 1] - o;
  L.active = l < L.K;
  cell = (long long)o + l;
  L.dz = L.active ? d.dz[cell] : 1.0;
  L.dzc = (L.active && l > 0) ? d.dzc[cell] : 1.0;
  L.phi = L.active ? d.phi[cell] : 0.5;
  G.phiI = L.active ? d.phiI[cell] : 0.5;
  L.cap = (L.active && d.fcap) ? d.fcap[cell] : 0.0;
  L.ze1 = (L.active && d.zw.on) ? d.zw.e1[cell] : 1.0;
  L.ze2 = (L.active && d.zw.on) ? d.zw.e2[cell] : 1.0;
  L.zcum = d.zw.on ? d.zw.cum[c] : 1.0;
  G.bfac = L.active ? d.bfac[cell] : 0.0;
  G.B = d.biot[c];
}

// load forcing record r of column c: bottom-water concentrations and top
// fluxes to the group's LDS slot, temperature-dependent coefficients to L
template <int ARITH>
__device__ __forceinline__ void load_forcing(const Dev &d, const sedcol_params &p, int c, int r, const Geo &G,
                                             Lane &L, double *sbc)
{
  if (L.l < NV) {
    sbc[L.l] = d.cbw[((long long)r * NV + L.l) * d.ncol + c];
    sbc[NV + L.l] = d.fup[(long long)r * d.fup_rec_stride + (long long)L.l * d.ncol + c];
  }
  const double T = d.temp[(long long)r * d.ncol + c];
  diffusivities<ARITH>(p, G.B, T, G.phiI, G.bfac, L);
}

template <int LPC>
__device__ __forceinline__ bool group_any(bool b)
{
  return lg_max<LPC>(b ? 1.0 : 0.0) > 0.0;
}

// ---------------------------------------------------------------------------
// single right-hand side (get_rhs) for parity and diagnostics
// ---------------------------------------------------------------------------

template <int LPC, int ARITH>
__global__ __launch_bounds__(BLOCK) void k_rhs(Dev d, const int *list, int nlist, sedcol_params p, int rec,
                                               const double *conc, double *rhs, double *flux, double *sinks)
{
  constexpr int G = BLOCK / LPC;
  __shared__ double sbc[G][2 * NV];
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int slot = blockIdx.x * G + g;
  const bool valid = slot < nlist;
  const int c = valid ? list[slot] : 0;
  Lane L;
  Geo Gm;
  long long cell = 0;
  if (valid) {
    setup_lane(d, c, l, L, Gm, cell);
    load_forcing<ARITH>(d, p, c, rec, Gm, L, sbc[g]);
  }
  __syncthreads();
  if (!valid) return;
  double x[NV];
#pragma unroll
  for (int n = 0; n < NV; ++n) x[n] = L.active ? conc[n * d.ncell + cell] : 0.0;
  double sC, sN;
  column_rhs<LPC, ARITH>(p, L, sbc[g], x, sC, sN, [&](int n, double r, double F) {
    if (L.active && rhs) rhs[n * d.ncell + cell] = r;
    if (l == 0 && flux) flux[(long long)n * d.ncol + c] = F;
  }, &d.zw);
  if (L.active && sinks) {
    sinks[cell] = sC;
    sinks[d.ncell + cell] = sN;
  }
}

// ---------------------------------------------------------------------------
// fixed-step window: EULER, RK4, RK4 3/8, EMP-1, EMP-2
// ---------------------------------------------------------------------------

template <int LPC, int METHOD, int ARITH, int PW = 0>
__global__ __launch_bounds__(BLOCK) void k_fixed(Dev d, const int *list, int nlist, sedcol_params p, int rec,
                                                 double dt, int nsteps, int *errflag)
{
  constexpr int G = BLOCK / LPC;
  __shared__ double sbc[G][2 * NV];
  __shared__ double sF[G][4][NV];
  // coefficients of the geometric-mean Patankar solve (one slot otherwise)
  __shared__ double sgeo[PW == 1 ? NV : 1][BLOCK];
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int slot = blockIdx.x * G + g;
  const bool valid = slot < nlist;
  const int c = valid ? list[slot] : 0;
  Lane L;
  Geo Gm;
  long long cell = 0;
  if (valid) {
    setup_lane(d, c, l, L, Gm, cell);
    load_forcing<ARITH>(d, p, c, rec, Gm, L, sbc[g]);
  }
  __syncthreads();
  if (!valid) return;
  const double *bc = sbc[g];
  double(*F)[NV] = sF[g];
  const bool act = L.active;
  const bool own = l < NV;  // lane l accumulates the top flux of variable l

  double x[NV];
#pragma unroll
  for (int n = 0; n < NV; ++n) x[n] = act ? d.conc[n * d.ncell + cell] : 0.0;
  double fint = own ? d.colacc[(long long)l * d.ncol + c] : 0.0;
  double fintc = own ? d.colacc[(long long)(NV + l) * d.ncol + c] : 0.0;
  double flast = own ? d.flux[(long long)l * d.ncol + c] : 0.0;
  double acc[5], accc[5];
#pragma unroll
  for (int a = 0; a < 5; ++a) {
    acc[a] = act ? d.cellacc[a * d.ncell + cell] : 0.0;
    accc[a] = act ? d.cellacc[(5 + a) * d.ncell + cell] : 0.0;
  }
  sedcol_coldiag dg = d.diag[c];
  int64_t zs = 0;
  const double third = 1.0 / 3.0;

  for (int s = 0; s < nsteps; ++s) {
    double sC, sN;
    if (METHOD == SEDCOL_EULER) {
      double r0[NV];
      column_rhs<LPC, ARITH>(p, L, bc, x, sC, sN, [&](int n, double r, double Fv) {
        r0[n] = r;
        if (l == 0) F[0][n] = Fv;
      }, &d.zw);
#pragma unroll
      for (int n = 0; n < NV; ++n) x[n] = x[n] + dt * r0[n];
      if (own) {
        kadd(fint, fintc, dt * F[0][l]);
        flast = F[0][l];
      }
      kadd(acc[0], accc[0], dt * sC);
      kadd(acc[1], accc[1], dt * sN);
    } else if (METHOD == SEDCOL_RUNGE_KUTTA_4) {
      // solver_library.F90:142-163
      double cs[NV], k[NV];
      double aC, aN;
      column_rhs<LPC, ARITH>(p, L, bc, x, sC, sN, [&](int n, double r, double Fv) {
        k[n] = 0.5 * r;
        cs[n] = x[n] + (0.5 * dt) * r;
        if (l == 0) F[0][n] = Fv;
      }, &d.zw);
      aC = 0.5 * sC;
      aN = 0.5 * sN;
      column_rhs<LPC, ARITH>(p, L, bc, cs, sC, sN, [&](int n, double r, double Fv) {
        k[n] = k[n] + r;
        cs[n] = x[n] + (0.5 * dt) * r;
        if (l == 0) F[1][n] = Fv;
      }, &d.zw);
      aC = aC + sC;
      aN = aN + sN;
      column_rhs<LPC, ARITH>(p, L, bc, cs, sC, sN, [&](int n, double r, double Fv) {
        k[n] = k[n] + r;
        cs[n] = x[n] + dt * r;
        if (l == 0) F[2][n] = Fv;
      }, &d.zw);
      aC = aC + sC;
      aN = aN + sN;
      column_rhs<LPC, ARITH>(p, L, bc, cs, sC, sN, [&](int n, double r, double Fv) {
        k[n] = k[n] + 0.5 * r;
        if (l == 0) F[3][n] = Fv;
      }, &d.zw);
      aC = aC + 0.5 * sC;
      aN = aN + 0.5 * sN;
      const double w = dt * third;
#pragma unroll
      for (int n = 0; n < NV; ++n) x[n] = x[n] + w * k[n];
      if (own) {
        kadd(fint, fintc, w * (((0.5 * F[0][l] + F[1][l]) + F[2][l]) + 0.5 * F[3][l]));
        flast = F[3][l];
      }
      kadd(acc[0], accc[0], w * aC);
      kadd(acc[1], accc[1], w * aN);
    } else if (METHOD == SEDCOL_RUNGE_KUTTA_4_38) {
      // solver_library.F90:164-185
      double cs[NV], k1[NV], k2[NV];
      double s1C, s1N, s2C, s2N;
      column_rhs<LPC, ARITH>(p, L, bc, x, sC, sN, [&](int n, double r, double Fv) {
        k1[n] = r;
        cs[n] = x[n] + (third * dt) * r;
        if (l == 0) F[0][n] = Fv;
      }, &d.zw);
      s1C = sC;
      s1N = sN;
      column_rhs<LPC, ARITH>(p, L, bc, cs, sC, sN, [&](int n, double r, double Fv) {
        k2[n] = r;
        cs[n] = x[n] + dt * (r - third * k1[n]);
        if (l == 0) F[1][n] = Fv;
      }, &d.zw);
      s2C = sC;
      s2N = sN;
      column_rhs<LPC, ARITH>(p, L, bc, cs, sC, sN, [&](int n, double r, double Fv) {
        cs[n] = x[n] + dt * ((k1[n] - k2[n]) + r);
        k1[n] = (k1[n] + 3.0 * k2[n]) + 3.0 * r;
        if (l == 0) F[2][n] = Fv;
      }, &d.zw);
      double aC = (s1C + 3.0 * s2C) + 3.0 * sC;
      double aN = (s1N + 3.0 * s2N) + 3.0 * sN;
      column_rhs<LPC, ARITH>(p, L, bc, cs, sC, sN, [&](int n, double r, double Fv) {
        k1[n] = k1[n] + r;
        if (l == 0) F[3][n] = Fv;
      }, &d.zw);
      aC = aC + sC;
      aN = aN + sN;
      const double w = (dt * 1.0) / 8.0;
#pragma unroll
      for (int n = 0; n < NV; ++n) x[n] = x[n] + w * k1[n];
      if (own) {
        kadd(fint, fintc, w * (((F[0][l] + 3.0 * F[1][l]) + 3.0 * F[2][l]) + F[3][l]));
        flast = F[3][l];
      }
      kadd(acc[0], accc[0], w * aC);
      kadd(acc[1], accc[1], w * aN);
    } else {  // EMP-1 / EMP-2 with a fixed step
      double f0[NV];
      column_rhs<LPC, ARITH>(p, L, bc, x, sC, sN, [&](int n, double r, double Fv) {
        f0[n] = r;
        if (l == 0) F[0][n] = Fv;
      }, &d.zw);
      const double p1 = pfactor<LPC, false, PW, BLOCK>(x, f0, x, dt, act, zs, dg.newton, &sgeo[0][threadIdx.x]);
      if (p1 < dg.p_min) dg.p_min = p1;
      if (METHOD == SEDCOL_EMP1) {
#pragma unroll
        for (int n = 0; n < NV; ++n) x[n] = x[n] + (dt * p1) * f0[n];
        if (own) {
          kadd(fint, fintc, (dt * p1) * F[0][l]);
          flast = F[0][l];
        }
        kadd(acc[0], accc[0], (dt * p1) * sC);
        kadd(acc[1], accc[1], (dt * p1) * sN);
      } else {
        double c1[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) {
          const double y = x[n] + (dt * p1) * f0[n];
          c1[n] = y < 0.0 ? 0.0 : y;
        }
        const double s0C = sC, s0N = sN;
        column_rhs<LPC, ARITH>(p, L, bc, c1, sC, sN, [&](int n, double r, double Fv) {
          f0[n] = (f0[n] + r) * 0.5;
          if (l == 0) F[1][n] = Fv;
        }, &d.zw);
        const double p2 = pfactor<LPC, true, PW, BLOCK>(x, f0, c1, dt, act, zs, dg.newton, &sgeo[0][threadIdx.x]);
        if (p2 < dg.p_min) dg.p_min = p2;
#pragma unroll
        for (int n = 0; n < NV; ++n) x[n] = x[n] + (dt * p2) * f0[n];
        if (own) {
          kadd(fint, fintc, (dt * p2) * ((F[0][l] + F[1][l]) * 0.5));
          flast = F[1][l];
        }
        kadd(acc[0], accc[0], (dt * p2) * ((s0C + sC) * 0.5));
        kadd(acc[1], accc[1], (dt * p2) * ((s0N + sN) * 0.5));
      }
    }
    // clamp: Patankar rounding guard (EMP), or the callers' minimum clamp
    if (METHOD >= SEDCOL_EMP1 || p.clamp) {
#pragma unroll
      for (int n = 0; n < NV; ++n)
        if (x[n] < 0.0) {
          if (elem(n) >= 0) kadd(acc[2 + elem(n)], accc[2 + elem(n)], 0.0 - x[n]);
          x[n] = 0.0;
        }
    }
    bool nan = false;
#pragma unroll
    for (int n = 0; n < NV; ++n) nan |= act && (x[n] != x[n]);
    dg.accepted += 1;
    dg.dt_last = dt;
    if (dt < dg.dt_minacc) dg.dt_minacc = dt;
    if (group_any<LPC>(nan)) {
      dg.status = 1;
      if (l == 0) atomicOr(errflag, 1);
      break;
    }
  }

#pragma unroll
  for (int n = 0; n < NV; ++n)
    if (act) d.conc[n * d.ncell + cell] = x[n];
  if (own) {
    d.colacc[(long long)l * d.ncol + c] = fint;
    d.colacc[(long long)(NV + l) * d.ncol + c] = fintc;
    d.flux[(long long)l * d.ncol + c] = flast;
  }
#pragma unroll
  for (int a = 0; a < 5; ++a)
    if (act) {
      d.cellacc[a * d.ncell + cell] = acc[a];
      d.cellacc[(5 + a) * d.ncell + cell] = accc[a];
    }
  const double zsum = lg_sum<LPC>((double)zs);
  if (l == 0) {
    dg.zero_sink += (int64_t)zsum;
    d.diag[c] = dg;
  }
}

// ---------------------------------------------------------------------------
// adaptive EMP: per-column step size, persistent over [t_start, t_end)
// ---------------------------------------------------------------------------

// minimum waves per SIMD requested from the register allocator: 2 fits the
// adaptive step in 256 VGPRs without scratch (1 lets it take 256 + AGPRs,
// 3 spills about 340 bytes per lane)
#ifndef SEDCOL_EMP_MIN_WAVES
#define SEDCOL_EMP_MIN_WAVES 2
#endif

// Work-queue kernel: a grid sized to the resident capacity, each lane group
// taking the next column from *next (atomic counter) when it finishes one,
// so columns needing many steps do not leave other lanes idle.  Every column
// is integrated over [t_start, t_end) record by record (record r covers
// [t0 + r dt_rec, t0 + (r+1) dt_rec), the first and last extend outward),
// with its own step size carried in diag.dt_next.  Each group exits when the
// queue is empty, so every wave reaches the end.
//
// Windows: with ctl.window > 0 a column stops after that many attempts in
// one launch and saves its position (record, time left in the record,
// attempts so far) in d.wrec / d.wrem / d.watt; *pending counts the columns
// so stopped.  With ctl.resume = 1 every column continues from its saved
// position (k_emp_begin sets the start position) and columns whose record
// is past r1 are skipped.  A column's operations are the same however its
// attempts are split over launches, so windowed and single-launch runs give
// identical results.
template <int LPC, int ARITH, int BLK, int PW>
__global__ __launch_bounds__(BLK, SEDCOL_EMP_MIN_WAVES) void k_emp_adaptive(
    const Dev *dptr, const int *list, int nlist, const sedcol_params *__restrict__ pptr,
    const EmpCtl *__restrict__ cptr, int *next, int *pending, int *work, int *errflag)
{
  // Domain, parameters and span are read through pointers to device copies
  // rather than passed by value: about 100 dwords of by-value arguments would
  // be held in SGPRs for the whole kernel and spill.
  const Dev &d = *dptr;
  const sedcol_params &p = *pptr;
  const EmpCtl &ctl = *cptr;
  constexpr int G = BLK / LPC;
  __shared__ double sbc[G][2 * NV];
  __shared__ double sF[G][2][NV];
  __shared__ int sslot[G];
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const bool own = l < NV;  // lane l accumulates the top flux of variable l
  double(*F)[NV] = sF[g];
  const double *bc = sbc[g];
  const int r1 = ctl.r1;

  // state of the group's current column
  int c = 0, r = 0;
  Lane L;
  Geo Gm;
  long long cell = 0;
  double x[NV];
  // Per-lane budget sums (with their compensations) and the last top flux:
  // touched only when an attempt is accepted, so they live in LDS and leave
  // the registers to the right-hand side.
  __shared__ double sacc[7][BLK];
  // coefficients of the geometric-mean Patankar solve (one slot otherwise)
  __shared__ double sgeo[PW == 1 ? NV : 1][BLK];
  double &sinkC = sacc[0][threadIdx.x], &sinkN = sacc[1][threadIdx.x], &fint = sacc[2][threadIdx.x];
  double &sinkCc = sacc[3][threadIdx.x], &sinkNc = sacc[4][threadIdx.x], &fintc = sacc[5][threadIdx.x];
  double &flast = sacc[6][threadIdx.x];
  sedcol_coldiag dg;
  int64_t zs = 0;
  long long attempts = 0;
  int wn = 0;
  double dt = 0.0, rem = 0.0;
  bool have = false, act = false;

  for (;;) {
    if (!have) {
      if (l == 0) sslot[g] = atomicAdd(next, 1);
      __builtin_amdgcn_wave_barrier();
      const int s = sslot[g];
      __builtin_amdgcn_wave_barrier();
      if (s >= nlist) break;
      c = list[s];
      if (ctl.resume) {
        r = d.wrec[c];
        if (r > r1) continue;  // finished in an earlier window
        rem = d.wrem[c];
        attempts = d.watt[c];
      } else {
        r = ctl.r0 - 1;
        rem = 0.0;
        attempts = 0;
      }
      setup_lane(d, c, l, L, Gm, cell);
      act = L.active;
#pragma unroll
      for (int n = 0; n < NV; ++n) x[n] = act ? d.conc[n * d.ncell + cell] : 0.0;
      sinkC = act ? d.cellacc[cell] : 0.0;
      sinkN = act ? d.cellacc[d.ncell + cell] : 0.0;
      sinkCc = act ? d.cellacc[5 * d.ncell + cell] : 0.0;
      sinkNc = act ? d.cellacc[6 * d.ncell + cell] : 0.0;
      fint = own ? d.colacc[(long long)l * d.ncol + c] : 0.0;
      fintc = own ? d.colacc[(long long)(NV + l) * d.ncol + c] : 0.0;
      flast = own ? d.flux[(long long)l * d.ncol + c] : 0.0;
      dg = d.diag[c];
      dt = dg.dt_next;
      if (!(dt > 0.0)) dt = p.emp_dt_max;
      zs = 0;
      wn = 0;
      have = true;
      if (rem > 0.0) {
        // resumed inside record r
        load_forcing<ARITH>(d, p, c, r, Gm, L, sbc[g]);
        __builtin_amdgcn_wave_barrier();
      }
    }
    bool finish = false;
    if (!(rem > 0.0)) {
      // next forcing record with a positive span, or the column is done
      ++r;
      if (r > r1) {
        finish = true;
      } else {
        const double lo = (r == 0) ? -INFINITY : d.t0 + (double)r * d.dt_rec;
        const double hi = (r == d.nrec - 1) ? INFINITY : d.t0 + (double)(r + 1) * d.dt_rec;
        const double a = ctl.t_start > lo ? ctl.t_start : lo;
        const double b = ctl.t_end < hi ? ctl.t_end : hi;
        rem = b - a;
        load_forcing<ARITH>(d, p, c, r, Gm, L, sbc[g]);
        __builtin_amdgcn_wave_barrier();
        continue;
      }
    }
    bool pause = false;
    if (!finish) {
      if (attempts >= ctl.max_attempts) {
        dg.status = 2;
        if (l == 0) atomicOr(errflag, 2);
        finish = true;
      } else if (ctl.window > 0 && wn >= ctl.window) {
        pause = true;
      }
    }
    if (!finish && !pause) {
      ++attempts;
      ++wn;
      const bool clipped = !(dt < rem);
      const double h = clipped ? rem : dt;
      double f0[NV], c1[NV];
      double s0C, s0N, s1C, s1N;
      column_rhs<LPC, ARITH>(p, L, bc, x, s0C, s0N, [&](int n, double rr, double Fv) {
        f0[n] = rr;
        if (l == 0) F[0][n] = Fv;
      }, &d.zw);
      const double p1 = pfactor<LPC, false, PW, BLK>(x, f0, x, h, act, zs, dg.newton, &sgeo[0][threadIdx.x]);
#pragma unroll
      for (int n = 0; n < NV; ++n) {
        const double y = x[n] + (h * p1) * f0[n];
        c1[n] = y < 0.0 ? 0.0 : y;
      }
      column_rhs<LPC, ARITH>(p, L, bc, c1, s1C, s1N, [&](int n, double rr, double Fv) {
        f0[n] = (f0[n] + rr) * 0.5;
        if (l == 0) F[1][n] = Fv;
      }, &d.zw);
      const double p2 = pfactor<LPC, true, PW, BLK>(x, f0, c1, h, act, zs, dg.newton, &sgeo[0][threadIdx.x]);
      // err = max |c2 - c1| / sc: the lane's largest ratio by
      // cross-multiplication over the variables, one division, then the
      // maximum over the layers
      double bn = 0.0, bd = 1.0;
      bool bad = false;
#pragma unroll
      for (int n = 0; n < NV; ++n) {
        const double y = x[n] + (h * p2) * f0[n];
        const double yp = y < 0.0 ? 0.0 : y;
        const double sc = p.emp_atol + p.emp_rtol * (yp > x[n] ? yp : x[n]);
        const double dv = fabs(yp - c1[n]);
        bad |= (dv != dv) || (sc != sc);
        if (dv * bd > bn * sc) {
          bn = dv;
          bd = sc;
        }
      }
      double el = bn / bd;
      bad |= el != el;
      if (!act) {
        el = 0.0;
        bad = false;
      }
      const double err = lg_max<LPC>(el);
      if (group_any<LPC>(bad)) {
        dg.status = 1;
        if (l == 0) atomicOr(errflag, 1);
        finish = true;
      } else {
        const bool accept = (err <= 1.0) || !(h > p.emp_dt_min);
        double fac = err > 0.0 ? 0.9 * drsqrt(err) : 5.0;
        if (fac < 0.2) fac = 0.2;
        if (fac > 5.0) fac = 5.0;
        if (p1 < dg.p_min) dg.p_min = p1;
        if (p2 < dg.p_min) dg.p_min = p2;
        if (accept) {
#pragma unroll
          for (int n = 0; n < NV; ++n) {
            double y = x[n] + (h * p2) * f0[n];
            if (y < 0.0) {
              // rounding guard; rare, so accumulated straight into the
              // lane's own cell of the budget array
              if (elem(n) >= 0 && act)
                kadd(d.cellacc[(2 + elem(n)) * d.ncell + cell], d.cellacc[(7 + elem(n)) * d.ncell + cell], 0.0 - y);
              y = 0.0;
            }
            x[n] = y;
          }
          if (own) {
            kadd(fint, fintc, (h * p2) * ((F[0][l] + F[1][l]) * 0.5));
            flast = F[1][l];
          }
          kadd(sinkC, sinkCc, (h * p2) * ((s0C + s1C) * 0.5));
          kadd(sinkN, sinkNc, (h * p2) * ((s0N + s1N) * 0.5));
          rem = rem - h;
          dg.accepted += 1;
          if (err > 1.0) dg.forced += 1;
          dg.dt_last = h;
          if (h < dg.dt_minacc) dg.dt_minacc = h;
          const double prop = h * fac;
          dt = (clipped && dt > prop) ? dt : prop;
        } else {
          dg.rejected += 1;
          dt = h * fac;
        }
        if (dt < p.emp_dt_min) dt = p.emp_dt_min;
        if (dt > p.emp_dt_max) dt = p.emp_dt_max;
      }
    }
    if (finish || pause) {
#pragma unroll
      for (int n = 0; n < NV; ++n)
        if (act) d.conc[n * d.ncell + cell] = x[n];
      if (act) {
        d.cellacc[cell] = sinkC;
        d.cellacc[d.ncell + cell] = sinkN;
        d.cellacc[5 * d.ncell + cell] = sinkCc;
        d.cellacc[6 * d.ncell + cell] = sinkNc;
      }
      if (own) {
        d.colacc[(long long)l * d.ncol + c] = fint;
        d.colacc[(long long)(NV + l) * d.ncol + c] = fintc;
        d.flux[(long long)l * d.ncol + c] = flast;
      }
      const double zsum = lg_sum<LPC>((double)zs);
      if (l == 0) {
        dg.zero_sink += (int64_t)zsum;
        dg.dt_next = dt;
        d.diag[c] = dg;
        work[c] = attempts < 0x7fffffffll ? (int)attempts : 0x7fffffff;
        // finished (or stopped by an error): record past r1
        d.wrec[c] = pause ? r : r1 + 1;
        d.wrem[c] = rem;
        d.watt[c] = attempts;
        if (pause) atomicAdd(pending, 1);
      }
      have = false;
    }
  }
}

// ---------------------------------------------------------------------------
// adaptive EMP with two layers per lane
// ---------------------------------------------------------------------------
//
// k_emp_adaptive with lane l of a group holding layers 2l and 2l+1
// (column_rhs2, psolve2): a group of LPC lanes takes columns of up to 2 LPC
// layers, so a wave holds twice the columns at the same lane count per
// layer.  Product weights, reference and fast forms.  Every operation per
// layer and every reduction tree are those of k_emp_adaptive, so the
// results are bitwise the same.  LPC >= 16: lanes 0..NV-1 own the top-flux
// accumulators.

// one wave per SIMD: two layers of state and stage values per lane exceed
// the 256 registers of two waves
#ifndef SEDCOL_EMP2_MIN_WAVES
#define SEDCOL_EMP2_MIN_WAVES 1
#endif

__device__ __forceinline__ void setup_lane2(const Dev &d, int c, int l, Lane (&L)[2], Geo (&G)[2],
                                            long long (&cell)[2])
{
  setup_lane(d, c, 2 * l, L[0], G[0], cell[0]);
  setup_lane(d, c, 2 * l + 1, L[1], G[1], cell[1]);
}

template <int ARITH>
__device__ __forceinline__ void load_forcing2(const Dev &d, const sedcol_params &p, int c, int r, int l,
                                              const Geo (&G)[2], Lane (&L)[2], double *sbc)
{
  if (l < NV) {
    sbc[l] = d.cbw[((long long)r * NV + l) * d.ncol + c];
    sbc[NV + l] = d.fup[(long long)r * d.fup_rec_stride + (long long)l * d.ncol + c];
  }
  const double T = d.temp[(long long)r * d.ncol + c];
  diffusivities<ARITH>(p, G[0].B, T, G[0].phiI, G[0].bfac, L[0]);
  diffusivities<ARITH>(p, G[1].B, T, G[1].phiI, G[1].bfac, L[1]);
}

template <int LPC, int ARITH, int BLK>
__global__ __launch_bounds__(BLK, SEDCOL_EMP2_MIN_WAVES) void k_emp_adaptive2(
    const Dev *dptr, const int *list, int nlist, const sedcol_params *__restrict__ pptr,
    const EmpCtl *__restrict__ cptr, int *next, int *pending, int *work, int *errflag)
{
  static_assert(LPC >= 16, "lanes 0..NV-1 own the top-flux accumulators");
  const Dev &d = *dptr;
  const sedcol_params &p = *pptr;
  const EmpCtl &ctl = *cptr;
  constexpr int G = BLK / LPC;
  __shared__ double sbc[G][2 * NV];
  __shared__ double sF[G][2][NV];
  __shared__ int sslot[G];
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const bool own = l < NV;
  double(*F)[NV] = sF[g];
  const double *bc = sbc[g];
  const int r1 = ctl.r1;

  int c = 0, r = 0;
  Lane L[2];
  Geo Gm[2];
  long long cell[2] = {0, 0};
  bool act[2] = {false, false};
  double x[2][NV];
  // per-cell budget sums and compensations of both sub-layers, and the
  // lane's top-flux accumulators, in LDS as in k_emp_adaptive
  __shared__ double scell[4][2][BLK];
  __shared__ double scol[3][BLK];
  double &fint = scol[0][threadIdx.x], &fintc = scol[1][threadIdx.x], &flast = scol[2][threadIdx.x];
  sedcol_coldiag dg;
  int64_t zs = 0;
  long long attempts = 0;
  int wn = 0;
  double dt = 0.0, rem = 0.0;
  bool have = false;

  for (;;) {
    if (!have) {
      if (l == 0) sslot[g] = atomicAdd(next, 1);
      __builtin_amdgcn_wave_barrier();
      const int sl = sslot[g];
      __builtin_amdgcn_wave_barrier();
      if (sl >= nlist) break;
      c = list[sl];
      if (ctl.resume) {
        r = d.wrec[c];
        if (r > r1) continue;
        rem = d.wrem[c];
        attempts = d.watt[c];
      } else {
        r = ctl.r0 - 1;
        rem = 0.0;
        attempts = 0;
      }
      setup_lane2(d, c, l, L, Gm, cell);
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        act[s] = L[s].active;
#pragma unroll
        for (int n = 0; n < NV; ++n) x[s][n] = act[s] ? d.conc[n * d.ncell + cell[s]] : 0.0;
        scell[0][s][threadIdx.x] = act[s] ? d.cellacc[cell[s]] : 0.0;
        scell[1][s][threadIdx.x] = act[s] ? d.cellacc[d.ncell + cell[s]] : 0.0;
        scell[2][s][threadIdx.x] = act[s] ? d.cellacc[5 * d.ncell + cell[s]] : 0.0;
        scell[3][s][threadIdx.x] = act[s] ? d.cellacc[6 * d.ncell + cell[s]] : 0.0;
      }
      fint = own ? d.colacc[(long long)l * d.ncol + c] : 0.0;
      fintc = own ? d.colacc[(long long)(NV + l) * d.ncol + c] : 0.0;
      flast = own ? d.flux[(long long)l * d.ncol + c] : 0.0;
      dg = d.diag[c];
      dt = dg.dt_next;
      if (!(dt > 0.0)) dt = p.emp_dt_max;
      zs = 0;
      wn = 0;
      have = true;
      if (rem > 0.0) {
        load_forcing2<ARITH>(d, p, c, r, l, Gm, L, sbc[g]);
        __builtin_amdgcn_wave_barrier();
      }
    }
    bool finish = false;
    if (!(rem > 0.0)) {
      ++r;
      if (r > r1) {
        finish = true;
      } else {
        const double lo = (r == 0) ? -INFINITY : d.t0 + (double)r * d.dt_rec;
        const double hi = (r == d.nrec - 1) ? INFINITY : d.t0 + (double)(r + 1) * d.dt_rec;
        const double a = ctl.t_start > lo ? ctl.t_start : lo;
        const double b = ctl.t_end < hi ? ctl.t_end : hi;
        rem = b - a;
        load_forcing2<ARITH>(d, p, c, r, l, Gm, L, sbc[g]);
        __builtin_amdgcn_wave_barrier();
        continue;
      }
    }
    bool pause = false;
    if (!finish) {
      if (attempts >= ctl.max_attempts) {
        dg.status = 2;
        if (l == 0) atomicOr(errflag, 2);
        finish = true;
      } else if (ctl.window > 0 && wn >= ctl.window) {
        pause = true;
      }
    }
    if (!finish && !pause) {
      ++attempts;
      ++wn;
      const bool clipped = !(dt < rem);
      const double h = clipped ? rem : dt;
      double f0[2][NV], c1[2][NV];
      double s0C[2], s0N[2], s1C[2], s1N[2];
      column_rhs2<ARITH>(p, L, bc, x, s0C, s0N, [&](int s, int n, double rr, double Fv) {
        f0[s][n] = rr;
        if (l == 0 && s == 0) F[0][n] = Fv;
      });
      const double p1 = psolve2<LPC, false>(x, f0, x, h, act, zs, dg.newton);
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int n = 0; n < NV; ++n) {
          const double y = x[s][n] + (h * p1) * f0[s][n];
          c1[s][n] = y < 0.0 ? 0.0 : y;
        }
      column_rhs2<ARITH>(p, L, bc, c1, s1C, s1N, [&](int s, int n, double rr, double Fv) {
        f0[s][n] = (f0[s][n] + rr) * 0.5;
        if (l == 0 && s == 0) F[1][n] = Fv;
      });
      const double p2 = psolve2<LPC, true>(x, f0, c1, h, act, zs, dg.newton);
      double el[2];
      bool bad = false;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        double bn = 0.0, bd = 1.0;
#pragma unroll
        for (int n = 0; n < NV; ++n) {
          const double y = x[s][n] + (h * p2) * f0[s][n];
          const double yp = y < 0.0 ? 0.0 : y;
          const double sc = p.emp_atol + p.emp_rtol * (yp > x[s][n] ? yp : x[s][n]);
          const double dv = fabs(yp - c1[s][n]);
          bad |= act[s] && ((dv != dv) || (sc != sc));
          if (dv * bd > bn * sc) {
            bn = dv;
            bd = sc;
          }
        }
        el[s] = bn / bd;
        bad |= act[s] && (el[s] != el[s]);
        if (!act[s]) el[s] = 0.0;
      }
      // the first tree level (lower layer first), then the lane butterfly
      const double elp = (el[1] > el[0] || el[1] != el[1]) ? el[1] : el[0];
      const double err = lg_max<LPC>(elp);
      if (group_any<LPC>(bad)) {
        dg.status = 1;
        if (l == 0) atomicOr(errflag, 1);
        finish = true;
      } else {
        const bool accept = (err <= 1.0) || !(h > p.emp_dt_min);
        double fac = err > 0.0 ? 0.9 * drsqrt(err) : 5.0;
        if (fac < 0.2) fac = 0.2;
        if (fac > 5.0) fac = 5.0;
        if (p1 < dg.p_min) dg.p_min = p1;
        if (p2 < dg.p_min) dg.p_min = p2;
        if (accept) {
#pragma unroll
          for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int n = 0; n < NV; ++n) {
              double y = x[s][n] + (h * p2) * f0[s][n];
              if (y < 0.0) {
                if (elem(n) >= 0 && act[s])
                  kadd(d.cellacc[(2 + elem(n)) * d.ncell + cell[s]], d.cellacc[(7 + elem(n)) * d.ncell + cell[s]],
                       0.0 - y);
                y = 0.0;
              }
              x[s][n] = y;
            }
          if (own) {
            kadd(fint, fintc, (h * p2) * ((F[0][l] + F[1][l]) * 0.5));
            flast = F[1][l];
          }
#pragma unroll
          for (int s = 0; s < 2; ++s) {
            kadd(scell[0][s][threadIdx.x], scell[2][s][threadIdx.x], (h * p2) * ((s0C[s] + s1C[s]) * 0.5));
            kadd(scell[1][s][threadIdx.x], scell[3][s][threadIdx.x], (h * p2) * ((s0N[s] + s1N[s]) * 0.5));
          }
          rem = rem - h;
          dg.accepted += 1;
          if (err > 1.0) dg.forced += 1;
          dg.dt_last = h;
          if (h < dg.dt_minacc) dg.dt_minacc = h;
          const double prop = h * fac;
          dt = (clipped && dt > prop) ? dt : prop;
        } else {
          dg.rejected += 1;
          dt = h * fac;
        }
        if (dt < p.emp_dt_min) dt = p.emp_dt_min;
        if (dt > p.emp_dt_max) dt = p.emp_dt_max;
      }
    }
    if (finish || pause) {
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        if (act[s]) {
#pragma unroll
          for (int n = 0; n < NV; ++n) d.conc[n * d.ncell + cell[s]] = x[s][n];
          d.cellacc[cell[s]] = scell[0][s][threadIdx.x];
          d.cellacc[d.ncell + cell[s]] = scell[1][s][threadIdx.x];
          d.cellacc[5 * d.ncell + cell[s]] = scell[2][s][threadIdx.x];
          d.cellacc[6 * d.ncell + cell[s]] = scell[3][s][threadIdx.x];
        }
      }
      if (own) {
        d.colacc[(long long)l * d.ncol + c] = fint;
        d.colacc[(long long)(NV + l) * d.ncol + c] = fintc;
        d.flux[(long long)l * d.ncol + c] = flast;
      }
      const double zsum = lg_sum<LPC>((double)zs);
      if (l == 0) {
        dg.zero_sink += (int64_t)zsum;
        dg.dt_next = dt;
        d.diag[c] = dg;
        work[c] = attempts < 0x7fffffffll ? (int)attempts : 0x7fffffff;
        d.wrec[c] = pause ? r : r1 + 1;
        d.wrem[c] = rem;
        d.watt[c] = attempts;
        if (pause) atomicAdd(pending, 1);
      }
      have = false;
    }
  }
}

// ---------------------------------------------------------------------------
// linearly implicit adaptive stepper (method 7): ROS2 as a W-method
// ---------------------------------------------------------------------------
//
//   W k1 = f(c),  W k2 = f(c + h k1) - 2 k1,  c+ = c + h (3/2 k1 + 1/2 k2),
//   error estimate h (k1 + k2)/2 against the first-order c + h k1,
// gamma = 1 + 1/sqrt(2) (Verwer, Spee, Blom and Hundsdorfer 1999).  W is
// the product of the per-species tridiagonal transport factor
// B = I - gamma h J_T (Thomas sweeps, one lane per species, through LDS)
// and the per-cell reaction factor A = I - gamma h J_R (block solve in the
// lane's registers), plus one correction with the full matrix:
//   k0 = A^-1 B^-1 r,  k = k0 + A^-1 B^-1 (r - (I - gamma h (J_T + J_R)) k0).
// The transport coefficients come from the same diffusivities as the
// right-hand side, whose arithmetic form (reference or fast) is the
// instance's.  Budgets: top fluxes integrate as h/2 [(F0 + F1) + g (T1 +
// T2)] with T the linearised Dirichlet top-flux terms of the two solves,
// the DIC and N2 sinks as h/2 [(s0 + s1) + g grad s . (k1 + k2)]: exactly
// what the update moves.  Positivity: a new state below -1e-3 atol
// rejects the step (zero_sink counts these); smaller negative values are
// clamped to 0 and booked as clamp mass.  Failed local solves (a pivot
// that is not positive) reject the step with a quarter step (newton counts
// these).  Same operation sequence as the oracle's col_ros2_adaptive.

constexpr double ROS_GAMMA = 1.7071067811865475;

// per-group LDS of the stepper
template <int LPC>
struct RosShared {
  double G[2][LPC];     // transport coefficients D/dzc (D/dz at the top): [0] particulate, [1] dissolved
  double W[LPC];        // 1/(phi dz)
  double ph[LPC];       // porosity
  // Species rows padded by one double: the sweep of species n runs in lane
  // n, so the lanes read one layer of 11 different rows at once; with the
  // padding those sit in different LDS banks (unpadded, the row pitch is a
  // multiple of the bank count and the 11 reads serialise)
  double T[NV][LPC + 1];    // right-hand sides and solutions of the species sweeps
  double inv[NV][LPC + 1];  // reciprocal Thomas pivots
  double A[NV][LPC + 1];    // sub-diagonal couplings of the forward sweep
  double E[NV][LPC + 1];    // super-diagonal coupling times reciprocal pivot (back substitution)
  double top[2][NV];    // linearised Dirichlet top-flux terms of the two stages
};

// Thomas factor of species n = l (lanes l < NV of the group): reciprocal
// pivots, the forward-sweep couplings a_k and the back-substitution
// couplings c_k / piv_k, so that a sweep is one dependent fma-and-multiply
// (forward) or multiply-and-add (backward) per layer with operands that do
// not depend on the chain
template <int LPC>
__device__ __forceinline__ void ros_factor(RosShared<LPC> &R, int l, int K, double g, unsigned dirmask)
{
  if (l < NV) {
    const bool pt = part(l);
    const double *__restrict__ G = R.G[pt ? 0 : 1];
    const double *__restrict__ W = R.W;
    const double *__restrict__ ph = R.ph;
    double *__restrict__ inv = R.inv[l];
    double *__restrict__ A = R.A[l];
    double *__restrict__ E = R.E[l];
    const bool dir = (dirmask >> l) & 1u;
    double e = 0.0;
    for (int k = 0; k < K; ++k) {
      const double qk = pt ? ph[k] : 1.0;
      const double s = ((k > 0 || dir) ? G[k] : 0.0) + (k < K - 1 ? G[k + 1] : 0.0);
      const double b = 1.0 + g * (W[k] * (s * qk));
      const double a = k > 0 ? g * ((W[k] * G[k]) * (pt ? ph[k - 1] : 1.0)) : 0.0;
      const double piv = k > 0 ? b - a * e : b;
      const double iv = 1.0 / piv;
      const double c = k < K - 1 ? g * ((W[k] * G[k + 1]) * (pt ? ph[k + 1] : 1.0)) : 0.0;
      e = c * iv;
      inv[k] = iv;
      A[k] = a;
      E[k] = e;
    }
  }
}

// B z = T in place, species n = l
template <int LPC>
__device__ __forceinline__ void ros_sweep(RosShared<LPC> &R, int l, int K)
{
  if (l < NV) {
    double *__restrict__ T = R.T[l];
    const double *__restrict__ iv = R.inv[l];
    const double *__restrict__ A = R.A[l];
    const double *__restrict__ E = R.E[l];
    double d = T[0] * iv[0];
    T[0] = d;
    for (int k = 1; k < K; ++k) {
      d = (T[k] + A[k] * d) * iv[k];
      T[k] = d;
    }
    for (int k = K - 2; k >= 0; --k) {
      d = T[k] + E[k] * d;
      T[k] = d;
    }
  }
}

// B^-1 r for the lane's layer: stage through LDS, sweep, read back
template <int LPC>
__device__ __forceinline__ void ros_bsolve(RosShared<LPC> &R, int l, int K, bool act, double g, const double (&r)[NV],
                                           double (&v)[NV])
{
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int n = 0; n < NV; ++n)
    if (act) R.T[n][l] = r[n];
  __builtin_amdgcn_wave_barrier();
  ros_sweep<LPC>(R, l, K);
  __builtin_amdgcn_wave_barrier();
#pragma unroll
  for (int n = 0; n < NV; ++n) v[n] = act ? R.T[n][l] : 0.0;
  __builtin_amdgcn_wave_barrier();
}

// k = W^-1 r (AMF and one correction); r is overwritten.  Lane 0 stores
// the stage's linearised Dirichlet top-flux terms in top[n].  Returns the
// lane's local solve status.
template <int LPC>
__device__ __forceinline__ bool ros_wsolve(RosShared<LPC> &R, const KJac &J, int l, int K, bool act, double g,
                                           unsigned dirmask, double phi, double (&r)[NV], double (&k)[NV],
                                           double *top)
{
  // the lane's transport coefficients, read from the group's LDS
  const double W = R.W[l];
  const double Gn_p = l + 1 < K ? R.G[0][l + 1] : 0.0, Gn_d = l + 1 < K ? R.G[1][l + 1] : 0.0;
  const double Gtop_p = R.G[0][0], Gtop_d = R.G[1][0];
  double a[NV];
  ros_bsolve<LPC>(R, l, K, act, g, r, a);
  bool ok = kjac_solve(J, g, a, k);  // k = k0
  kjac_apply(J, k, a);               // a = J_R k0
  // residual (r - k0) + g (J_T k0 + J_R k0), into r
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const bool pt = part(n);
    const double q = pt ? phi : 1.0;
    const double zp = k[n] * q;
    const double zpn = lg_dn(zp);
    const double Gn = pt ? Gn_p : Gn_d;
    const double Fd = l < K - 1 ? -(Gn * (zpn - zp)) : 0.0;
    double Fu = lg_up(Fd);
    if (l == 0) Fu = ((dirmask >> n) & 1u) ? -((pt ? Gtop_p : Gtop_d) * zp) : 0.0;
    const double t = (Fu - Fd) * W;
    r[n] = (r[n] - k[n]) + g * (t + a[n]);
  }
  ros_bsolve<LPC>(R, l, K, act, g, r, a);  // a = v2
  if (l == 0) {
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      const bool pt = part(n);
      top[n] = ((dirmask >> n) & 1u) ? 0.0 - (((pt ? Gtop_p : Gtop_d) * (pt ? phi : 1.0)) * (k[n] + a[n])) : 0.0;
    }
  }
  ok = kjac_solve(J, g, a, r) && ok;  // r = kc
#pragma unroll
  for (int n = 0; n < NV; ++n) k[n] = k[n] + r[n];
  return ok || !act;
}

// one wave per SIMD: the stepper's live state (state, two stage vectors,
// the reaction Jacobian) exceeds the 256 registers of two waves; with one
// wave the overflow goes to the accumulation registers instead of scratch
#ifndef SEDCOL_ROS_MIN_WAVES
#define SEDCOL_ROS_MIN_WAVES 1
#endif

template <int LPC, int ARITH, int BLK>
__global__ __launch_bounds__(BLK, SEDCOL_ROS_MIN_WAVES) void k_ros_adaptive(
    const Dev *dptr, const int *list, int nlist, const sedcol_params *__restrict__ pptr,
    const EmpCtl *__restrict__ cptr, int *next, int *pending, int *work, int *errflag)
{
  const Dev &d = *dptr;
  const sedcol_params &p = *pptr;
  const EmpCtl &ctl = *cptr;
  constexpr int G = BLK / LPC;
  __shared__ double sbc[G][2 * NV];
  __shared__ double sF[G][2][NV];
  __shared__ int sslot[G];
  __shared__ RosShared<LPC> sR[G];
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const bool own = l < NV;
  double(*F)[NV] = sF[g];
  const double *bc = sbc[g];
  RosShared<LPC> &R = sR[g];
  const int r1 = ctl.r1;
  const unsigned dirmask = (p.bcup_part == 2 ? 0x41Fu : 0u) | (p.bcup_dis == 2 ? 0x3E0u : 0u);
  const double negtol = 1.0e-3 * p.emp_atol;

  int c = 0, r = 0;
  Lane L;
  Geo Gm;
  long long cell = 0;
  double x[NV];
  __shared__ double sacc[7][BLK];
  double &sinkC = sacc[0][threadIdx.x], &sinkN = sacc[1][threadIdx.x], &fint = sacc[2][threadIdx.x];
  double &sinkCc = sacc[3][threadIdx.x], &sinkNc = sacc[4][threadIdx.x], &fintc = sacc[5][threadIdx.x];
  double &flast = sacc[6][threadIdx.x];
  sedcol_coldiag dg;
  long long attempts = 0;
  int wn = 0;
  double dt = 0.0, rem = 0.0;
  bool have = false, act = false;

  auto coefficients = [&]() {
    // W-matrix coefficients of the record (the right-hand side's
    // diffusivities) into the group's LDS
    double Gp, Gd, Wl;
    if (ARITH == ARITH_FAST) {
      Gp = L.Gb;
      Gd = L.Gd;
      Wl = L.W;
    } else {
      const double len = l == 0 ? L.dz : L.dzc;
      Gp = L.Dbio / len;
      Gd = L.Ddis / len;
      Wl = 1.0 / (L.phi * L.dz);
    }
    __builtin_amdgcn_wave_barrier();
    if (act) {
      R.G[0][l] = Gp;
      R.G[1][l] = Gd;
      R.W[l] = Wl;
      R.ph[l] = L.phi;
    }
    __builtin_amdgcn_wave_barrier();
  };

  for (;;) {
    if (!have) {
      if (l == 0) sslot[g] = atomicAdd(next, 1);
      __builtin_amdgcn_wave_barrier();
      const int s = sslot[g];
      __builtin_amdgcn_wave_barrier();
      if (s >= nlist) break;
      c = list[s];
      if (ctl.resume) {
        r = d.wrec[c];
        if (r > r1) continue;
        rem = d.wrem[c];
        attempts = d.watt[c];
      } else {
        r = ctl.r0 - 1;
        rem = 0.0;
        attempts = 0;
      }
      setup_lane(d, c, l, L, Gm, cell);
      act = L.active;
#pragma unroll
      for (int n = 0; n < NV; ++n) x[n] = act ? d.conc[n * d.ncell + cell] : 0.0;
      sinkC = act ? d.cellacc[cell] : 0.0;
      sinkN = act ? d.cellacc[d.ncell + cell] : 0.0;
      sinkCc = act ? d.cellacc[5 * d.ncell + cell] : 0.0;
      sinkNc = act ? d.cellacc[6 * d.ncell + cell] : 0.0;
      fint = own ? d.colacc[(long long)l * d.ncol + c] : 0.0;
      fintc = own ? d.colacc[(long long)(NV + l) * d.ncol + c] : 0.0;
      flast = own ? d.flux[(long long)l * d.ncol + c] : 0.0;
      dg = d.diag[c];
      dt = dg.dt_next;
      if (!(dt > 0.0)) dt = p.emp_dt_max;
      wn = 0;
      have = true;
      if (rem > 0.0) {
        load_forcing<ARITH>(d, p, c, r, Gm, L, sbc[g]);
        coefficients();
      }
    }
    bool finish = false;
    if (!(rem > 0.0)) {
      ++r;
      if (r > r1) {
        finish = true;
      } else {
        const double lo = (r == 0) ? -INFINITY : d.t0 + (double)r * d.dt_rec;
        const double hi = (r == d.nrec - 1) ? INFINITY : d.t0 + (double)(r + 1) * d.dt_rec;
        const double a = ctl.t_start > lo ? ctl.t_start : lo;
        const double b = ctl.t_end < hi ? ctl.t_end : hi;
        rem = b - a;
        load_forcing<ARITH>(d, p, c, r, Gm, L, sbc[g]);
        coefficients();
        continue;
      }
    }
    bool pause = false;
    if (!finish) {
      if (attempts >= ctl.max_attempts) {
        dg.status = 2;
        if (l == 0) atomicOr(errflag, 2);
        finish = true;
      } else if (ctl.window > 0 && wn >= ctl.window) {
        pause = true;
      }
    }
    if (!finish && !pause) {
      ++attempts;
      ++wn;
      const bool clipped = !(dt < rem);
      const double h = clipped ? rem : dt;
      const double gh = ROS_GAMMA * h;
      const int K = L.K;
      double f[NV], k1[NV], k2[NV];
      double s0C, s0N, s1C = 0.0, s1N = 0.0;
      column_rhs<LPC, ARITH>(p, L, bc, x, s0C, s0N, [&](int n, double rr, double Fv) {
        f[n] = rr;
        if (l == 0) F[0][n] = Fv;
      });
      __builtin_amdgcn_wave_barrier();
      ros_factor<LPC>(R, l, K, gh, dirmask);
      __builtin_amdgcn_wave_barrier();
      // the Jacobian is evaluated again for the second stage (same values)
      // rather than kept live across the second right-hand side
      double gC = 0.0, gN = 0.0;
      bool ok;
      {
        KJac J;
        kin_jac(p, L.fT, x, J);
        ok = ros_wsolve<LPC>(R, J, l, K, act, gh, dirmask, L.phi, f, k1, R.top[0]);
        gC = kjac_sinkC(J, k1);
        gN = kjac_sinkN(J, k1);
      }
      bool fail = group_any<LPC>(!ok);
      if (!fail) {
#pragma unroll
        for (int n = 0; n < NV; ++n) {
          const double y = x[n] + h * k1[n];
          f[n] = y < 0.0 ? 0.0 : y;
        }
        double c1[NV];
#pragma unroll
        for (int n = 0; n < NV; ++n) c1[n] = f[n];
        column_rhs<LPC, ARITH>(p, L, bc, c1, s1C, s1N, [&](int n, double rr, double Fv) {
          f[n] = rr - 2.0 * k1[n];
          if (l == 0) F[1][n] = Fv;
        });
        KJac J;
        kin_jac(p, L.fT, x, J);
        ok = ros_wsolve<LPC>(R, J, l, K, act, gh, dirmask, L.phi, f, k2, R.top[1]);
        fail = group_any<LPC>(!ok);
        gC = gC + kjac_sinkC(J, k2);
        gN = gN + kjac_sinkN(J, k2);
      }
      double bn = 0.0, bd = 1.0;
      bool bad = false, neg = false;
      if (!fail) {
#pragma unroll
        for (int n = 0; n < NV; ++n) {
          const double y = x[n] + h * (1.5 * k1[n] + 0.5 * k2[n]);
          const double yp = y < 0.0 ? 0.0 : y;
          const double sc = p.emp_atol + p.emp_rtol * (yp > x[n] ? yp : x[n]);
          const double dv = fabs(h * (0.5 * k1[n] + 0.5 * k2[n]));
          bad |= (dv != dv) || (sc != sc);
          neg |= y < 0.0 - negtol;
          if (dv * bd > bn * sc) {
            bn = dv;
            bd = sc;
          }
        }
      }
      double el = bn / bd;
      bad |= el != el;
      if (!act) {
        el = 0.0;
        bad = false;
        neg = false;
      }
      const double err = lg_max<LPC>(el);
      neg = group_any<LPC>(neg);
      if (group_any<LPC>(bad) || (fail && !(h > p.emp_dt_min))) {
        dg.status = 1;
        if (l == 0) atomicOr(errflag, 1);
        finish = true;
      } else {
        const bool accept = !fail && ((err <= 1.0 && !neg) || !(h > p.emp_dt_min));
        double fac = err > 0.0 ? 0.9 * drsqrt(err) : 5.0;
        if (fac < 0.2) fac = 0.2;
        if (fac > 5.0) fac = 5.0;
        if (accept) {
#pragma unroll
          for (int n = 0; n < NV; ++n) {
            double y = x[n] + h * (1.5 * k1[n] + 0.5 * k2[n]);
            if (y < 0.0) {
              if (elem(n) >= 0 && act)
                kadd(d.cellacc[(2 + elem(n)) * d.ncell + cell], d.cellacc[(7 + elem(n)) * d.ncell + cell], 0.0 - y);
              y = 0.0;
            }
            x[n] = y;
          }
          __builtin_amdgcn_wave_barrier();
          if (own) {
            kadd(fint, fintc, (h * 0.5) * ((F[0][l] + F[1][l]) + gh * (R.top[0][l] + R.top[1][l])));
            flast = F[1][l];
          }
          __builtin_amdgcn_wave_barrier();
          kadd(sinkC, sinkCc, (h * 0.5) * ((s0C + s1C) + gh * gC));
          kadd(sinkN, sinkNc, (h * 0.5) * ((s0N + s1N) + gh * gN));
          rem = rem - h;
          dg.accepted += 1;
          if (err > 1.0 || neg) dg.forced += 1;
          dg.dt_last = h;
          if (h < dg.dt_minacc) dg.dt_minacc = h;
          const double prop = h * fac;
          dt = (clipped && dt > prop) ? dt : prop;
        } else {
          dg.rejected += 1;
          if (fail) {
            dg.newton += 1;
            dt = h * 0.25;
          } else if (err <= 1.0) {
            dg.zero_sink += 1;
            dt = h * 0.5;
          } else {
            dt = h * fac;
          }
        }
        if (dt < p.emp_dt_min) dt = p.emp_dt_min;
        if (dt > p.emp_dt_max) dt = p.emp_dt_max;
      }
    }
    if (finish || pause) {
#pragma unroll
      for (int n = 0; n < NV; ++n)
        if (act) d.conc[n * d.ncell + cell] = x[n];
      if (act) {
        d.cellacc[cell] = sinkC;
        d.cellacc[d.ncell + cell] = sinkN;
        d.cellacc[5 * d.ncell + cell] = sinkCc;
        d.cellacc[6 * d.ncell + cell] = sinkNc;
      }
      if (own) {
        d.colacc[(long long)l * d.ncol + c] = fint;
        d.colacc[(long long)(NV + l) * d.ncol + c] = fintc;
        d.flux[(long long)l * d.ncol + c] = flast;
      }
      if (l == 0) {
        dg.dt_next = dt;
        d.diag[c] = dg;
        work[c] = attempts < 0x7fffffffll ? (int)attempts : 0x7fffffff;
        d.wrec[c] = pause ? r : r1 + 1;
        d.wrem[c] = rem;
        d.watt[c] = attempts;
        if (pause) atomicAdd(pending, 1);
      }
      have = false;
    }
  }
}

// ---------------------------------------------------------------------------
// ADAPTIVE_EULER (solver_library.F90:104-140): the reject rule is global, so
// each sub-step is a right-hand side, a domain-wide check and a commit
// ---------------------------------------------------------------------------

template <int LPC, int ARITH>
__global__ __launch_bounds__(BLOCK) void k_rhs_store(Dev d, const int *list, int nlist, sedcol_params p, int rec,
                                                     double *r_out, double *s_out)
{
  constexpr int G = BLOCK / LPC;
  __shared__ double sbc[G][2 * NV];
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int slot = blockIdx.x * G + g;
  const bool valid = slot < nlist;
  const int c = valid ? list[slot] : 0;
  Lane L;
  Geo Gm;
  long long cell = 0;
  if (valid) {
    setup_lane(d, c, l, L, Gm, cell);
    load_forcing<ARITH>(d, p, c, rec, Gm, L, sbc[g]);
  }
  __syncthreads();
  if (!valid) return;
  double x[NV];
#pragma unroll
  for (int n = 0; n < NV; ++n) x[n] = L.active ? d.conc[n * d.ncell + cell] : 0.0;
  double sC, sN;
  column_rhs<LPC, ARITH>(p, L, sbc[g], x, sC, sN, [&](int n, double r, double F) {
    if (L.active) r_out[n * d.ncell + cell] = r;
    if (l == 0) d.flux[(long long)n * d.ncol + c] = F;
  }, &d.zw);
  if (L.active) {
    s_out[cell] = sC;
    s_out[d.ncell + cell] = sN;
  }
}

template <int LPC>
__global__ __launch_bounds__(BLOCK) void k_ae_check(Dev d, const int *list, int nlist, double dt_red, double rcm,
                                                    const double *r_in, int *viol)
{
  constexpr int G = BLOCK / LPC;
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int slot = blockIdx.x * G + g;
  if (slot >= nlist) return;
  const int c = list[slot];
  const int o = d.off[c], K = d.off[c + 1] - o;
  if (l >= K) return;
  const long long cell = (long long)o + l;
  bool v = false;
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const double x = d.conc[n * d.ncell + cell];
    const double c1 = x + dt_red * r_in[n * d.ncell + cell];
    v |= (c1 - (1.0 + rcm) * x) < 0.0;
  }
  if (v) atomicOr(viol, 1);
}

template <int LPC>
__global__ __launch_bounds__(BLOCK) void k_ae_commit(Dev d, const int *list, int nlist, double dt_red,
                                                     const double *r_in, const double *s_in)
{
  constexpr int G = BLOCK / LPC;
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int slot = blockIdx.x * G + g;
  if (slot >= nlist) return;
  const int c = list[slot];
  const int o = d.off[c], K = d.off[c + 1] - o;
  if (l < NV)
    kadd(d.colacc[(long long)l * d.ncol + c], d.colacc[(long long)(NV + l) * d.ncol + c],
         dt_red * d.flux[(long long)l * d.ncol + c]);
  if (l == 0) {
    sedcol_coldiag dg = d.diag[c];
    dg.accepted += 1;
    dg.dt_last = dt_red;
    if (dt_red < dg.dt_minacc) dg.dt_minacc = dt_red;
    d.diag[c] = dg;
  }
  if (l >= K) return;
  const long long cell = (long long)o + l;
#pragma unroll
  for (int n = 0; n < NV; ++n) d.conc[n * d.ncell + cell] = d.conc[n * d.ncell + cell] + dt_red * r_in[n * d.ncell + cell];
  kadd(d.cellacc[cell], d.cellacc[5 * d.ncell + cell], dt_red * s_in[cell]);
  kadd(d.cellacc[d.ncell + cell], d.cellacc[6 * d.ncell + cell], dt_red * s_in[d.ncell + cell]);
}

// ADAPTIVE_EULER steps of a small domain in one workgroup: every wet column
// of the list is a lane group of the block, so the reference's global rule
// (the step is quartered while any cell of the domain fails the
// relative-change test, solver_library.F90:100-140) is a block-wide vote
// instead of a flag read by the host.  Per attempt the same operations as
// k_rhs_store, k_ae_check and k_ae_commit, and k_clamp after each step, so
// the results are those of the multi-launch path bit for bit.  The state
// stays in registers; the top fluxes of the last right-hand side go to LDS
// for the lanes that accumulate them.  *min_dt receives the smallest step
// taken.  A launch stops after kAeFusedMaxAttempts attempts (errflag bit 2).
template <int LPC, int ARITH>
__global__ __launch_bounds__(kAeFusedBlock) void k_ae_fused(const Dev *dptr, const int *list, int nlist,
                                                           const sedcol_params *__restrict__ pptr, int rec, double dt,
                                                           int nsteps, int *errflag, double *min_dt)
{
  // domain and parameters through pointers to device copies: passed by
  // value they would sit in scalar registers for the whole run and spill
  const Dev &d = *dptr;
  const sedcol_params &p = *pptr;
  constexpr int G = kAeFusedBlock / LPC;
  __shared__ double sbc[G][2 * NV];
  __shared__ double sF[G][NV];
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const bool valid = g < nlist;
  const int c = valid ? list[g] : 0;
  Lane L;
  Geo Gm;
  long long cell = 0;
  if (valid) {
    setup_lane(d, c, l, L, Gm, cell);
    load_forcing<ARITH>(d, p, c, rec, Gm, L, sbc[g]);
  } else {
    L.l = l;
    L.K = 0;
    L.active = false;
  }
  __syncthreads();
  const bool act = valid && L.active;
  const bool own = valid && l < NV;
  double x[NV], r[NV];
#pragma unroll
  for (int n = 0; n < NV; ++n) x[n] = act ? d.conc[n * d.ncell + cell] : 0.0;
  double sC = 0.0, sN = 0.0;
  // sums touched only by accepted attempts live in LDS, leaving the
  // registers to the right-hand side: the DIC and N2 sinks with their
  // compensations, the lane's top-flux sum and compensation, and the
  // column's diagnostics (updated by its top lane)
  __shared__ double sacc[6][kAeFusedBlock];
  __shared__ sedcol_coldiag sdg[G];
  double(&acc)[6][kAeFusedBlock] = sacc;
  const int t = threadIdx.x;
  sedcol_coldiag &dg = sdg[g];
  acc[0][t] = act ? d.cellacc[cell] : 0.0;
  acc[1][t] = act ? d.cellacc[d.ncell + cell] : 0.0;
  acc[2][t] = act ? d.cellacc[5 * d.ncell + cell] : 0.0;
  acc[3][t] = act ? d.cellacc[6 * d.ncell + cell] : 0.0;
  acc[4][t] = own ? d.colacc[(long long)l * d.ncol + c] : 0.0;
  acc[5][t] = own ? d.colacc[(long long)(NV + l) * d.ncol + c] : 0.0;
  if (valid && l == 0) dg = d.diag[c];
  const double rcm = p.relative_change_min;
  double mind = INFINITY;
  int attempts = 0;
  bool stop = false;
  for (int s = 0; s < nsteps && !stop; ++s) {
    double dt_int = 0.0, dt_red = dt;
    bool need = true;
    while (dt_int < dt) {
      if (++attempts > kAeFusedMaxAttempts) {
        if (threadIdx.x == 0) atomicOr(errflag, 2);
        stop = true;
        break;
      }
      if (need) {
        if (valid) {
          column_rhs<LPC, ARITH>(p, L, sbc[g], x, sC, sN, [&](int n, double rr, double F) {
            r[n] = rr;
            if (l == 0) {
              sF[g][n] = F;
              d.flux[(long long)n * d.ncol + c] = F;
            }
          }, &d.zw);
        }
        need = false;
        __syncthreads();
      }
      bool v = false;
      if (act) {
#pragma unroll
        for (int n = 0; n < NV; ++n) {
          const double c1 = x[n] + dt_red * r[n];
          v |= (c1 - (1.0 + rcm) * x[n]) < 0.0;
        }
      }
      const int viol = __syncthreads_or(v);
      if (viol && dt_red > p.dt_min) {
        dt_red = dt_red * 0.25;  // the state is unchanged: the right-hand side is reused
      } else {
        if (dt_red < mind) mind = dt_red;
        if (own) kadd(acc[4][t], acc[5][t], dt_red * sF[g][l]);
        if (valid && l == 0) {
          dg.accepted += 1;
          dg.dt_last = dt_red;
          if (dt_red < dg.dt_minacc) dg.dt_minacc = dt_red;
        }
        if (act) {
#pragma unroll
          for (int n = 0; n < NV; ++n) x[n] = x[n] + dt_red * r[n];
          kadd(acc[0][t], acc[2][t], dt_red * sC);
          kadd(acc[1][t], acc[3][t], dt_red * sN);
        }
        dt_int = dt_int + dt_red;
        need = true;
        // the lanes reading sF are done before the next evaluation writes it
        __syncthreads();
      }
    }
    if (stop) break;
    if (p.clamp && act) {
      bool nan = false;
#pragma unroll
      for (int n = 0; n < NV; ++n) {
        nan |= x[n] != x[n];
        if (x[n] < 0.0) {
          if (elem(n) >= 0)
            kadd(d.cellacc[(2 + elem(n)) * d.ncell + cell], d.cellacc[(7 + elem(n)) * d.ncell + cell], 0.0 - x[n]);
          x[n] = 0.0;
        }
      }
      if (nan) atomicOr(errflag, 1);
    }
  }
  if (act) {
#pragma unroll
    for (int n = 0; n < NV; ++n) d.conc[n * d.ncell + cell] = x[n];
    d.cellacc[cell] = acc[0][t];
    d.cellacc[d.ncell + cell] = acc[1][t];
    d.cellacc[5 * d.ncell + cell] = acc[2][t];
    d.cellacc[6 * d.ncell + cell] = acc[3][t];
  }
  if (own) {
    d.colacc[(long long)l * d.ncol + c] = acc[4][t];
    d.colacc[(long long)(NV + l) * d.ncol + c] = acc[5][t];
  }
  if (valid && l == 0) d.diag[c] = dg;
  if (threadIdx.x == 0) *min_dt = mind;
}

// minimum clamp after an ode_solver call (fabm_sediment_component.F90:1726-1731)
template <int LPC>
__global__ __launch_bounds__(BLOCK) void k_clamp(Dev d, const int *list, int nlist, int *errflag)
{
  constexpr int G = BLOCK / LPC;
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int slot = blockIdx.x * G + g;
  if (slot >= nlist) return;
  const int c = list[slot];
  const int o = d.off[c], K = d.off[c + 1] - o;
  if (l >= K) return;
  const long long cell = (long long)o + l;
  bool nan = false;
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const double x = d.conc[n * d.ncell + cell];
    nan |= x != x;
    if (x < 0.0) {
      if (elem(n) >= 0)
        kadd(d.cellacc[(2 + elem(n)) * d.ncell + cell], d.cellacc[(7 + elem(n)) * d.ncell + cell], 0.0 - x);
      d.conc[n * d.ncell + cell] = 0.0;
    }
  }
  if (nan) atomicOr(errflag, 1);
}

// ---------------------------------------------------------------------------
// column reductions: budget, inventory, steady-state residual
// ---------------------------------------------------------------------------

template <int LPC>
__global__ __launch_bounds__(BLOCK) void k_budget(Dev d, const int *list, int nlist, double *budget)
{
  constexpr int G = BLOCK / LPC;
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int slot = blockIdx.x * G + g;
  if (slot >= nlist) return;
  const int c = list[slot];
  const int o = d.off[c], K = d.off[c + 1] - o;
  const bool act = l < K;
  const long long cell = (long long)o + l;
  const double w = act ? d.phi[cell] * d.dz[cell] : 0.0;
  double *b = budget + (long long)c * SEDCOL_NBUD;
  if (l < NV) b[l] = d.colacc[(long long)l * d.ncol + c] + d.colacc[(long long)(NV + l) * d.ncol + c];
#pragma unroll
  for (int a = 0; a < 5; ++a) {
    const double v = act ? w * (d.cellacc[a * d.ncell + cell] + d.cellacc[(5 + a) * d.ncell + cell]) : 0.0;
    const double s = lg_sum<LPC>(v);
    if (l == 0) b[NV + a] = s;
  }
}

template <int LPC>
__global__ __launch_bounds__(BLOCK) void k_inventory(Dev d, const int *list, int nlist, double *inv)
{
  constexpr int G = BLOCK / LPC;
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int slot = blockIdx.x * G + g;
  if (slot >= nlist) return;
  const int c = list[slot];
  const int o = d.off[c], K = d.off[c + 1] - o;
  const bool act = l < K;
  const long long cell = (long long)o + l;
  const double w = act ? d.phi[cell] * d.dz[cell] : 0.0;
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const double v = act ? w * d.conc[n * d.ncell + cell] : 0.0;
    const double s = lg_sum<LPC>(v);
    if (l == 0) inv[(long long)n * d.ncol + c] = s;
  }
}

template <int LPC, int ARITH>
__global__ __launch_bounds__(BLOCK) void k_residual(Dev d, const int *list, int nlist, sedcol_params p, int rec,
                                                    const double *cfloor, double *resid, double *flux)
{
  constexpr int G = BLOCK / LPC;
  __shared__ double sbc[G][2 * NV];
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int slot = blockIdx.x * G + g;
  const bool valid = slot < nlist;
  const int c = valid ? list[slot] : 0;
  Lane L;
  Geo Gm;
  long long cell = 0;
  if (valid) {
    setup_lane(d, c, l, L, Gm, cell);
    load_forcing<ARITH>(d, p, c, rec, Gm, L, sbc[g]);
  }
  __syncthreads();
  if (!valid) return;
  double x[NV];
#pragma unroll
  for (int n = 0; n < NV; ++n) x[n] = L.active ? d.conc[n * d.ncell + cell] : 0.0;
  double sC, sN, m = 0.0;
  column_rhs<LPC, ARITH>(p, L, sbc[g], x, sC, sN, [&](int n, double r, double F) {
    const double ax = fabs(x[n]);
    const double den = ax > cfloor[n] ? ax : cfloor[n];
    const double e = fabs(r) / den;
    if (L.active && (e > m || e != e)) m = e;
    if (l == 0 && flux) flux[(long long)n * d.ncol + c] = F;
  }, &d.zw);
  m = lg_max<LPC>(m);
  if (l == 0) resid[c] = m;
}

// ---------------------------------------------------------------------------
// boundary conditions from the pelagic fields at the soil surface
// (get_boundary_conditions, fabm_sediment_component.F90:1865-2030): one
// thread per listed column, writing forcing record 0
// ---------------------------------------------------------------------------

__global__ void k_assemble(Dev d, const int *list, int nlist, double D0, int bcup_dis, unsigned present,
                           const double *ts, const double *cs, const double *ws, double *temp, double *cbw,
                           double *fup)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlist) return;
  const int c = list[i];
  const long long top = d.off[c];
  if (ts) temp[c] = ts[c];  // :1903-1939
  for (int n = 0; n < NV; ++n) {
    if (!((present >> n) & 1u)) continue;  // field absent: left as it is (:1963-1967)
    const long long k = (long long)n * d.ncol + c;
    if (part(n)) {
      fup[k] = -(cs[k] * ws[k]);  // :1978-1986, downward positive
    } else {
      cbw[k] = cs[k];  // :1989-2000
      if (bcup_dis == 1) {
        // :2012-2014, in its operation order
        double x = (d.conc[n * d.ncell + top] - cbw[k]) / d.dz[top];
        x = x * ((d.biot[c] + D0) + temp[c] * 0.035);
        x = x * d.phi[top];
        x = x / 86400.0;
        x = x / 10000.0;
        fup[k] = -x;
      } else {
        fup[k] = 0.0;  // :2021
      }
    }
  }
}

void launch_assemble(hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, unsigned present,
                     const double *ts, const double *cs, const double *ws, double *temp, double *cbw, double *fup)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_assemble, dim3((n + 255) / 256), dim3(256), 0, st, d, list, n, p.D0, (int)p.bcup_dis, present,
                     ts, cs, ws, temp, cbw, fup);
}

// ---------------------------------------------------------------------------
// launch helpers (host)
// ---------------------------------------------------------------------------

template <int LPC>
static inline dim3 grid_for(int n)
{
  constexpr int G = BLOCK / LPC;
  return dim3((unsigned)((n + G - 1) / G));
}

// kernel instance for the parameters: the extended reference form for
// BcUp = 4 and bioturbation profile 3 (the host marks the latter with
// arith = ARITH_EXT), the fast form where it applies, else the reference form
static inline int arith_form(const sedcol_params &p)
{
  if (p.arith == ARITH_EXT || p.bcup_part == 4) return ARITH_EXT;
  if (p.arith == SEDCOL_ARITH_FAST && p.bcdown != 2) return ARITH_FAST;
  return ARITH_REF;
}

#define SC_DISPATCH_LPC(lpc, ...)                                \
  switch (lpc) {                                                 \
    case 16: { constexpr int LPC = 16; __VA_ARGS__; } break;     \
    case 32: { constexpr int LPC = 32; __VA_ARGS__; } break;     \
    default: { constexpr int LPC = 64; __VA_ARGS__; } break;     \
  }
#define SC_DISPATCH_ARITH(p, ...)                                          \
  switch (arith_form(p)) {                                                 \
    case ARITH_FAST: { constexpr int AR = ARITH_FAST; __VA_ARGS__; } break; \
    case ARITH_EXT: { constexpr int AR = ARITH_EXT; __VA_ARGS__; } break;   \
    default: { constexpr int AR = ARITH_REF; __VA_ARGS__; } break;          \
  }

void launch_rhs(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, int rec,
                const double *conc, double *rhs, double *flux, double *sinks)
{
  SC_DISPATCH_ARITH(p, SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_rhs<LPC, AR>), grid_for<LPC>(n), dim3(BLOCK), 0,
                                                               st, d, list, n, p, rec, conc, rhs, flux, sinks)));
}

void launch_fixed(int lpc, int method, hipStream_t st, const Dev &d, const int *list, int n,
                  const sedcol_params &p, int rec, double dt, int nsteps, int *errflag)
{
#define SC_FIX(M, W)                                                                                             \
  SC_DISPATCH_ARITH(p, SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_fixed<LPC, M, AR, W>), grid_for<LPC>(n),        \
                                                               dim3(BLOCK), 0, st, d, list, n, p, rec, dt, nsteps, \
                                                               errflag)))
  const bool geo = p.emp_weights == 1;
  switch (method) {
    case SEDCOL_EULER: SC_FIX(SEDCOL_EULER, 0); break;
    case SEDCOL_RUNGE_KUTTA_4: SC_FIX(SEDCOL_RUNGE_KUTTA_4, 0); break;
    case SEDCOL_RUNGE_KUTTA_4_38: SC_FIX(SEDCOL_RUNGE_KUTTA_4_38, 0); break;
    case SEDCOL_EMP1:
      if (geo) {
        SC_FIX(SEDCOL_EMP1, 1);
      } else {
        SC_FIX(SEDCOL_EMP1, 0);
      }
      break;
    default:
      if (geo) {
        SC_FIX(SEDCOL_EMP2, 1);
      } else {
        SC_FIX(SEDCOL_EMP2, 0);
      }
      break;
  }
#undef SC_FIX
}

// blocks of a work-queue kernel that are resident at once on the current
// device, cached per device and kernel instance (threads that race on a
// first query store the same value).  PW = 2 selects the linearly implicit
// kernel k_ros_adaptive, PW = 3 k_emp_adaptive2 (two layers per lane),
// otherwise k_emp_adaptive with Patankar weights PW.
template <int LPC, int AR, int BLK, int PW>
static int emp_resident()
{
  constexpr int kMaxDev = 64;
  static std::atomic<int> cached[kMaxDev];  // static storage: starts at zero
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  std::atomic<int> *slot = (dev >= 0 && dev < kMaxDev) ? &cached[dev] : nullptr;
  int v = slot ? slot->load(std::memory_order_relaxed) : 0;
  if (v > 0) return v;
  hipDeviceProp_t prop;
  int per_cu = 0;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 1024;
  hipError_t e;
  if constexpr (PW == 2)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_ros_adaptive<LPC, AR, BLK>, BLK, 0);
  else if constexpr (PW == 3)
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_emp_adaptive2<LPC, AR, BLK>, BLK, 0);
  else
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_emp_adaptive<LPC, AR, BLK, PW>, BLK, 0);
  if (e != hipSuccess) per_cu = 1;
  v = (per_cu > 0 ? per_cu : 1) * prop.multiProcessorCount;
  if (slot) slot->store(v, std::memory_order_relaxed);
  return v;
}

#define SC_DISPATCH_BLK(blk, ...)                                \
  switch (blk) {                                                 \
    case 64: { constexpr int BLK = 64; __VA_ARGS__; } break;     \
    case 128: { constexpr int BLK = 128; __VA_ARGS__; } break;   \
    case 256: { constexpr int BLK = 256; __VA_ARGS__; } break;   \
    default: return false;                                       \
  }
#define SC_DISPATCH_PW(p, ...)                                   \
  switch ((p).emp_weights) {                                     \
    case 1: { constexpr int PW = 1; __VA_ARGS__; } break;        \
    default: { constexpr int PW = 0; __VA_ARGS__; } break;       \
  }

static bool emp_shape_ok(int lpc, int blk)
{
  return (lpc == 16 || lpc == 32 || lpc == 64) && (blk == 64 || blk == 128 || blk == 256);
}

// the linearly implicit kernel has the reference and fast forms only
#define SC_DISPATCH_ROS_ARITH(p, ...)                                          \
  switch (arith_form(p)) {                                                     \
    case ARITH_FAST: { constexpr int AR = ARITH_FAST; __VA_ARGS__; } break;     \
    case ARITH_REF: { constexpr int AR = ARITH_REF; __VA_ARGS__; } break;       \
    default: return false;                                                     \
  }

static bool ros_selected(const sedcol_params &p) { return p.adaptive_method == SEDCOL_ROS2_ADAPTIVE; }

bool emp2_supported(const sedcol_params &p)
{
  return !ros_selected(p) && p.emp_weights == 0 && arith_form(p) != ARITH_EXT;
}

int emp_resident_blocks(int lpc, int blk, int lpl, const sedcol_params &p)
{
  if (!emp_shape_ok(lpc, blk)) return 0;
  if (lpl == 2 && (lpc < 16 || !emp2_supported(p))) return 0;
  int out = 0;
  auto body = [&]() -> bool {
    if (lpl == 2) {
      SC_DISPATCH_ROS_ARITH(p, SC_DISPATCH_LPC(lpc, SC_DISPATCH_BLK(blk, out = emp_resident<LPC, AR, BLK, 3>())));
    } else if (ros_selected(p)) {
      SC_DISPATCH_ROS_ARITH(p, SC_DISPATCH_LPC(lpc, SC_DISPATCH_BLK(blk, out = emp_resident<LPC, AR, BLK, 2>())));
    } else {
      SC_DISPATCH_PW(p, SC_DISPATCH_ARITH(p, SC_DISPATCH_LPC(lpc, SC_DISPATCH_BLK(blk, out = emp_resident<LPC, AR, BLK, PW>()))));
    }
    return true;
  };
  body();
  return out;
}

bool launch_emp_adaptive(int lpc, int blk, int lpl, hipStream_t st, const Dev *d_dev, const int *list, int n,
                         const sedcol_params &p, const sedcol_params *d_params, const EmpCtl *d_ctl, int *next,
                         int *pending, int *work, int *errflag)
{
  if (!emp_shape_ok(lpc, blk)) return false;
  if (lpl == 2) {
    if (lpc < 16 || !emp2_supported(p)) return false;
    SC_DISPATCH_ROS_ARITH(p, SC_DISPATCH_LPC(lpc, SC_DISPATCH_BLK(blk, {
      const int G = BLK / LPC;
      const unsigned need = (unsigned)((n + G - 1) / G);
      const unsigned cap = (unsigned)emp_resident<LPC, AR, BLK, 3>();
      hipLaunchKernelGGL((k_emp_adaptive2<LPC, AR, BLK>), dim3(need < cap ? need : cap), dim3(BLK), 0, st, d_dev,
                         list, n, d_params, d_ctl, next, pending, work, errflag);
    })));
    return true;
  }
  if (lpl != 1) return false;
  if (ros_selected(p)) {
    SC_DISPATCH_ROS_ARITH(p, SC_DISPATCH_LPC(lpc, SC_DISPATCH_BLK(blk, {
      const int G = BLK / LPC;
      const unsigned need = (unsigned)((n + G - 1) / G);
      const unsigned cap = (unsigned)emp_resident<LPC, AR, BLK, 2>();
      hipLaunchKernelGGL((k_ros_adaptive<LPC, AR, BLK>), dim3(need < cap ? need : cap), dim3(BLK), 0, st, d_dev,
                         list, n, d_params, d_ctl, next, pending, work, errflag);
    })));
    return true;
  }
  SC_DISPATCH_PW(p, SC_DISPATCH_ARITH(p, SC_DISPATCH_LPC(lpc, SC_DISPATCH_BLK(blk, {
    const int G = BLK / LPC;
    const unsigned need = (unsigned)((n + G - 1) / G);
    const unsigned cap = (unsigned)emp_resident<LPC, AR, BLK, PW>();
    hipLaunchKernelGGL((k_emp_adaptive<LPC, AR, BLK, PW>), dim3(need < cap ? need : cap), dim3(BLK), 0, st, d_dev,
                       list, n, d_params, d_ctl, next, pending, work, errflag);
  }))));
  return true;
}

__global__ void k_emp_begin(const int *list, int n, int r0, int *wrec, double *wrem, long long *watt)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = list[i];
    wrec[c] = r0 - 1;
    wrem[c] = 0.0;
    watt[c] = 0;
  }
}

// zc from sequential sums of dz (init_grid :161-168), exp(100 k_l zc),
// exp(200 k_l zc), and the depth to the last layer's top (the cumdepth left by
// the light loop of get_rhs :593-598); one thread per column, in the
// oracle's operation order
__global__ void k_zw_geometry(Dev d, const int *list, int n, double k_l, double *e1, double *e2, double *cum)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = list[i];
  const int o = d.off[c], K = d.off[c + 1] - o;
  double zi = 0.0, cd = 0.0;
  for (int k = 0; k < K; ++k) {
    const double dz = d.dz[o + k];
    const double zc = zi + dz * 0.5;
    e1[o + k] = dexp((zc * 100.0) * k_l);
    e2[o + k] = dexp((zc * 200.0) * k_l);
    zi = zi + dz;
    if (k < K - 1) cd = cd + dz;
  }
  cum[c] = cd;
}

void launch_zw_geometry(hipStream_t st, const Dev &d, const int *list, int n, double k_l, double *e1, double *e2,
                        double *cum)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zw_geometry, dim3((n + 255) / 256), dim3(256), 0, st, d, list, n, k_l, e1, e2, cum);
}

__global__ void k_fill_int(int *p, int n, int v)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_emp_begin(hipStream_t st, const int *list, int n, int r0, int *wrec, double *wrem, long long *watt)
{
  if (n <= 0) return;
  const int blocks = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
  hipLaunchKernelGGL(k_emp_begin, dim3(blocks), dim3(256), 0, st, list, n, r0, wrec, wrem, watt);
}

// ---------------------------------------------------------------------------
// warm start from log10 profiles on a geometric node grid
// ---------------------------------------------------------------------------

// ratio r of an n-layer geometric grid, dz1 (r^n - 1)/(r - 1) = total, by
// Newton's method on g(r) = dz1 (r^n - 1) - total (r - 1), which is convex
// for r > 0 with the roots 1 and the ratio.  A growing grid (n dz1 < total)
// has the larger root: from r0 = (total/dz1)^(1/(n-1)), where g(r0) =
// total - dz1 > 0, the iterates decrease monotonically onto it.  A
// shrinking grid has the smaller root: from r = 0 they increase onto it.
// The iteration stops when an iterate no longer moves in its direction
// (at most a few ulp from the root; the host's bisection lands as close).
__device__ double geometric_ratio_dev(int n, double dz1, double total)
{
  const double fn = (double)n;
  if (n <= 1 || fabs(fn * dz1 - total) <= 1e-15 * total) return 1.0;
  const bool grow = fn * dz1 < total;
  double r = grow ? pow(total / dz1, 1.0 / (fn - 1.0)) : 0.0;
  for (int it = 0; it < 100; ++it) {
    const double rm = pow(r, fn - 1.0);  // r^(n-1)
    const double g = dz1 * (rm * r - 1.0) - total * (r - 1.0);
    const double dg = (dz1 * fn) * rm - total;
    const double rn = r - g / dg;
    if (!(grow ? rn < r : rn > r)) break;
    r = rn;
  }
  return r;
}

// One column per group of LPC lanes, lane l = layer l.  The group builds
// the node centres in LDS (lanes stride over the nodes), then each layer
// interpolates every variable's profile at its own centre.
template <int LPC>
__global__ __launch_bounds__(BLOCK) void k_warm_profiles(Dev d, const int *list, int nlist, WarmProfiles w)
{
  constexpr int G = BLOCK / LPC;
  __shared__ double sdz[G][kWarmMaxNodes];
  __shared__ double szc[G][kWarmMaxNodes];
  const int g = threadIdx.x / LPC, l = threadIdx.x % LPC;
  const int slot = blockIdx.x * G + g;
  const bool live = slot < nlist;
  const int c = live ? list[slot] : 0;
  const int o = live ? d.off[c] : 0, K = live ? d.off[c + 1] - o : 0;
  if (live) {
    double depth = 0.0;
    if (w.depth)
      depth = w.depth[c];
    else
      for (int k = 0; k < K; ++k) depth = depth + d.dz[o + k];
    const double r = geometric_ratio_dev(w.nz, w.dz1, depth);
    for (int j = l; j < w.nz; j += LPC) sdz[g][j] = w.dz1 * pow(r, (double)j);
  }
  __syncthreads();
  if (live)
    for (int j = l; j < w.nz; j += LPC) {
      double zi = 0.0;
      for (int q = 0; q < j; ++q) zi = zi + sdz[g][q];
      szc[g][j] = zi + 0.5 * sdz[g][j];
    }
  __syncthreads();
  if (!live || l >= K) return;
  double zi = 0.0;
  for (int q = 0; q < l; ++q) zi = zi + d.dz[o + q];
  const double zk = zi + 0.5 * d.dz[o + l];
  int j = -1;
  for (int q = 0; q < w.nz; ++q) j += zk >= szc[g][q] ? 1 : 0;
  j = j < 0 ? 0 : (j > w.nz - 2 ? w.nz - 2 : j);
  const double z0 = szc[g][j], z1 = szc[g][j + 1];
  double wt = (zk - z0) / (z1 - z0);
  wt = wt < 0.0 ? 0.0 : (wt > 1.0 ? 1.0 : wt);
  const float *row = w.prof + (long long)c * w.ld;
  const long long cell = (long long)o + l;
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const int i0 = n * w.nz + j;
    const double f0 = (double)row[i0] * w.scale[i0] + w.offset[i0];
    const double f1 = (double)row[i0 + 1] * w.scale[i0 + 1] + w.offset[i0 + 1];
    const double x = exp10(f0 + wt * (f1 - f0)) - w.eps[n];
    d.conc[n * d.ncell + cell] = x < 0.0 ? 0.0 : x;  // NaN stays NaN
  }
}

void launch_warm_profiles(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const WarmProfiles &w)
{
  SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_warm_profiles<LPC>), grid_for<LPC>(n), dim3(BLOCK), 0, st, d, list, n,
                                          w));
}

void launch_fill_int(hipStream_t st, int *p, int n, int v)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
}

__global__ void k_fill_dbl(double *p, long long n, double v)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill_dbl(hipStream_t st, double *p, long long n, double v)
{
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_dbl, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
}

// Fortran state conc(i,j,k,n) = f[(n K + k) ncol + c] against the packed
// state x[n ncell + c K + k] of a uniform grid (ncell = ncol K); one thread
// per element, in Fortran order (reads or writes of f coalesced).
__global__ void k_layout(const double *src, double *dst, int ncol, int K, int to_packed)
{
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long plane = (long long)ncol * K;
  if (i >= (long long)NV * plane) return;
  const long long n = i / plane;
  const long long r = i - n * plane;
  const long long k = r / ncol;
  const long long c = r - k * ncol;
  const long long j = n * plane + c * K + k;
  if (to_packed) dst[j] = src[i];
  else dst[i] = src[j];
}

void launch_layout(hipStream_t st, const double *src, double *dst, int ncol, int K, bool to_packed)
{
  const long long n = (long long)NV * ncol * K;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_layout, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, src, dst, ncol, K,
                     to_packed ? 1 : 0);
}

void launch_rhs_store(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, int rec,
                      double *r_out, double *s_out)
{
  SC_DISPATCH_ARITH(p, SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_rhs_store<LPC, AR>), grid_for<LPC>(n), dim3(BLOCK),
                                                               0, st, d, list, n, p, rec, r_out, s_out)));
}

void launch_ae_check(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double dt_red, double rcm,
                     const double *r_in, int *viol)
{
  SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_ae_check<LPC>), grid_for<LPC>(n), dim3(BLOCK), 0, st, d, list, n,
                                          dt_red, rcm, r_in, viol));
}

void launch_ae_commit(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double dt_red,
                      const double *r_in, const double *s_in)
{
  SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_ae_commit<LPC>), grid_for<LPC>(n), dim3(BLOCK), 0, st, d, list, n,
                                          dt_red, r_in, s_in));
}

void launch_ae_fused(int lpc, hipStream_t st, const Dev *d_dev, const int *list, int n, const sedcol_params &p,
                     const sedcol_params *d_params, int rec, double dt, int nsteps, int *errflag, double *min_dt)
{
  SC_DISPATCH_ARITH(p, SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_ae_fused<LPC, AR>), dim3(1), dim3(kAeFusedBlock),
                                                                0, st, d_dev, list, n, d_params, rec, dt, nsteps,
                                                                errflag, min_dt)));
}

void launch_clamp(int lpc, hipStream_t st, const Dev &d, const int *list, int n, int *errflag)
{
  SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_clamp<LPC>), grid_for<LPC>(n), dim3(BLOCK), 0, st, d, list, n,
                                          errflag));
}

void launch_budget(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double *budget)
{
  SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_budget<LPC>), grid_for<LPC>(n), dim3(BLOCK), 0, st, d, list, n,
                                          budget));
}

void launch_inventory(int lpc, hipStream_t st, const Dev &d, const int *list, int n, double *inv)
{
  SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_inventory<LPC>), grid_for<LPC>(n), dim3(BLOCK), 0, st, d, list, n,
                                          inv));
}

void launch_residual(int lpc, hipStream_t st, const Dev &d, const int *list, int n, const sedcol_params &p, int rec,
                     const double *cfloor, double *resid, double *flux)
{
  SC_DISPATCH_ARITH(p, SC_DISPATCH_LPC(lpc, hipLaunchKernelGGL((k_residual<LPC, AR>), grid_for<LPC>(n), dim3(BLOCK),
                                                               0, st, d, list, n, p, rec, cfloor, resid, flux)));
}

}  // namespace sc
