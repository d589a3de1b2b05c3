// sedcol_device.hpp -- device building blocks of the sediment-column kernels.
//
// Execution model: one column per lane group of LPC lanes (LPC = 16, 32 or
// 64, a power of two >= K), lane l holding layer l with all NV state
// variables in VGPRs.  Vertical neighbours move between lanes by shuffles;
// column sums, products and maxima use an xor butterfly, i.e. a balanced
// pairwise tree whose grouping does not depend on LPC (padding lanes hold the
// identity).
//
// Floating point: built with -ffp-contract=off; fused multiply-adds are
// explicit fma() calls.  Operation order follows the reference Fortran where
// the reference defines the arithmetic (cited per function).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sedcol.h"

namespace sc {

constexpr int NV = SEDCOL_NVAR;

// particulate flag per state variable (fabm_sediment_driver.F90:658)
__host__ __device__ constexpr bool part(int n) { return n <= 4 || n == 10; }
// element index for clamp accounting: 0 C, 1 N, 2 P, -1 none
__host__ __device__ constexpr int elem(int n)
{
  return (n == 0 || n == 1) ? 0 : (n == 2 || n == 3 || n == 6 || n == 7) ? 1 : (n == 4 || n == 9 || n == 10) ? 2 : -1;
}

__device__ __forceinline__ double nn(double c) { return c < 0.0 ? 0.0 : c; }

// exp(x): Cody-Waite reduction with the fdlibm two-part ln2, degree-13 Taylor
// polynomial in Horner form, 2^n through the exponent field.  Normal range
// only.  The fixed sequence of IEEE operations makes the result identical on
// every IEEE-754 FP64 unit.
__device__ __forceinline__ double dexp(double x)
{
  if (x != x) return x;
  if (x > 709.0) return __longlong_as_double(0x7FF0000000000000ll);
  if (x < -708.0) return 0.0;
  const double log2e = 1.4426950408889634;
  const double ln2hi = 6.93147180369123816490e-01;
  const double ln2lo = 1.90821492927058770002e-10;
  double n = rint(x * log2e);
  double r = (x - n * ln2hi) - n * ln2lo;
  double s = 1.6059043836821613e-10;
  s = fma(s, r, 2.08767569878681e-09);
  s = fma(s, r, 2.505210838544172e-08);
  s = fma(s, r, 2.755731922398589e-07);
  s = fma(s, r, 2.7557319223985893e-06);
  s = fma(s, r, 2.48015873015873e-05);
  s = fma(s, r, 0.0001984126984126984);
  s = fma(s, r, 0.001388888888888889);
  s = fma(s, r, 0.008333333333333333);
  s = fma(s, r, 0.041666666666666664);
  s = fma(s, r, 0.16666666666666666);
  s = fma(s, r, 0.5);
  s = fma(s, r, 1.0);
  s = fma(s, r, 1.0);
  long long e = (long long)n + 1023;
  return s * __longlong_as_double(e << 52);
}

// 1/sqrt(x), x > 0 finite: exponent-halving guess plus four Newton steps
// (step-size controller only).
__device__ __forceinline__ double drsqrt(double x)
{
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  b = 0x5FE6EB50C7B537A9ull - (b >> 1);
  double y = __longlong_as_double((long long)b);
  double h = 0.5 * x;
#pragma unroll
  for (int i = 0; i < 4; ++i) y = y * (1.5 - (h * y) * y);
  return y;
}

// temperature factor of get_rhs, fabm_sediment_driver.F90:648
__device__ __forceinline__ double temp_factor(double T)
{
  return dexp(-4500.0 * (1.0 / (T + 273.0) - (1.0 / 288.0)));
}

// Local reaction rates at one cell: OMEXDIA-type C/N/P/O2/ODU kinetics
// (Soetaert et al. 1996) with the parameter names of
// examples/standalone/omexdia_p/fabm_sed.nml:51-77.  The reference obtains
// these rates from FABM's hzg_omexdia_p through fabm_do
// (fabm_sediment_driver.F90:700); that module is not part of the reference,
// so this is the library's own definition.  Each destruction term is
// proportional to the destroyed species.  sC = DIC release, sN = N2 release.
__device__ __forceinline__ void kinetics(const sedcol_params &p, double fT, const double (&c)[NV],
                                         double (&R)[NV], double &sC, double &sN)
{
  double ldC = nn(c[0]), sdC = nn(c[1]), ldN = nn(c[2]), sdN = nn(c[3]);
  double dP = nn(c[4]), O2 = nn(c[5]), NO3 = nn(c[6]), NH4 = nn(c[7]);
  double ODU = nn(c[8]), PO4 = nn(c[9]), aP = nn(c[10]);

  double kL = fT * p.rL;
  double kS = fT * p.rS;
  double dL = kL * ldC;
  double dS = kS * sdC;
  double lim = p.CprodMax / (p.CprodMax + (dL + dS));
  double CpL = dL * lim;
  double CpS = dS * lim;
  double Cprod = CpL + CpS;
  double NpL = (kL * ldN) * lim;
  double NpS = (kS * sdN) * lim;
  double Pprod = (kL * dP) * lim;

  double oxlim = O2 / (O2 + p.ksO2oxic);
  double dnlim = (p.kinO2denit / (O2 + p.kinO2denit)) * (NO3 / (NO3 + p.ksNO3denit));
  double anlim = (p.kinO2anox / (O2 + p.kinO2anox)) * (p.kinNO3anox / (NO3 + p.kinNO3anox));
  double rescale = 1.0 / ((oxlim + dnlim) + anlim);
  double OxicMin = (Cprod * oxlim) * rescale;
  double Denit = (Cprod * dnlim) * rescale;
  double AnoxMin = (Cprod * anlim) * rescale;

  double Nitri = ((fT * p.rnit) * NH4) * (O2 / (O2 + p.ksO2nitri));
  double o2odu = O2 / (O2 + p.ksO2oduox);
  double OduOx = ((fT * p.rODUox) * ODU) * o2odu;
  double kP = fT * p.rPads;
  double Pads = (kP * PO4) * o2odu;
  double Pdes = (kP * aP) * (ODU / (ODU + p.PAdsODU));

  R[0] = -CpL;
  R[1] = -CpS;
  R[2] = -NpL;
  R[3] = -NpS;
  R[4] = -Pprod;
  R[5] = -((OxicMin + 2.0 * Nitri) + OduOx);
  R[6] = Nitri - 0.8 * Denit;
  R[7] = (NpL + NpS) - Nitri;
  R[8] = AnoxMin - OduOx;
  R[9] = (Pprod - Pads) + Pdes;
  R[10] = Pads - Pdes;
  sC = Cprod;
  sN = 0.8 * Denit;
}

// ---------------------------------------------------------------------------
// lane-group exchange (groups of LPC lanes inside a wave64)
// ---------------------------------------------------------------------------

// value held by lane l-1 of the group (lane 0: own value, never used)
template <int LPC>
__device__ __forceinline__ double lg_up(double v)
{
  return __shfl_up(v, 1u, LPC);
}
// value held by lane l+1 of the group (last lane: own value, never used)
template <int LPC>
__device__ __forceinline__ double lg_dn(double v)
{
  return __shfl_down(v, 1u, LPC);
}
template <int LPC>
__device__ __forceinline__ double lg_bcast0(double v)
{
  return __shfl(v, 0, LPC);
}

// butterfly sum: every lane receives the balanced-tree sum over the group
template <int LPC>
__device__ __forceinline__ double lg_sum(double v)
{
#pragma unroll
  for (int s = 1; s < LPC; s <<= 1) v = v + __shfl_xor(v, s, LPC);
  return v;
}
template <int LPC>
__device__ __forceinline__ double lg_max(double v)
{
#pragma unroll
  for (int s = 1; s < LPC; s <<= 1) {
    double o = __shfl_xor(v, s, LPC);
    v = o > v ? o : v;
  }
  return v;
}
// (P, D) product-with-derivative pairs: (P1,D1)(P2,D2) = (P1 P2, D1 P2 + P1 D2)
template <int LPC>
__device__ __forceinline__ void lg_pd(double &P, double &D)
{
#pragma unroll
  for (int s = 1; s < LPC; s <<= 1) {
    double Po = __shfl_xor(P, s, LPC);
    double Do = __shfl_xor(D, s, LPC);
    D = D * Po + P * Do;
    P = P * Po;
  }
}

// ---------------------------------------------------------------------------
// per-lane column data
// ---------------------------------------------------------------------------

struct Lane {
  int l;        // layer index (lane in group)
  int K;        // layers of the column
  bool active;  // l < K
  double dz, dzc, phi;
  double Dbio, Ddis;  // interface diffusivities for the current temperature
  double fT;
};

// Interface diffusivities of get_rhs (fabm_sediment_driver.F90:652-653 and
// :682-683, division order as written there).
__device__ __forceinline__ void diffusivities(const sedcol_params &p, double B, double T, double phiI,
                                             double bfac, Lane &L)
{
  L.fT = temp_factor(T);
  L.Dbio = ((((B * L.fT) / 86400.0) / 10000.0) * (1.0 - phiI)) * bfac;
  L.Ddis = L.Dbio + (((p.D0 + T * 0.035) * phiI) / 86400.0) / 10000.0;
}

// One get_rhs evaluation for the lane's layer.  sbc[0..NV) bottom-water
// concentrations, sbc[NV..2NV) prescribed top fluxes (LDS, per group).
// For each variable n the callback receives (n, rhs_n, F) where F is the flux
// through the lane's upper interface (lane 0: the top flux intFlux(:,:,1)).
// Transport: diff3d (fabm_sediment_driver.F90:739-825) with the particulate
// bulk/volume-fraction handling of get_rhs :658-678 and dissolved :680-692.
template <int LPC, class CB>
__device__ __forceinline__ void column_rhs(const sedcol_params &p, const Lane &L, const double *sbc,
                                           const double (&c)[NV], double &sC, double &sN, CB &&cb)
{
  double R[NV];
  kinetics(p, L.fT, c, R, sC, sN);
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const bool pt = part(n);
    const double x = pt ? c[n] * L.phi : c[n];
    const double xup = lg_up<LPC>(x);
    const double D = pt ? L.Dbio : L.Ddis;
    double F;
    if (L.l == 0) {
      const int bc = pt ? p.bcup_part : p.bcup_dis;
      F = bc == 1 ? sbc[NV + n] : bc == 2 ? -((D * (x - sbc[n])) / L.dz) : 0.0;
    } else {
      F = -((D * (x - xup)) / L.dzc);
    }
    double Fdn = lg_dn<LPC>(F);
    if (L.l == L.K - 1) Fdn = p.bcdown == 2 ? -((D * (0.0 - x)) / L.dz) : 0.0;
    const double VF = pt ? (1.0 - L.phi) : L.phi;
    const double dC = (F - Fdn) / (VF * L.dz);
    const double Q = pt ? (dC * (1.0 - L.phi)) / L.phi : dC;
    cb(n, R[n] + Q, F);
  }
}

// Patankar factor of one column: root p in (0,1] of
// p = prod_{j: g_j<0} (1 + h g_j p / c_j) (Bruggeman et al. 2007), by Newton
// from p = 0 on the convex decreasing h(p) = prod - p.
template <int LPC>
__device__ __forceinline__ double psolve(const double (&c)[NV], const double (&g)[NV], double h, bool active,
                                         int64_t &zero_sink)
{
  double a[NV];
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    double av = 0.0;
    if (active && g[n] < 0.0) {
      if (c[n] > 0.0) av = (h * g[n]) / c[n];
      else zero_sink += 1;
    }
    a[n] = av;
  }
  double pp = 0.0;
  for (int it = 0; it < 60; ++it) {
    double P = 1.0, D = 0.0;
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      double f = fma(a[n], pp, 1.0);
      D = D * f + P * a[n];
      P = P * f;
    }
    lg_pd<LPC>(P, D);
    double hv = P - pp;
    double hp = D - 1.0;
    double pn = pp - hv / hp;
    if (!(pn > pp)) break;
    bool conv = (pn - pp) <= 0x1p-50 * pn;
    pp = pn;
    if (conv) break;
  }
  return pp > 1.0 ? 1.0 : pp;
}

}  // namespace sc
