// sedcol_device.hpp -- device building blocks of the sediment-column kernels.
//
// Execution model: one column per lane group of LPC lanes (LPC = 16, 32 or
// 64, a power of two >= K), lane l holding layer l with all NV state
// variables in VGPRs.  Vertical neighbours move between lanes with DPP
// wave shifts; column sums, products and maxima use a butterfly (DPP quad
// permutes and row mirrors, a 32-lane swizzle, and a bpermute for the
// 64-lane level), i.e. a balanced pairwise tree whose grouping does not
// depend on LPC (padding lanes hold the identity).
//
// Floating point: built with -ffp-contract=off; fused multiply-adds are
// explicit fma() calls.  Two arithmetic forms exist (template ARITH):
//   SEDCOL_ARITH_REFERENCE  the operation order of the reference Fortran
//                           (fabm_sediment_driver.F90:652-692, :776-820);
//   SEDCOL_ARITH_FAST       geometry divisions replaced by per-lane
//                           coefficients computed once per forcing record
//                           (D/dzc, 1/(phi dz)) and merged Monod quotients;
//                           same mathematics, different rounding.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sedcol.h"

namespace sc {

constexpr int NV = SEDCOL_NVAR;
constexpr int ARITH_REF = SEDCOL_ARITH_REFERENCE;
constexpr int ARITH_FAST = SEDCOL_ARITH_FAST;
// the reference form with the transport options that need per-evaluation
// work (BcUp = 4, bioturbation profile 3); a kernel instance of its own, so
// that the plain reference form keeps its register budget
constexpr int ARITH_EXT = 2;

// particulate flag per state variable (fabm_sediment_driver.F90:658)
__host__ __device__ constexpr bool part(int n) { return n <= 4 || n == 10; }
// element index for clamp accounting: 0 C, 1 N, 2 P, -1 none
__host__ __device__ constexpr int elem(int n)
{
  return (n == 0 || n == 1) ? 0 : (n == 2 || n == 3 || n == 6 || n == 7) ? 1 : (n == 4 || n == 9 || n == 10) ? 2 : -1;
}

__device__ __forceinline__ double nn(double c) { return c < 0.0 ? 0.0 : c; }

// Neumaier-compensated running sum: s + c carries the total of every x added.
// Same operation order as the oracle's kadd, so sums and compensations match
// bit for bit.
__device__ __forceinline__ void kadd(double &s, double &c, double x)
{
  const double t = s + x;
  c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
  s = t;
}

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

// Natural logarithm with the oracle's operation sequence (oc_log): x = m 2^e,
// m in [sqrt(1/2), sqrt(2)], 2 atanh((m-1)/(m+1)) to s^25, plus e ln2.
__device__ __forceinline__ double dlog(double x)
{
  if (x != x) return x;
  if (!(x > 0.0)) return x == 0.0 ? -__longlong_as_double(0x7FF0000000000000ll) : __longlong_as_double(0x7FF8000000000000ll);
  if (x == __longlong_as_double(0x7FF0000000000000ll)) return x;
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  double k = 0.0;
  if (((b >> 52) & 0x7ffull) == 0) {
    x = x * 0x1p54;
    b = (unsigned long long)__double_as_longlong(x);
    k = -54.0;
  }
  long long e = (long long)((b >> 52) & 0x7ffull) - 1023;
  double m = __longlong_as_double((long long)((b & 0x000fffffffffffffull) | 0x3ff0000000000000ull));
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e = e + 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double s2 = s * s;
  double t = 1.0 / 25.0;
  t = fma(t, s2, 1.0 / 23.0);
  t = fma(t, s2, 1.0 / 21.0);
  t = fma(t, s2, 1.0 / 19.0);
  t = fma(t, s2, 1.0 / 17.0);
  t = fma(t, s2, 1.0 / 15.0);
  t = fma(t, s2, 1.0 / 13.0);
  t = fma(t, s2, 1.0 / 11.0);
  t = fma(t, s2, 1.0 / 9.0);
  t = fma(t, s2, 1.0 / 7.0);
  t = fma(t, s2, 1.0 / 5.0);
  t = fma(t, s2, 1.0 / 3.0);
  t = fma(t, s2, 1.0);
  const double lm = (2.0 * s) * t;
  k = k + (double)e;
  const double ln2hi = 6.93147180369123816490e-01;
  const double ln2lo = 1.90821492927058770002e-10;
  return fma(k, ln2hi, fma(k, ln2lo, lm));
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
template <int ARITH>
__device__ __forceinline__ void kinetics(const sedcol_params &p, double fT, const double (&c)[NV], double (&R)[NV],
                                         double &sC, double &sN)
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

  double oxlim, dnlim, anlim, nitlim, o2odu, odulim;
  if (ARITH != ARITH_FAST) {
    oxlim = O2 / (O2 + p.ksO2oxic);
    dnlim = (p.kinO2denit / (O2 + p.kinO2denit)) * (NO3 / (NO3 + p.ksNO3denit));
    anlim = (p.kinO2anox / (O2 + p.kinO2anox)) * (p.kinNO3anox / (NO3 + p.kinNO3anox));
    nitlim = O2 / (O2 + p.ksO2nitri);
    o2odu = O2 / (O2 + p.ksO2oduox);
    odulim = ODU / (ODU + p.PAdsODU);
  } else {
    // the six Monod denominators share one reciprocal of their product,
    // split back with prefix/suffix products (range: concentrations < 1e30,
    // half-saturation constants >= 1e-6)
    const double q2 = O2 + p.ksO2oxic;
    const double q3 = (O2 + p.kinO2denit) * (NO3 + p.ksNO3denit);
    const double q4 = (O2 + p.kinO2anox) * (NO3 + p.kinNO3anox);
    const double q6 = O2 + p.ksO2nitri;
    const double q7 = O2 + p.ksO2oduox;
    const double q8 = ODU + p.PAdsODU;
    const double a23 = q2 * q3, a234 = a23 * q4, a2346 = a234 * q6, a23467 = a2346 * q7;
    const double inv = 1.0 / (a23467 * q8);
    const double b87 = q8 * q7, b876 = b87 * q6, b8764 = b876 * q4, b87643 = b8764 * q3;
    oxlim = O2 * (inv * b87643);
    dnlim = (p.kinO2denit * NO3) * ((inv * q2) * b8764);
    anlim = (p.kinO2anox * p.kinNO3anox) * ((inv * a23) * b876);
    nitlim = O2 * ((inv * a234) * b87);
    o2odu = O2 * ((inv * a2346) * q8);
    odulim = ODU * (inv * a23467);
  }
  double rescale = 1.0 / ((oxlim + dnlim) + anlim);
  double OxicMin = (Cprod * oxlim) * rescale;
  double Denit = (Cprod * dnlim) * rescale;
  double AnoxMin = (Cprod * anlim) * rescale;

  double Nitri = ((fT * p.rnit) * NH4) * nitlim;
  double OduOx = ((fT * p.rODUox) * ODU) * o2odu;
  double kP = fT * p.rPads;
  double Pads = (kP * PO4) * o2odu;
  double Pdes = (kP * aP) * odulim;

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
// Jacobian of the local rates (linearly implicit stepper, method 7)
// ---------------------------------------------------------------------------

// Nonzero partial derivatives of the processes of kinetics<> with respect to
// the state, at max(c, 0), from the reference-form expressions (the W
// matrix of the stepper needs no more than that; both arithmetic forms use
// it).  Same operation sequence as the oracle's oc_kin_jac.
struct KJac {
  double cl0, cl1, cs0, cs1;  // dCpL/d(ldC,sdC), dCpS/d(ldC,sdC)
  double nl0, nl1, nl2;       // dNpL/d(ldC,sdC,ldN)
  double ns0, ns1, ns3;       // dNpS/d(ldC,sdC,sdN)
  double pp0, pp1, pp4;       // dPprod/d(ldC,sdC,detP)
  double ox0, ox1, ox5, ox6;  // dOxicMin/d(ldC,sdC,O2,NO3)
  double dn0, dn1, dn5, dn6;  // dDenit
  double an0, an1, an5, an6;  // dAnoxMin
  double ni5, ni7;            // dNitri/d(O2,NH4)
  double od5, od8;            // dOduOx/d(O2,ODU)
  double pa5, pa9;            // dPads/d(O2,PO4)
  double pd8, pd10;           // dPdes/d(ODU,adsP)
  double gc0, gc1;            // dCprod/d(ldC,sdC)
};

__device__ __forceinline__ void kin_jac(const sedcol_params &p, double fT, const double (&c)[NV], KJac &J)
{
  const double ldC = nn(c[0]), sdC = nn(c[1]), ldN = nn(c[2]), sdN = nn(c[3]);
  const double dP = nn(c[4]), O2 = nn(c[5]), NO3 = nn(c[6]), NH4 = nn(c[7]);
  const double ODU = nn(c[8]), PO4 = nn(c[9]), aP = nn(c[10]);
  const double kL = fT * p.rL, kS = fT * p.rS;
  const double dL = kL * ldC, dS = kS * sdC;
  const double den = p.CprodMax + (dL + dS);
  const double lim = p.CprodMax / den;
  const double q = lim / den;
  J.cl0 = kL * (lim - dL * q);
  J.cl1 = -((dL * kS) * q);
  J.cs0 = -((dS * kL) * q);
  J.cs1 = kS * (lim - dS * q);
  J.gc0 = J.cl0 + J.cs0;
  J.gc1 = J.cl1 + J.cs1;
  const double aN = kL * ldN, bN = kS * sdN, aP4 = kL * dP;
  J.nl0 = -((aN * kL) * q);
  J.nl1 = -((aN * kS) * q);
  J.nl2 = kL * lim;
  J.ns0 = -((bN * kL) * q);
  J.ns1 = -((bN * kS) * q);
  J.ns3 = kS * lim;
  J.pp0 = -((aP4 * kL) * q);
  J.pp1 = -((aP4 * kS) * q);
  J.pp4 = kL * lim;
  const double Cprod = dL * lim + dS * lim;
  const double qo = O2 + p.ksO2oxic;
  const double oxl = O2 / qo;
  const double oxl5 = p.ksO2oxic / (qo * qo);
  const double qd = O2 + p.kinO2denit, qn = NO3 + p.ksNO3denit;
  const double wd = p.kinO2denit / qd, mn = NO3 / qn;
  const double dnl = wd * mn;
  const double dnl5 = -(dnl / qd);
  const double dnl6 = wd * (p.ksNO3denit / (qn * qn));
  const double qa = O2 + p.kinO2anox, qb = NO3 + p.kinNO3anox;
  const double anl = (p.kinO2anox / qa) * (p.kinNO3anox / qb);
  const double anl5 = -(anl / qa);
  const double anl6 = -(anl / qb);
  const double rs = 1.0 / ((oxl + dnl) + anl);
  const double S5 = (oxl5 + dnl5) + anl5, S6 = dnl6 + anl6;
  const double shox = oxl * rs, shdn = dnl * rs, shan = anl * rs;
  J.ox0 = J.gc0 * shox;
  J.ox1 = J.gc1 * shox;
  J.ox5 = Cprod * ((oxl5 - shox * S5) * rs);
  J.ox6 = Cprod * ((0.0 - shox * S6) * rs);
  J.dn0 = J.gc0 * shdn;
  J.dn1 = J.gc1 * shdn;
  J.dn5 = Cprod * ((dnl5 - shdn * S5) * rs);
  J.dn6 = Cprod * ((dnl6 - shdn * S6) * rs);
  J.an0 = J.gc0 * shan;
  J.an1 = J.gc1 * shan;
  J.an5 = Cprod * ((anl5 - shan * S5) * rs);
  J.an6 = Cprod * ((anl6 - shan * S6) * rs);
  const double rn = fT * p.rnit, qi = O2 + p.ksO2nitri;
  J.ni7 = rn * (O2 / qi);
  J.ni5 = (rn * NH4) * (p.ksO2nitri / (qi * qi));
  const double ro = fT * p.rODUox, qu = O2 + p.ksO2oduox;
  const double o2u = O2 / qu, o2u5 = p.ksO2oduox / (qu * qu);
  J.od8 = ro * o2u;
  J.od5 = (ro * ODU) * o2u5;
  const double kP = fT * p.rPads, qp = ODU + p.PAdsODU;
  J.pa9 = kP * o2u;
  J.pa5 = (kP * PO4) * o2u5;
  J.pd10 = kP * (ODU / qp);
  J.pd8 = (kP * aP) * (p.PAdsODU / (qp * qp));
}

// o = J_R z
__device__ __forceinline__ void kjac_apply(const KJac &J, const double (&z)[NV], double (&o)[NV])
{
  o[0] = -(J.cl0 * z[0] + J.cl1 * z[1]);
  o[1] = -(J.cs0 * z[0] + J.cs1 * z[1]);
  o[2] = -((J.nl0 * z[0] + J.nl1 * z[1]) + J.nl2 * z[2]);
  o[3] = -((J.ns0 * z[0] + J.ns1 * z[1]) + J.ns3 * z[3]);
  o[4] = -((J.pp0 * z[0] + J.pp1 * z[1]) + J.pp4 * z[4]);
  const double ox = (J.ox0 * z[0] + J.ox1 * z[1]) + (J.ox5 * z[5] + J.ox6 * z[6]);
  const double dn = (J.dn0 * z[0] + J.dn1 * z[1]) + (J.dn5 * z[5] + J.dn6 * z[6]);
  const double an = (J.an0 * z[0] + J.an1 * z[1]) + (J.an5 * z[5] + J.an6 * z[6]);
  const double ni = J.ni5 * z[5] + J.ni7 * z[7];
  const double od = J.od5 * z[5] + J.od8 * z[8];
  const double pa = J.pa5 * z[5] + J.pa9 * z[9];
  const double pd = J.pd8 * z[8] + J.pd10 * z[10];
  o[5] = -((ox + 2.0 * ni) + od);
  o[6] = ni - 0.8 * dn;
  o[7] = (0.0 - (o[2] + o[3])) - ni;
  o[8] = an - od;
  o[9] = ((0.0 - o[4]) - pa) + pd;
  o[10] = pa - pd;
}

__device__ __forceinline__ double kjac_sinkC(const KJac &J, const double (&z)[NV]) { return J.gc0 * z[0] + J.gc1 * z[1]; }
__device__ __forceinline__ double kjac_sinkN(const KJac &J, const double (&z)[NV])
{
  return 0.8 * ((J.dn0 * z[0] + J.dn1 * z[1]) + (J.dn5 * z[5] + J.dn6 * z[6]));
}

// (I - g J_R) w = z by the block structure (detritus C 2x2, detritus N and
// P, O2/NO3 2x2 after eliminating NH4 and ODU, PO4/adsP 2x2); false if a
// pivot is not positive
__device__ __forceinline__ bool kjac_solve(const KJac &J, double g, const double (&z)[NV], double (&w)[NV])
{
  bool ok = true;
  const double a00 = 1.0 + g * J.cl0, a01 = g * J.cl1, a10 = g * J.cs0, a11 = 1.0 + g * J.cs1;
  const double det = a00 * a11 - a01 * a10;
  ok = ok && det > 0.0;
  const double idet = 1.0 / det;
  w[0] = (a11 * z[0] - a01 * z[1]) * idet;
  w[1] = (a00 * z[1] - a10 * z[0]) * idet;
  w[2] = (z[2] - g * (J.nl0 * w[0] + J.nl1 * w[1])) / (1.0 + g * J.nl2);
  w[3] = (z[3] - g * (J.ns0 * w[0] + J.ns1 * w[1])) / (1.0 + g * J.ns3);
  w[4] = (z[4] - g * (J.pp0 * w[0] + J.pp1 * w[1])) / (1.0 + g * J.pp4);
  const double r5 = z[5] - g * (J.ox0 * w[0] + J.ox1 * w[1]);
  const double r6 = z[6] - g * (0.8 * (J.dn0 * w[0] + J.dn1 * w[1]));
  const double r7 = z[7] + g * (((J.nl0 + J.ns0) * w[0] + (J.nl1 + J.ns1) * w[1]) + (J.nl2 * w[2] + J.ns3 * w[3]));
  const double r8 = z[8] + g * (J.an0 * w[0] + J.an1 * w[1]);
  const double A55 = 1.0 + g * ((J.ox5 + 2.0 * J.ni5) + J.od5), A56 = g * J.ox6;
  const double A57 = g * (2.0 * J.ni7), A58 = g * J.od8;
  const double A65 = 0.0 - g * (J.ni5 - 0.8 * J.dn5), A66 = 1.0 + g * (0.8 * J.dn6), A67 = 0.0 - g * J.ni7;
  const double A75 = g * J.ni5, A77 = 1.0 + g * J.ni7;
  const double A85 = 0.0 - g * (J.an5 - J.od5), A86 = 0.0 - g * J.an6, A88 = 1.0 + g * J.od8;
  const double i77 = 1.0 / A77, i88 = 1.0 / A88;
  const double m75 = A75 * i77, r7s = r7 * i77, r8s = r8 * i88;
  const double b55 = (A55 - A57 * m75) - A58 * (A85 * i88);
  const double b56 = A56 - A58 * (A86 * i88);
  const double b65 = A65 - A67 * m75;
  const double rr5 = (r5 - A57 * r7s) - A58 * r8s;
  const double rr6 = r6 - A67 * r7s;
  const double det2 = b55 * A66 - b56 * b65;
  ok = ok && det2 > 0.0;
  const double id2 = 1.0 / det2;
  w[5] = (A66 * rr5 - b56 * rr6) * id2;
  w[6] = (b55 * rr6 - b65 * rr5) * id2;
  w[7] = (r7 - A75 * w[5]) * i77;
  w[8] = ((r8 - A85 * w[5]) - A86 * w[6]) * i88;
  const double r9 = z[9] + g * ((((J.pp0 * w[0] + J.pp1 * w[1]) + J.pp4 * w[4]) - J.pa5 * w[5]) + J.pd8 * w[8]);
  const double r10 = z[10] + g * (J.pa5 * w[5] - J.pd8 * w[8]);
  const double A99 = 1.0 + g * J.pa9, A9a = 0.0 - g * J.pd10, Aa9 = 0.0 - g * J.pa9, Aaa = 1.0 + g * J.pd10;
  const double det3 = A99 * Aaa - A9a * Aa9;
  ok = ok && det3 > 0.0;
  const double id3 = 1.0 / det3;
  w[9] = (Aaa * r9 - A9a * r10) * id3;
  w[10] = (A99 * r10 - Aa9 * r9) * id3;
  return ok;
}

// ---------------------------------------------------------------------------
// lane exchange
// ---------------------------------------------------------------------------

// bound_ctrl set: a lane whose source is out of range or disabled reads 0,
// so no destination register needs clearing first
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v)
{
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

template <int PATTERN>
__device__ __forceinline__ double swz_d(double v)
{
  const int lo = __builtin_amdgcn_ds_swizzle(__double2loint(v), PATTERN);
  const int hi = __builtin_amdgcn_ds_swizzle(__double2hiint(v), PATTERN);
  return __hiloint2double(hi, lo);
}

// value held by the previous lane (wave_shr:1); lane 0 of a group never uses it
__device__ __forceinline__ double lg_up(double v) { return dpp_d<0x138>(v); }
// value held by the next lane (wave_shl:1); the last layer never uses it
__device__ __forceinline__ double lg_dn(double v) { return dpp_d<0x130>(v); }

// Partner value at butterfly level S (S = 1, 2, 4, ...).  Levels 4 and 8 use
// row mirrors, which pair the same sub-groups as lane^S once every sub-group
// of size S holds one value, which the butterfly guarantees because it runs
// the levels in increasing order.
template <int S>
__device__ __forceinline__ double bfly(double v)
{
  if constexpr (S == 1) return dpp_d<0xB1>(v);        // quad_perm [1,0,3,2]
  else if constexpr (S == 2) return dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  else if constexpr (S == 4) return dpp_d<0x141>(v);  // row_half_mirror
  else if constexpr (S == 8) return dpp_d<0x140>(v);  // row_mirror
  else if constexpr (S == 16) return swz_d<0x401F>(v); // swizzle xor 16 within 32 lanes
  else return __shfl_xor(v, 32, 64);
}

template <int LPC, int S = 1>
__device__ __forceinline__ double lg_sum(double v)
{
  if constexpr (S >= LPC) return v;
  else return lg_sum<LPC, 2 * S>(v + bfly<S>(v));
}
// maximum over the group; a NaN in any lane gives NaN
template <int LPC, int S = 1>
__device__ __forceinline__ double lg_max(double v)
{
  if constexpr (S >= LPC) return v;
  else {
    const double o = bfly<S>(v);
    return lg_max<LPC, 2 * S>((o > v || o != o) ? o : v);
  }
}
// (P, D) product-with-derivative pairs: (P1,D1)(P2,D2) = (P1 P2, D1 P2 + P1 D2)
template <int LPC, int S = 1>
__device__ __forceinline__ void lg_pd(double &P, double &D)
{
  if constexpr (S < LPC) {
    const double Po = bfly<S>(P);
    const double Do = bfly<S>(D);
    D = D * Po + P * Do;
    P = P * Po;
    lg_pd<LPC, 2 * S>(P, D);
  }
}

// ---------------------------------------------------------------------------
// per-lane column data
// ---------------------------------------------------------------------------

struct Lane {
  int l;        // layer index (lane in group)
  int K;        // layers of the column
  bool active;  // l < K
  double phi;
  double fT;
  // reference form: geometry and interface diffusivities
  double dz, dzc, Dbio, Ddis;
  // fast form: Dbio/dzc, Ddis/dzc (D/dz on the top layer), 1/(phi dz)
  double Gb, Gd, W;
  // distributed-POM flux cap of the layer (bcup_part = 4; 0 when unset)
  double cap;
  // bioturbation profile 3: exp(100 k_l zc), exp(200 k_l zc), the
  // column's depth to its last layer, (1/86400/1e4)(1 - phiI) and the
  // molecular part of the solute diffusivity
  double ze1, ze2, zcum, zq, Dmol;
};

// parameters of bioturbation profile 3 (Zhang & Wirtz 2017) as the kernels
// see them: on = 0 leaves the set_geometry factor in use
struct ZwDev {
  int on;
  double L1, L2, beta, b, f_labile, f_semilabile;
  const double *e1, *e2, *cum;  // ncell, ncell, ncol
};

// Interface diffusivities of get_rhs (fabm_sediment_driver.F90:652-653 and
// :682-683, division order as written there) and, for the fast form, the
// per-lane transport coefficients derived from them.
template <int ARITH>
__device__ __forceinline__ void diffusivities(const sedcol_params &p, double B, double T, double phiI, double bfac,
                                              Lane &L)
{
  L.fT = temp_factor(T);
  const double Dbio = ((((B * L.fT) / 86400.0) / 10000.0) * (1.0 - phiI)) * bfac;
  const double Dmol = (((p.D0 + T * 0.035) * phiI) / 86400.0) / 10000.0;
  const double Ddis = Dbio + Dmol;
  if (ARITH != ARITH_FAST) {
    L.Dbio = Dbio;
    L.Ddis = Ddis;
    if (ARITH == ARITH_EXT) {
      // profile 3: B = f_T = 1 (fabm_sediment_driver.F90:622-623)
      L.zq = (((1.0 * 1.0) / 86400.0) / 10000.0) * (1.0 - phiI);
      L.Dmol = Dmol;
    }
  } else {
    const double len = L.l == 0 ? L.dz : L.dzc;
    L.Gb = Dbio / len;
    L.Gd = Ddis / len;
    L.W = 1.0 / (L.phi * L.dz);
  }
}

// One get_rhs evaluation for the lane's layer.  sbc[0..NV) bottom-water
// concentrations, sbc[NV..2NV) prescribed top fluxes (LDS, per group).
// For each variable n the callback receives (n, rhs_n, F) where F is the flux
// through the lane's upper interface (lane 0: the top flux intFlux(:,:,1)).
// Transport: diff3d (fabm_sediment_driver.F90:739-825) with the particulate
// bulk/volume-fraction handling of get_rhs :658-678 and dissolved :680-692.
// The fast form requires bcdown != 2.
template <int LPC, int ARITH, class CB>
__device__ __forceinline__ void column_rhs(const sedcol_params &p, const Lane &L, const double *sbc,
                                           const double (&c)[NV], double &sC, double &sN, CB &&cb,
                                           const ZwDev *zw = nullptr)
{
  // Callbacks may stage values in LDS for other lanes of the same wave; the
  // wave barriers keep those accesses ordered around the evaluation.
  __builtin_amdgcn_wave_barrier();
  double R[NV];
  kinetics<ARITH>(p, L.fT, c, R, sC, sN);
  const bool top = L.l == 0;
  const bool bottom = L.l == L.K - 1;
  double Dbio = ARITH != ARITH_FAST ? L.Dbio : 0.0;
  double Ddis = ARITH != ARITH_FAST ? L.Ddis : 0.0;
  if (ARITH == ARITH_EXT && zw && zw->on) {
    // bioturbation profile 3 (fabm_sediment_driver.F90:618-645): the factor
    // follows the infauna biomass implied by the weighted TOC profile of the
    // two POC classes (:380-386), averaged over the column
    const double phi = L.phi;
    double w = 0.0;
    w = w + ((zw->f_labile * phi) / (1.0 - phi)) * c[0];
    w = w + ((zw->f_semilabile * phi) / (1.0 - phi)) * c[1];
    const double avg = lg_sum<LPC>(L.active ? L.dz * w : 0.0) / L.zcum;
    const double bio = ((w * L.ze1) * avg) / (zw->L1 + zw->L2 * L.ze2);
    const double bf = L.active ? (zw->beta * dexp(zw->b * dlog(bio))) / w : 0.0;
    Dbio = L.zq * bf;
    Ddis = Dbio + L.Dmol;
  }
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const bool pt = part(n);
    const double x = pt ? c[n] * L.phi : c[n];
    const double xup = lg_up(x);
    const int bc = pt ? p.bcup_part : p.bcup_dis;
    double F;
    if (ARITH != ARITH_FAST) {
      const double D = pt ? Dbio : Ddis;
      {
        // one flux expression for the top (bcup = 2, against the bottom
        // water over dz) and the interior (against the layer above over
        // dzc), then a select: the same operations as two branches, without
        // the divergent control flow around every variable
        const double xo = top ? sbc[n] : xup;
        const double den = top ? L.dz : L.dzc;
        const double Fd = -((D * (x - xo)) / den);
        const double Ft = (bc == 1 || bc == 4) ? sbc[NV + n] : bc == 2 ? Fd : 0.0;
        F = top ? Ft : Fd;
      }
      if (ARITH == ARITH_EXT && bc == 4) {
        // diff3d BcUp = 4 (fabm_sediment_driver.F90:791-802): the top flux
        // beyond the first layer's cap is added to the interfaces below,
        // each layer's cap taken off in turn, sequentially as written there;
        // every lane of the group runs the same loop on broadcast caps
        double rest = sbc[NV + n] - __shfl(L.cap, 0, LPC);
        bool running = true;  // every check so far passed
#pragma unroll 1
        for (int k = 1; k < LPC; ++k) {
          const double ck = __shfl(L.cap, k, LPC);
          if (k < L.K) {
            running = running && rest > 0.0;
            if (running) {
              if (L.l == k) F = F + rest;
              rest = rest - ck;
            }
          }
        }
        // the loop ran past the last layer: the rest goes to its interface
        if (running && bottom) F = F + rest;
      }
      double Fdn = lg_dn(F);
      if (bottom) Fdn = p.bcdown == 2 ? -((D * (0.0 - x)) / L.dz) : 0.0;
      const double VF = pt ? (1.0 - L.phi) : L.phi;
      const double dC = (F - Fdn) / (VF * L.dz);
      const double Q = pt ? (dC * (1.0 - L.phi)) / L.phi : dC;
      cb(n, R[n] + Q, F);
    } else {
      const double G = pt ? L.Gb : L.Gd;
      {
        const double xo = top ? sbc[n] : xup;
        const double Fd = -(G * (x - xo));
        const double Ft = bc == 1 ? sbc[NV + n] : bc == 2 ? Fd : 0.0;
        F = top ? Ft : Fd;
      }
      // the lane exchange runs on every lane (a lane outside the exec mask
      // reads as 0 to the lane above it), then the bottom value replaces it;
      // written as a select on lg_dn's result it would only be evaluated
      // on the lanes that use it
      double Fdn = lg_dn(F);
      if (bottom) Fdn = 0.0;
      const double Q = (F - Fdn) * L.W;
      cb(n, R[n] + Q, F);
    }
  }
  __builtin_amdgcn_wave_barrier();
}

template <int LPC, int S = 1>
__device__ __forceinline__ double lg_prod(double v)
{
  if constexpr (S >= LPC) return v;
  else return lg_prod<LPC, 2 * S>(v * bfly<S>(v));
}

// Patankar factor of one column (Bruggeman et al. 2007), by Newton from
// p = 0 on the convex decreasing h(p) = S prod (1 + a_j p) - p,
// a_j = h g_j / c_j over the destroyed components (g_j < 0).
//   first stage (STAGE2 = false): S = 1, root in (0, 1];
//   second stage: p = prod (c_j + h g_j p) / c1_j with c1 the first-stage
//   values (den), i.e. S = prod c_j / c1_j (factors with c1_j <= 0 left
//   out); the c1 denominators make the two-stage scheme second order (the
//   weights of MPRK22, Burchard et al. 2003, in single-factor form).
template <int LPC, bool STAGE2>
__device__ __forceinline__ double psolve(const double (&c)[NV], const double (&g)[NV], const double (&den)[NV],
                                         double h, bool active, int64_t &zero_sink, int64_t &newton)
{
  // a single scale S per column keeps the Newton loop at one coefficient per
  // component (a per-component scale costs 22 more live VGPRs)
  double a[NV];
  double Sl = 1.0;
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    double av = 0.0;
    // branch, not select: a wave skips the divisions when none of its lanes
    // destroys component n
    if (active && g[n] < 0.0) {
      if (c[n] > 0.0) {
        av = (h * g[n]) / c[n];
        if (STAGE2 && den[n] > 0.0) Sl = Sl * (c[n] / den[n]);
      } else {
        zero_sink += 1;
      }
    }
    a[n] = av;
  }
  const double Sc = STAGE2 ? lg_prod<LPC>(Sl) : 1.0;
  // quadratic convergence from below (positivity and conservation hold for
  // every iterate); the stopping rule leaves a relative error near 1e-9 at
  // most
  // First step, from p = 0: every factor is 1, so the product is 1 and its
  // derivative the plain sum of the a_j.  The butterfly sum below takes the
  // same operands in the same order as lg_pd would (x * 1 is exact), so the
  // iterate equals the general step's bit for bit.
  double D0 = 0.0;
#pragma unroll
  for (int n = 0; n < NV; ++n) D0 = D0 + a[n];
  D0 = lg_sum<LPC>(D0);
  newton += 1;
  double pp = 0.0 - Sc / (Sc * D0 - 1.0);
  if (!(pp > 0.0)) return 0.0;
  if (pp <= 0x1p-20 * pp) return (!STAGE2 && pp > 1.0) ? 1.0 : pp;
  double dprev = pp;  // the last correction
  for (int it = 1; it < 60; ++it) {
    newton += 1;
    double P = 1.0, D = 0.0;
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      const double f = fma(a[n], pp, 1.0);
      D = fma(D, f, P * a[n]);
      P = P * f;
    }
    lg_pd<LPC>(P, D);
    const double hv = Sc * P - pp;
    const double hp = Sc * D - 1.0;
    const double pn = pp - hv / hp;
    if (!(pn > pp)) break;
    // Stop once the correction is below 2^-20 p, or once it is below
    // 2^-10 p and the next one, dl^3 / dprev^2 by the quadratic
    // convergence, would be below 2^-30 p.
    const double dl = pn - pp;
    const bool conv = dl <= 0x1p-20 * pn || (dl <= 0x1p-10 * pn && (dl * dl) * dl <= (0x1p-30 * pn) * (dprev * dprev));
    dprev = dl;
    pp = pn;
    if (conv) break;
  }
  return (!STAGE2 && pp > 1.0) ? 1.0 : pp;
}

// Patankar factor with geometric-mean weights: p = (S prod_j (1 + a_j p))^(1/N)
// over the N destroyed components of the column, with the a_j and S of
// psolve.  The product form's p falls below 1 by about the sum of the |a_j|,
// which grows with the number of components (layers x variables) and
// throttles long columns; here it falls by about their mean.  The scheme's
// properties stay: one factor scales every rate (conservation), the root
// lies below 1/|a_j| for every j (positivity), and p = 1 + O(h) in the
// first stage and 1 + O(h^2) in the second, as in the product form.
// With q = 1/p = qmin + d, qmin = max(-a_j) and b_j = qmin + a_j >= 0, the
// root solves L = ln S + sum_j ln(b_j + d) = 0.  L is convex and increasing
// in w = ln d, so Newton in w from d0 = S^(-1/N) (L >= 0) decreases d
// monotonically and quadratically to the root, every iterate giving
// p <= p*, also when the root lies next to qmin (a component driven almost
// to zero in one step, d*/qmin far below rounding).  Per lane the factors
// (b_j + d) s, s = 1/(qmin + d0) <= 1, are multiplied with their
// derivative; L and dL/dw = d sum 1/(b_j + d) come from one log and one
// division per lane.
// The b_j live in LDS (bs[j * RS], the lane's slots): they are read once per
// iteration and would otherwise hold 22 VGPRs through the loop.
template <int LPC, bool STAGE2, int RS>
__device__ __forceinline__ double psolve_geo(const double (&c)[NV], const double (&g)[NV], const double (&den)[NV],
                                             double h, bool active, int64_t &zero_sink, int64_t &newton,
                                             double *bs)
{
  double b[NV];
  double Sl = 1.0, nl = 0.0, am = 0.0;
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    double av = 1.0;  // > 0 marks a component that is not destroyed
    if (active && g[n] < 0.0) {
      if (c[n] > 0.0) {
        av = (h * g[n]) / c[n];
        nl = nl + 1.0;
        if (0.0 - av > am) am = 0.0 - av;
        if (STAGE2 && den[n] > 0.0) Sl = Sl * (c[n] / den[n]);
      } else {
        zero_sink += 1;
      }
    }
    b[n] = av;
  }
  const double N = lg_sum<LPC>(nl);
  if (!(N > 0.0)) return 1.0;
  const double lnS = STAGE2 ? lg_sum<LPC>(active ? dlog(Sl) : 0.0) : 0.0;
  const double qmin = lg_max<LPC>(am);
#pragma unroll
  for (int n = 0; n < NV; ++n) bs[n * RS] = b[n] > 0.0 ? -1.0 : qmin + b[n];  // b < 0: not destroyed
  double d = STAGE2 ? dexp((0.0 - lnS) / N) : 1.0;
  const double sc = 1.0 / (qmin + d);
  const double lsc = dlog(sc);
  for (int it = 0; it < 60; ++it) {
    newton += 1;
    // the slots are re-read in every iteration, not held in registers
    __asm__ volatile("" ::: "memory");
    double P = 1.0, D = 0.0;
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      const double bn = bs[n * RS];
      const bool on = bn >= 0.0;
      const double f = on ? (bn + d) * sc : 1.0;
      D = fma(D, f, on ? P * sc : 0.0);
      P = P * f;
    }
    const double Ls = lg_sum<LPC>(active ? dlog(P) : 0.0);
    const double Rs = lg_sum<LPC>(active ? D / P : 0.0);
    const double L = (lnS + Ls) - N * lsc;
    const double Lw = d * Rs;
    const double dn = d * dexp(0.0 - L / Lw);
    if (!(dn > 0.0) || !(dn < 1.0e300)) break;  // NaN, underflow or overflow: keep d
    if (it > 0 && !(dn < d)) break;              // decreasing after the first step
    const double step = d - dn;
    const bool conv = step <= 0x1p-20 * dn || step <= 0x1p-40 * (qmin + dn);
    d = dn;
    if (conv) break;
  }
  const double pg = 1.0 / (qmin + d);
  return (!STAGE2 && pg > 1.0) ? 1.0 : pg;
}

// The Patankar factor of a weighting (sedcol_params.emp_weights), fixed per
// kernel instance: PW = 0 the product over the destroyed components, PW = 1
// their geometric mean.
template <int LPC, bool STAGE2, int PW, int RS>
__device__ __forceinline__ double pfactor(const double (&c)[NV], const double (&g)[NV], const double (&den)[NV],
                                          double h, bool active, int64_t &zero_sink, int64_t &newton, double *bs)
{
  if constexpr (PW == 1) return psolve_geo<LPC, STAGE2, RS>(c, g, den, h, active, zero_sink, newton, bs);
  else return psolve<LPC, STAGE2>(c, g, den, h, active, zero_sink, newton);
}

// ---------------------------------------------------------------------------
// two layers per lane
// ---------------------------------------------------------------------------
//
// Sub-layer s of lane l holds layer 2l + s.  The right-hand side takes the
// neighbour above sub-layer 0 from the previous lane's sub-layer 1 and the
// flux below sub-layer 1 from the next lane's sub-layer 0; the other two
// neighbours are the lane's own.  Layer reductions combine the two
// sub-layers first, which is the first level of the pairwise tree over the
// layers (lower layer first), and then run the lane butterfly: every
// reduction is the tree of the one-layer-per-lane kernels, bit for bit.

// get_rhs for the lane's two layers, reference and fast forms (bioturbation
// profile 3 and BcUp 4 need the one-layer kernels); the callback receives
// (s, n, rhs, F) with F the flux through sub-layer s's upper interface
template <int ARITH, class CB>
__device__ __forceinline__ void column_rhs2(const sedcol_params &p, const Lane (&L)[2], const double *sbc,
                                            const double (&c)[2][NV], double (&sC)[2], double (&sN)[2], CB &&cb)
{
  static_assert(ARITH == ARITH_REF || ARITH == ARITH_FAST, "two layers per lane: reference and fast forms");
  __builtin_amdgcn_wave_barrier();
  double R0[NV], R1[NV];
  kinetics<ARITH>(p, L[0].fT, c[0], R0, sC[0], sN[0]);
  kinetics<ARITH>(p, L[1].fT, c[1], R1, sC[1], sN[1]);
  const bool top = L[0].l == 0;
  const bool bottom0 = L[0].l == L[0].K - 1, bottom1 = L[1].l == L[1].K - 1;
#pragma unroll
  for (int n = 0; n < NV; ++n) {
    const bool pt = part(n);
    const double x0 = pt ? c[0][n] * L[0].phi : c[0][n];
    const double x1 = pt ? c[1][n] * L[1].phi : c[1][n];
    const double xup = lg_up(x1);
    const int bc = pt ? p.bcup_part : p.bcup_dis;
    if (ARITH == ARITH_REF) {
      const double D0 = pt ? L[0].Dbio : L[0].Ddis;
      const double D1 = pt ? L[1].Dbio : L[1].Ddis;
      double F0;
      if (top) F0 = (bc == 1 || bc == 4) ? sbc[NV + n] : bc == 2 ? -((D0 * (x0 - sbc[n])) / L[0].dz) : 0.0;
      else F0 = -((D0 * (x0 - xup)) / L[0].dzc);
      const double F1 = -((D1 * (x1 - x0)) / L[1].dzc);
      double Fdn0 = F1;
      if (bottom0) Fdn0 = p.bcdown == 2 ? -((D0 * (0.0 - x0)) / L[0].dz) : 0.0;
      double Fdn1 = lg_dn(F0);
      if (bottom1) Fdn1 = p.bcdown == 2 ? -((D1 * (0.0 - x1)) / L[1].dz) : 0.0;
      const double VF0 = pt ? (1.0 - L[0].phi) : L[0].phi;
      const double VF1 = pt ? (1.0 - L[1].phi) : L[1].phi;
      const double dC0 = (F0 - Fdn0) / (VF0 * L[0].dz);
      const double dC1 = (F1 - Fdn1) / (VF1 * L[1].dz);
      const double Q0 = pt ? (dC0 * (1.0 - L[0].phi)) / L[0].phi : dC0;
      const double Q1 = pt ? (dC1 * (1.0 - L[1].phi)) / L[1].phi : dC1;
      cb(0, n, R0[n] + Q0, F0);
      cb(1, n, R1[n] + Q1, F1);
    } else {
      const double G0 = pt ? L[0].Gb : L[0].Gd;
      const double G1 = pt ? L[1].Gb : L[1].Gd;
      double F0;
      if (top) F0 = bc == 1 ? sbc[NV + n] : bc == 2 ? -(G0 * (x0 - sbc[n])) : 0.0;
      else F0 = -(G0 * (x0 - xup));
      const double F1 = -(G1 * (x1 - x0));
      double Fdn0 = F1;
      if (bottom0) Fdn0 = 0.0;
      double Fdn1 = lg_dn(F0);
      if (bottom1) Fdn1 = 0.0;
      cb(0, n, R0[n] + (F0 - Fdn0) * L[0].W, F0);
      cb(1, n, R1[n] + (F1 - Fdn1) * L[1].W, F1);
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// psolve for two layers per lane (product weights), the same operation
// sequence per layer and the same reduction tree
template <int LPC, bool STAGE2>
__device__ __forceinline__ double psolve2(const double (&c)[2][NV], const double (&g)[2][NV],
                                          const double (&den)[2][NV], double h, const bool (&active)[2],
                                          int64_t &zero_sink, int64_t &newton)
{
  double a[2][NV];
  double Sl[2] = {1.0, 1.0};
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int n = 0; n < NV; ++n) {
      double av = 0.0;
      if (active[s] && g[s][n] < 0.0) {
        if (c[s][n] > 0.0) {
          av = (h * g[s][n]) / c[s][n];
          if (STAGE2 && den[s][n] > 0.0) Sl[s] = Sl[s] * (c[s][n] / den[s][n]);
        } else {
          zero_sink += 1;
        }
      }
      a[s][n] = av;
    }
  const double Sc = STAGE2 ? lg_prod<LPC>(Sl[0] * Sl[1]) : 1.0;
  double Ds[2] = {0.0, 0.0};
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int n = 0; n < NV; ++n) Ds[s] = Ds[s] + a[s][n];
  const double D0 = lg_sum<LPC>(Ds[0] + Ds[1]);
  newton += 1;
  double pp = 0.0 - Sc / (Sc * D0 - 1.0);
  if (!(pp > 0.0)) return 0.0;
  if (pp <= 0x1p-20 * pp) return (!STAGE2 && pp > 1.0) ? 1.0 : pp;
  double dprev = pp;
  for (int it = 1; it < 60; ++it) {
    newton += 1;
    double Ps[2] = {1.0, 1.0}, Dq[2] = {0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int n = 0; n < NV; ++n) {
        const double f = fma(a[s][n], pp, 1.0);
        Dq[s] = fma(Dq[s], f, Ps[s] * a[s][n]);
        Ps[s] = Ps[s] * f;
      }
    double D = Dq[0] * Ps[1] + Ps[0] * Dq[1];
    double P = Ps[0] * Ps[1];
    lg_pd<LPC>(P, D);
    const double hv = Sc * P - pp;
    const double hp = Sc * D - 1.0;
    const double pn = pp - hv / hp;
    if (!(pn > pp)) break;
    const double dl = pn - pp;
    const bool conv = dl <= 0x1p-20 * pn || (dl <= 0x1p-10 * pn && (dl * dl) * dl <= (0x1p-30 * pn) * (dprev * dprev));
    dprev = dl;
    pp = pn;
    if (conv) break;
  }
  return (!STAGE2 && pp > 1.0) ? 1.0 : pp;
}

}  // namespace sc
