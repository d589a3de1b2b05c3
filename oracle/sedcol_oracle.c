/*
 * sedcol_oracle.c -- CPU oracle for the sediment-column hot path.
 *
 * TEST INFRASTRUCTURE ONLY (see sedcol_oracle.h).  Only tests/, the smoke
 * check and the cpu_baseline leg of bench.py load this library, and only as
 * the checker.
 *
 * Floating point: compiled with -ffp-contract=off; every fused multiply-add
 * is an explicit fma() call, so the operation order written here is the one
 * executed.  Reductions over the layers of a column use a balanced pairwise
 * tree over the next power of two of K (identity padding), the order a
 * lane-per-layer butterfly produces.
 */
#include "sedcol_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NV OC_NVAR
#define OC_KMAX 128

/* particulate flag per state variable: detritus pools and adsorbed P are
 * attached to the solid phase (fabm_sediment_driver.F90:658 'particulate') */
static const int oc_part[NV] = {1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 1};
/* element membership for clamp accounting: 0 = C, 1 = N, 2 = P, -1 none */
static const int oc_elem[NV] = {0, 0, 1, 1, 2, -1, 1, 1, -1, 2, 2};

/* ------------------------------------------------------------------------ */
/* elementary functions with a fixed operation sequence                      */
/* ------------------------------------------------------------------------ */

/* exp(x): Cody-Waite reduction x = n ln2 + r with the fdlibm two-part ln2,
 * then the Taylor series of e^r to degree 13 in Horner form (|r| <= 0.347,
 * truncation < 3e-18 relative), scaled by 2^n through the exponent field.
 * Normal-range results only (x in [-708, 709]). */
double oc_exp(double x)
{
  if (x != x) return x;
  if (x > 709.0) return INFINITY;
  if (x < -708.0) return 0.0;
  const double log2e = 1.4426950408889634;
  const double ln2hi = 6.93147180369123816490e-01;
  const double ln2lo = 1.90821492927058770002e-10;
  double n = rint(x * log2e);
  double r = (x - n * ln2hi) - n * ln2lo;
  double s = 1.6059043836821613e-10;          /* 1/13! */
  s = fma(s, r, 2.08767569878681e-09);        /* 1/12! */
  s = fma(s, r, 2.505210838544172e-08);       /* 1/11! */
  s = fma(s, r, 2.755731922398589e-07);       /* 1/10! */
  s = fma(s, r, 2.7557319223985893e-06);      /* 1/9!  */
  s = fma(s, r, 2.48015873015873e-05);        /* 1/8!  */
  s = fma(s, r, 0.0001984126984126984);       /* 1/7!  */
  s = fma(s, r, 0.001388888888888889);        /* 1/6!  */
  s = fma(s, r, 0.008333333333333333);        /* 1/5!  */
  s = fma(s, r, 0.041666666666666664);        /* 1/4!  */
  s = fma(s, r, 0.16666666666666666);         /* 1/3!  */
  s = fma(s, r, 0.5);
  s = fma(s, r, 1.0);
  s = fma(s, r, 1.0);
  int64_t e = (int64_t)n + 1023;
  uint64_t bits = (uint64_t)e << 52;
  double two_n;
  memcpy(&two_n, &bits, sizeof two_n);
  return s * two_n;
}

/* Natural logarithm, the in-house counterpart of oc_exp: x = m 2^e with m in
 * [sqrt(1/2), sqrt(2)], log m = 2 atanh(s) with s = (m-1)/(m+1) summed to
 * s^25 (|s| <= 0.1716), plus e ln2 in two parts.  The kernels use the same
 * operation sequence. */
double oc_log(double x)
{
  if (x != x) return x;
  if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
  if (x == INFINITY) return x;
  uint64_t b;
  memcpy(&b, &x, sizeof b);
  double k = 0.0;
  if (((b >> 52) & 0x7ff) == 0) { /* subnormal */
    x = x * 0x1p54;
    memcpy(&b, &x, sizeof b);
    k = -54.0;
  }
  int64_t e = (int64_t)((b >> 52) & 0x7ff) - 1023;
  uint64_t mb = (b & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
  double m;
  memcpy(&m, &mb, sizeof m);
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

/* 1/sqrt(x) for finite x > 0: exponent-halving initial guess followed by four
 * Newton steps y <- y (1.5 - (x/2) y^2).  Used only by the step-size
 * controller, where determinism matters more than the last bit. */
double oc_rsqrt(double x)
{
  uint64_t bits;
  memcpy(&bits, &x, sizeof bits);
  bits = 0x5FE6EB50C7B537A9ull - (bits >> 1);
  double y;
  memcpy(&y, &bits, sizeof y);
  double h = 0.5 * x;
  for (int i = 0; i < 4; ++i) y = y * (1.5 - (h * y) * y);
  return y;
}

static inline double nn(double c) { return c < 0.0 ? 0.0 : c; }

/* Planted mistakes, for the tests that show the kinetics pins catch them:
 * oc_plant = 0 (the default) is the definition; each other value replaces
 * one term by a plausible wrong variant (see oc_set_plant).  Nothing but
 * those tests sets it. */
static int oc_plant = 0;

/* 0 definition; 1 nitrification O2 term as inhibition ks/(O2+ks); 2 hard
 * production cap min(1, CprodMax/(dL+dS)); 3 one O2 per nitrified N; 4 one
 * NO3 (and one N2-N) per denitrified C; 5 anoxic O2 inhibition with
 * kinO2denit; 6 detritus N decaying without the production cap; 7 P release
 * term PAdsODU/(ODU+PAdsODU); 8 temperature factor referenced to 20 degC;
 * 9 oxic mineralisation half-saturation kinO2anox; 10 ODU oxidation
 * half-saturation ksO2oxic; 11 labile detritus P decaying at the semi-labile
 * rate.  Returns the previous value. */
int oc_set_plant(int which)
{
  const int prev = oc_plant;
  oc_plant = which;
  return prev;
}

/* temperature factor, fabm_sediment_driver.F90:648 (the _ONE_* factor is an
 * exact multiplication by one) */
static inline double oc_fT(double T)
{
  if (oc_plant == 8) return oc_exp(-4500.0 * (1.0 / (T + 273.0) - (1.0 / 293.0)));
  return oc_exp(-4500.0 * (1.0 / (T + 273.0) - (1.0 / 288.0)));
}

/* ------------------------------------------------------------------------ */
/* kinetics (project definition; FABM hzg_omexdia_p is not in the reference) */
/* ------------------------------------------------------------------------ */

/* Local rates R[n] (mmol m-3 pore water s-1) at one cell, following the
 * OMEXDIA scheme of Soetaert et al. (1996): first-order detritus decay with a
 * saturating production cap (CprodMax), partition of carbon production into
 * oxic mineralisation, denitrification and anoxic mineralisation by
 * Monod/inhibition terms normalised to one, nitrification and ODU
 * re-oxidation, and P adsorption onto / release from oxidised iron.
 * Parameter names follow examples/standalone/omexdia_p/fabm_sed.nml:51-77.
 * Every destruction term of a species is proportional to that species, which
 * is what the Patankar steppers require.  sinkC = carbon leaving as DIC,
 * sinkN = nitrogen leaving as N2. */
void oc_kinetics(const oc_params *p, double fT, const double *c, double *R,
                 double *sinkC, double *sinkN)
{
  double ldC = nn(c[OC_LDETC]), sdC = nn(c[OC_SDETC]);
  double ldN = nn(c[OC_LDETN]), sdN = nn(c[OC_SDETN]);
  double dP = nn(c[OC_DETP]), O2 = nn(c[OC_O2]), NO3 = nn(c[OC_NO3]);
  double NH4 = nn(c[OC_NH4]), ODU = nn(c[OC_ODU]), PO4 = nn(c[OC_PO4]);
  double aP = nn(c[OC_ADSP]);

  double kL = fT * p->rL;
  double kS = fT * p->rS;
  double dL = kL * ldC;
  double dS = kS * sdC;
  double lim = p->CprodMax / (p->CprodMax + (dL + dS));
  if (oc_plant == 2) lim = (dL + dS) > p->CprodMax ? p->CprodMax / (dL + dS) : 1.0;
  double CpL = dL * lim;
  double CpS = dS * lim;
  double Cprod = CpL + CpS;
  double NpL = (kL * ldN) * lim;
  double NpS = (kS * sdN) * lim;
  double Pprod = (kL * dP) * lim;
  if (oc_plant == 6) {
    NpL = kL * ldN;
    NpS = kS * sdN;
  }
  if (oc_plant == 11) Pprod = (kS * dP) * lim;

  double oxlim, dnlim, anlim, nitlim, o2odu, odulim;
  if (p->arith == 0) {
    oxlim = O2 / (O2 + p->ksO2oxic);
    dnlim = (p->kinO2denit / (O2 + p->kinO2denit)) * (NO3 / (NO3 + p->ksNO3denit));
    anlim = (p->kinO2anox / (O2 + p->kinO2anox)) * (p->kinNO3anox / (NO3 + p->kinNO3anox));
    nitlim = O2 / (O2 + p->ksO2nitri);
    o2odu = O2 / (O2 + p->ksO2oduox);
    odulim = ODU / (ODU + p->PAdsODU);
    if (oc_plant == 1) nitlim = p->ksO2nitri / (O2 + p->ksO2nitri);
    if (oc_plant == 5) anlim = (p->kinO2denit / (O2 + p->kinO2denit)) * (p->kinNO3anox / (NO3 + p->kinNO3anox));
    if (oc_plant == 7) odulim = p->PAdsODU / (ODU + p->PAdsODU);
    if (oc_plant == 9) oxlim = O2 / (O2 + p->kinO2anox);
    if (oc_plant == 10) o2odu = O2 / (O2 + p->ksO2oxic);
  } else {
    /* fast form: the six Monod denominators share one reciprocal, taken of
     * their product and split back with prefix/suffix products (valid while
     * the product stays in range: concentrations < 1e30, half-saturation
     * constants >= 1e-6) */
    double q2 = O2 + p->ksO2oxic;
    double q3 = (O2 + p->kinO2denit) * (NO3 + p->ksNO3denit);
    double q4 = (O2 + p->kinO2anox) * (NO3 + p->kinNO3anox);
    double q6 = O2 + p->ksO2nitri;
    double q7 = O2 + p->ksO2oduox;
    double q8 = ODU + p->PAdsODU;
    double a23 = q2 * q3, a234 = a23 * q4, a2346 = a234 * q6, a23467 = a2346 * q7;
    double inv = 1.0 / (a23467 * q8);
    double b87 = q8 * q7, b876 = b87 * q6, b8764 = b876 * q4, b87643 = b8764 * q3;
    oxlim = O2 * (inv * b87643);
    dnlim = (p->kinO2denit * NO3) * ((inv * q2) * b8764);
    anlim = (p->kinO2anox * p->kinNO3anox) * ((inv * a23) * b876);
    nitlim = O2 * ((inv * a234) * b87);
    o2odu = O2 * ((inv * a2346) * q8);
    odulim = ODU * (inv * a23467);
  }
  double rescale = 1.0 / ((oxlim + dnlim) + anlim);
  double OxicMin = (Cprod * oxlim) * rescale;
  double Denit = (Cprod * dnlim) * rescale;
  double AnoxMin = (Cprod * anlim) * rescale;

  double Nitri = ((fT * p->rnit) * NH4) * nitlim;
  double OduOx = ((fT * p->rODUox) * ODU) * o2odu;
  double kP = fT * p->rPads;
  double Pads = (kP * PO4) * o2odu;
  double Pdes = (kP * aP) * odulim;

  R[OC_LDETC] = -CpL;
  R[OC_SDETC] = -CpS;
  R[OC_LDETN] = -NpL;
  R[OC_SDETN] = -NpS;
  R[OC_DETP] = -Pprod;
  R[OC_O2] = -((OxicMin + 2.0 * Nitri) + OduOx);
  R[OC_NO3] = Nitri - 0.8 * Denit;
  R[OC_NH4] = (NpL + NpS) - Nitri;
  R[OC_ODU] = AnoxMin - OduOx;
  R[OC_PO4] = (Pprod - Pads) + Pdes;
  R[OC_ADSP] = Pads - Pdes;
  *sinkC = Cprod;
  *sinkN = 0.8 * Denit;
  if (oc_plant == 3) R[OC_O2] = -((OxicMin + Nitri) + OduOx);
  if (oc_plant == 4) {
    R[OC_NO3] = Nitri - Denit;
    *sinkN = Denit;
  }
}

/* ------------------------------------------------------------------------ */
/* column view and column right-hand side                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
  int K;
  const double *dz, *dzc, *phi, *phiI, *bfac, *fcap;
  double B, T;
  double cbw[NV], fup[NV];
  const oc_zw *zw;
} colview;

static void make_view(const oc_domain *d, int c, colview *v)
{
  int o = d->off[c];
  v->K = d->off[c + 1] - o;
  v->dz = d->dz + o;
  v->dzc = d->dzc + o;
  v->phi = d->phi + o;
  v->phiI = d->phiI + o;
  v->bfac = d->bfac + o;
  v->fcap = d->fcap ? d->fcap + o : NULL;
  v->B = d->biot[c];
  v->T = d->temp[c];
  v->zw = d->zw;
  for (int n = 0; n < NV; ++n) {
    v->cbw[n] = d->cbw[(size_t)n * d->ncol + c];
    v->fup[n] = d->fluxup[(size_t)n * d->ncol + c];
  }
}

/* Fast form of the column transport (p->arith == 1): the same fluxes with the
 * divisions by dzc (dz at the top) folded into G = D/dzc and the divergence
 * (F_k - F_k+1)/(VF dz), rescaled by (1-phi)/phi for particulates, written
 * as (F_k - F_k+1) * W with W = 1/(phi dz) for every variable. */
static void col_transport_fast(const oc_params *p, const colview *v,
                               const double *c, long ks, double *rhs, long rs,
                               double *F1, const double *Dbio, const double *Ddis)
{
  const int K = v->K;
  double Gb[OC_KMAX], Gd[OC_KMAX], W[OC_KMAX], x[OC_KMAX], F[OC_KMAX + 1];
  for (int k = 0; k < K; ++k) {
    double len = k == 0 ? v->dz[k] : v->dzc[k];
    Gb[k] = Dbio[k] / len;
    Gd[k] = Ddis[k] / len;
    W[k] = 1.0 / (v->phi[k] * v->dz[k]);
  }
  for (int n = 0; n < NV; ++n) {
    const int part = oc_part[n];
    const double *G = part ? Gb : Gd;
    const double *cn = c + (long)n * ks;
    for (int k = 0; k < K; ++k) x[k] = part ? cn[k] * v->phi[k] : cn[k];
    for (int k = 1; k < K; ++k) F[k] = -(G[k] * (x[k] - x[k - 1]));
    int bc = part ? p->bcup_part : p->bcup_dis;
    if (bc == 1) F[0] = v->fup[n];
    else if (bc == 2) F[0] = -(G[0] * (x[0] - v->cbw[n]));
    else F[0] = 0.0;
    F[K] = 0.0;
    double *rn = rhs + (long)n * rs;
    for (int k = 0; k < K; ++k) rn[k] = (F[k] - F[k + 1]) * W[k];
    F1[n] = F[0];
  }
}

static double tree_sum(double *v, int K);

/* Layer-centre depths zc (sequential sums of dz from the top, as init_grid
 * :161-168), exp(100 k_l zc), exp(200 k_l zc) and the depth to the last
 * layer's top (cumdepth after the light loop of :593-598: the dz of layers
 * 1..K-1), all from the geometry. */
void oc_zw_geometry(const oc_zw *zw, const double *dz, int K, double *e1, double *e2, double *cum)
{
  double zi = 0.0, cd = 0.0;
  for (int k = 0; k < K; ++k) {
    const double zc = zi + dz[k] * 0.5;
    e1[k] = oc_exp((zc * 100.0) * zw->k_l);
    e2[k] = oc_exp((zc * 200.0) * zw->k_l);
    zi = zi + dz[k];
    if (k < K - 1) cd = cd + dz[k];
  }
  *cum = cd;
}

/* bioturbation factor of profile 3 at every layer (fabm_sediment_driver.F90:
 * 624-642): weighted TOC of the two POC classes, its depth average over the
 * column (sum of dz * toc, in the pairwise tree, over cumdepth), biomass and
 * beta biomass^b / toc */
static void zw_bfac(const colview *v, const double *c, long ks, double *bf)
{
  const oc_zw *zw = v->zw;
  const int K = v->K;
  double wt[OC_KMAX], t[OC_KMAX], e1[OC_KMAX], e2[OC_KMAX], cum;
  oc_zw_geometry(zw, v->dz, K, e1, e2, &cum);
  for (int k = 0; k < K; ++k) {
    const double phi = v->phi[k];
    double w = 0.0;
    w = w + ((zw->f_labile * phi) / (1.0 - phi)) * c[(long)OC_LDETC * ks + k];
    w = w + ((zw->f_semilabile * phi) / (1.0 - phi)) * c[(long)OC_SDETC * ks + k];
    wt[k] = w;
    t[k] = v->dz[k] * w;
  }
  const double avg = tree_sum(t, K) / cum;
  for (int k = 0; k < K; ++k) {
    const double bio = ((wt[k] * e1[k]) * avg) / (zw->L1 + zw->L2 * e2[k]);
    bf[k] = (zw->beta * oc_exp(zw->b * oc_log(bio))) / wt[k];
  }
}

/* One get_rhs for one column.  c and rhs are [NV][K] with stride ks.
 * F1[n] = top interface flux (intFlux(:,:,1)); sC/sN per-cell sinks.
 * fabm_sediment_driver.F90:575-717 and diff3d :739-825. */
static void col_rhs(const oc_params *p, const colview *v, const double *c,
                    long ks, double *rhs, long rs, double *F1, double *sC,
                    double *sN)
{
  const int K = v->K;
  double Dbio[OC_KMAX], Ddis[OC_KMAX], x[OC_KMAX], F[OC_KMAX + 1];
  double fT = oc_fT(v->T);
  double bfz[OC_KMAX];
  const double *bfac = v->bfac;
  double B = v->B;
  if (v->zw) {
    /* profile 3 (:618-645): bioturbation from the infauna biomass that the
     * TOC profile implies; B and f_T are 1 */
    zw_bfac(v, c, ks, bfz);
    bfac = bfz;
    B = 1.0;
    fT = 1.0;
  }
  /* :652-653, then :682-683 for dissolved variables */
  for (int k = 0; k < K; ++k) {
    Dbio[k] = ((((B * fT) / 86400.0) / 10000.0) * (1.0 - v->phiI[k])) * bfac[k];
    Ddis[k] = Dbio[k] + (((p->D0 + v->T * 0.035) * v->phiI[k]) / 86400.0) / 10000.0;
  }
  fT = oc_fT(v->T); /* the kinetics keep the temperature factor */
  /* the fast form covers BcUp 1..3 and BcDown 1, 3 */
  const int fast = p->arith == 1 && p->bcdown != 2 && p->bcup_part != 4 && !v->zw;
  if (fast) col_transport_fast(p, v, c, ks, rhs, rs, F1, Dbio, Ddis);
  for (int n = 0; n < NV && !fast; ++n) {
    const int part = oc_part[n];
    const double *D = part ? Dbio : Ddis;
    const double *cn = c + (long)n * ks;
    for (int k = 0; k < K; ++k) x[k] = part ? cn[k] * v->phi[k] : cn[k];
    /* interior interfaces :776-778 */
    for (int k = 1; k < K; ++k) F[k] = -((D[k] * (x[k] - x[k - 1])) / v->dzc[k]);
    /* upper boundary :782-803 */
    int bc = part ? p->bcup_part : p->bcup_dis;
    if (bc == 1) {
      F[0] = v->fup[n];
    } else if (bc == 2) {
      F[0] = -((D[0] * (x[0] - v->cbw[n])) / v->dz[0]);
    } else if (bc == 3) {
      F[0] = 0.0;
    } else { /* 4: capped flux, overflow cascades downward */
      F[0] = v->fup[n];
      int k = 1;
      double rest = F[0] - v->fcap[0];
      while (rest > 0.0 && k < K) {
        F[k] = F[k] + rest;
        rest = rest - v->fcap[k];
        k = k + 1;
      }
      if (k >= K) F[K - 1] = F[K - 1] + rest;
    }
    /* lower boundary :806-815 (fluxdown = Cdown = 0 in get_rhs) */
    if (p->bcdown == 1) F[K] = 0.0;
    else if (p->bcdown == 2) F[K] = -((D[K - 1] * (0.0 - x[K - 1])) / v->dz[K - 1]);
    else F[K] = 0.0;
    /* divergence :818-820 and particulate rescale :677-678 */
    double *rn = rhs + (long)n * rs;
    for (int k = 0; k < K; ++k) {
      double VF = part ? (1.0 - v->phi[k]) : v->phi[k];
      double dC = (F[k] - F[k + 1]) / (VF * v->dz[k]);
      rn[k] = part ? (dC * (1.0 - v->phi[k])) / v->phi[k] : dC;
    }
    F1[n] = F[0];
  }
  /* local rates, then rhs = R + transport (:697-715); where the fast form
   * does not apply the whole evaluation is in the reference form */
  oc_params pk = *p;
  pk.arith = fast ? p->arith : 0;
  for (int k = 0; k < K; ++k) {
    double cc[NV], R[NV], a, b;
    for (int n = 0; n < NV; ++n) cc[n] = c[(long)n * ks + k];
    oc_kinetics(&pk, fT, cc, R, &a, &b);
    for (int n = 0; n < NV; ++n) rhs[(long)n * rs + k] = R[n] + rhs[(long)n * rs + k];
    sC[k] = a;
    sN[k] = b;
  }
}

/* ------------------------------------------------------------------------ */
/* layer reductions (balanced pairwise tree, identity padding)              */
/* ------------------------------------------------------------------------ */

static int pow2ceil(int K)
{
  int q = 1;
  while (q < K) q <<= 1;
  return q;
}

static double tree_sum(double *v, int K)
{
  int q = pow2ceil(K);
  for (int k = K; k < q; ++k) v[k] = 0.0;
  for (int s = 1; s < q; s <<= 1)
    for (int l = 0; l < q; l += 2 * s) v[l] = v[l] + v[l + s];
  return v[0];
}

static double tree_prod(double *v, int K)
{
  int q = pow2ceil(K);
  for (int k = K; k < q; ++k) v[k] = 1.0;
  for (int s = 1; s < q; s <<= 1)
    for (int l = 0; l < q; l += 2 * s) v[l] = v[l] * v[l + s];
  return v[0];
}

static void tree_pd(double *P, double *D, int K)
{
  int q = pow2ceil(K);
  for (int k = K; k < q; ++k) { P[k] = 1.0; D[k] = 0.0; }
  for (int s = 1; s < q; s <<= 1)
    for (int l = 0; l < q; l += 2 * s) {
      double p1 = P[l], d1 = D[l], p2 = P[l + s], d2 = D[l + s];
      P[l] = p1 * p2;
      D[l] = d1 * p2 + p1 * d2;
    }
}

/* ------------------------------------------------------------------------ */
/* Patankar factor                                                           */
/* ------------------------------------------------------------------------ */

/* Root p in (0,1] of p = prod_{j: g_j<0} (1 + dt g_j p / c_j) over all
 * components of one column (Bruggeman et al. 2007, eq. for EMP).  h(p) =
 * prod - p is convex and decreasing on the admissible interval, so Newton
 * from p = 0 increases monotonically to the root; the product and its
 * derivative are carried together as (P, D) pairs. */
/* Second stage (den != NULL): p = prod_{j: g_j<0} (c_j + dt g_j p) / den_j
 * with den = the first-stage values c1, i.e. p = S prod (1 + a_j p) with the
 * constant S = prod c_j/den_j (a factor is left out when den_j <= 0).  Using
 * c1 rather than c in the denominators is what makes the two-stage scheme
 * second order (the weights c_j^{n+1}/c_j^{(1)} of MPRK22, Burchard et al.
 * 2003, in the single-factor form of Bruggeman et al. 2007).  The same
 * Newton iteration applies; S > 1 allows p > 1. */
static double col_psolve2(const double *c, long ks, const double *g, long gs,
                          const double *den, long dens, double dt, int K,
                          int64_t *zero_sink, int64_t *newton)
{
  double a[NV][OC_KMAX], S[OC_KMAX];
  for (int k = 0; k < K; ++k) S[k] = 1.0;
  for (int n = 0; n < NV; ++n)
    for (int k = 0; k < K; ++k) {
      double gv = g[(long)n * gs + k], cv = c[(long)n * ks + k];
      double av = 0.0;
      if (gv < 0.0) {
        if (cv > 0.0) {
          av = (dt * gv) / cv;
          if (den) {
            double dv = den[(long)n * dens + k];
            if (dv > 0.0) S[k] = S[k] * (cv / dv);
          }
        } else {
          ++*zero_sink;
        }
      }
      a[n][k] = av;
    }
  double Sc = 1.0;
  if (den) Sc = tree_prod(S, K);
  double pp = 0.0;
  double P[OC_KMAX], D[OC_KMAX];
  /* Newton converges quadratically from below (p below the root keeps
   * every component positive, and c + dt p f conserves for any p); the
   * stopping rule below leaves a relative error near 1e-9 at most. */
  double dprev = INFINITY; /* the last correction */
  for (int it = 0; it < 60; ++it) {
    ++*newton;
    for (int k = 0; k < K; ++k) {
      double Pl = 1.0, Dl = 0.0;
      for (int n = 0; n < NV; ++n) {
        double av = a[n][k];
        double f = fma(av, pp, 1.0);
        Dl = fma(Dl, f, Pl * av);
        Pl = Pl * f;
      }
      P[k] = Pl;
      D[k] = Dl;
    }
    tree_pd(P, D, K);
    double h = Sc * P[0] - pp;
    double hp = Sc * D[0] - 1.0;
    double pn = pp - h / hp;
    if (!(pn > pp)) break;
    /* stop once the correction is below 2^-20 p, or once it is below
     * 2^-10 p and the next one, dl^3 / dprev^2 by the quadratic convergence
     * (e_{k+1} = C e_k^2 with C estimated from the last two corrections),
     * would be below 2^-30 p */
    double dl = pn - pp;
    int conv = dl <= 0x1p-20 * pn || (dl <= 0x1p-10 * pn && (dl * dl) * dl <= (0x1p-30 * pn) * (dprev * dprev));
    dprev = dl;
    pp = pn;
    if (conv) break;
  }
  if (!den && pp > 1.0) pp = 1.0;
  return pp;
}

/* Geometric-mean weights (sedcol_params.emp_weights = 1): the root p of
 * p = (S prod_{j: g_j<0} (1 + a_j p))^(1/N) over the N destroyed components,
 * with the a_j and S of col_psolve2.  One factor for all rates keeps mass
 * conservation, the root below every 1/|a_j| keeps positivity, and p - 1 is
 * O(dt) in the first stage and O(dt^2) in the second, as for the product;
 * but p falls below 1 by about the mean of the |a_j| rather than their sum,
 * so long columns are not throttled by their number of components.
 * q = 1/p = qmin + d with qmin = max(-a_j), b_j = qmin + a_j >= 0: the root
 * solves L = ln S + sum_j ln(b_j + d) = 0, convex and increasing in ln d, so
 * Newton in ln d from d0 = S^(-1/N) (L >= 0) decreases d monotonically and
 * quadratically to the root with p <= p* throughout, also when the root lies
 * next to qmin.  Per cell the scaled factors (b_j + d)/(qmin + d0) and their
 * derivative; one log and one division per cell and iteration. */
static double col_psolve_geo(const double *c, long ks, const double *g, long gs,
                             const double *den, long dens, double dt, int K,
                             int64_t *zero_sink, int64_t *newton)
{
  double b[NV][OC_KMAX], Sl[OC_KMAX], nl[OC_KMAX], am[OC_KMAX], t[OC_KMAX];
  for (int k = 0; k < K; ++k) { Sl[k] = 1.0; nl[k] = 0.0; am[k] = 0.0; }
  for (int n = 0; n < NV; ++n)
    for (int k = 0; k < K; ++k) {
      double gv = g[(long)n * gs + k], cv = c[(long)n * ks + k];
      double av = 1.0; /* > 0: not destroyed */
      if (gv < 0.0) {
        if (cv > 0.0) {
          av = (dt * gv) / cv;
          nl[k] = nl[k] + 1.0;
          if (0.0 - av > am[k]) am[k] = 0.0 - av;
          if (den) {
            double dv = den[(long)n * dens + k];
            if (dv > 0.0) Sl[k] = Sl[k] * (cv / dv);
          }
        } else {
          ++*zero_sink;
        }
      }
      b[n][k] = av;
    }
  const double N = tree_sum(nl, K);
  if (!(N > 0.0)) return 1.0;
  double lnS = 0.0;
  if (den) {
    for (int k = 0; k < K; ++k) t[k] = oc_log(Sl[k]);
    lnS = tree_sum(t, K);
  }
  double qmin = 0.0;
  for (int k = 0; k < K; ++k)
    if (am[k] > qmin) qmin = am[k];
  for (int n = 0; n < NV; ++n)
    for (int k = 0; k < K; ++k) b[n][k] = b[n][k] > 0.0 ? -1.0 : qmin + b[n][k];
  double d = den ? oc_exp((0.0 - lnS) / N) : 1.0;
  const double sc = 1.0 / (qmin + d);
  const double lsc = oc_log(sc);
  double lp[OC_KMAX], r[OC_KMAX];
  for (int it = 0; it < 60; ++it) {
    ++*newton;
    for (int k = 0; k < K; ++k) {
      double P = 1.0, D = 0.0;
      for (int n = 0; n < NV; ++n) {
        const int on = b[n][k] >= 0.0;
        const double f = on ? (b[n][k] + d) * sc : 1.0;
        D = fma(D, f, on ? P * sc : 0.0);
        P = P * f;
      }
      lp[k] = oc_log(P);
      r[k] = D / P;
    }
    const double Ls = tree_sum(lp, K);
    const double Rs = tree_sum(r, K);
    const double L = (lnS + Ls) - N * lsc;
    const double Lw = d * Rs;
    const double dn = d * oc_exp(0.0 - L / Lw);
    if (!(dn > 0.0) || !(dn < 1.0e300)) break;
    if (it > 0 && !(dn < d)) break;
    const double step = d - dn;
    const int conv = step <= 0x1p-20 * dn || step <= 0x1p-40 * (qmin + dn);
    d = dn;
    if (conv) break;
  }
  const double pg = 1.0 / (qmin + d);
  return (!den && pg > 1.0) ? 1.0 : pg;
}

static double col_pfactor(int weights, const double *c, long ks, const double *g, long gs,
                          const double *den, long dens, double dt, int K,
                          int64_t *zero_sink, int64_t *newton)
{
  if (weights == 1) return col_psolve_geo(c, ks, g, gs, den, dens, dt, K, zero_sink, newton);
  return col_psolve2(c, ks, g, gs, den, dens, dt, K, zero_sink, newton);
}

/* ------------------------------------------------------------------------ */
/* accumulators                                                              */
/* cellacc[a*ncell + cell]: a = 0 sinkC, 1 sinkN, 2 clampC, 3 clampN, 4 clampP */
/* colacc[n*ncol + c]: time-integrated top flux per variable                  */
/* Both are running sums with Neumaier compensation: the compensation of     */
/* cellacc[a*ncell + cell] is cellacc[(5+a)*ncell + cell], that of            */
/* colacc[n*ncol + c] is colacc[(NV+n)*ncol + c]; a total is sum + comp.      */

static void kadd(double *sum, double *comp, double x)
{
  const double t = *sum + x;
  *comp += fabs(*sum) >= fabs(x) ? (*sum - t) + x : (x - t) + *sum;
  *sum = t;
}
#define ACC_CELL(acc, ncell, a, k, x) kadd(&(acc)[(long)(a) * (ncell) + (k)], &(acc)[(long)(5 + (a)) * (ncell) + (k)], (x))
#define ACC_COL(acc, ncol, n, x) kadd(&(acc)[(long)(n) * (ncol)], &(acc)[(long)(NV + (n)) * (ncol)], (x))
/* ------------------------------------------------------------------------ */

static void clamp_col(const colview *v, double *c, long ks, double *cellacc,
                      long ncell)
{
  for (int n = 0; n < NV; ++n)
    for (int k = 0; k < v->K; ++k) {
      double *x = c + (long)n * ks + k;
      if (*x < 0.0) {
        double add = 0.0 - *x;
        if (cellacc && oc_elem[n] >= 0) ACC_CELL(cellacc, ncell, 2 + oc_elem[n], k, add);
        *x = 0.0;
      }
    }
}

/* ------------------------------------------------------------------------ */
/* fixed-step methods for one column (0, 1, 3, 4, 5)                         */
/* ------------------------------------------------------------------------ */

typedef struct {
  double r[4][NV * OC_KMAX];
  double s[4][2][OC_KMAX];
  double F[4][NV];
  double c1[NV * OC_KMAX];
} colwork;

/* One ode_solver step of one column.  c: [NV][K] stride ks; cellacc points
 * at this column's first cell (stride ncell); colacc at this column (stride
 * ncol).  Returns 0, or 1 if the new state has a NaN. */
static int col_step(const oc_params *p, const colview *v, int method,
                    double dt, double *c, long ks, double *F1out,
                    double *cellacc, long ncell, double *colacc, long ncol,
                    oc_coldiag *dg, colwork *w)
{
  const int K = v->K;
  const long S = OC_KMAX; /* stride of the work arrays */
  const double third = 1.0 / 3.0;
  double *c1 = w->c1;
  double sc = 0.0, sn = 0.0; /* unused placeholders */
  (void)sc; (void)sn;
  if (method == OC_EULER) {
    col_rhs(p, v, c, ks, w->r[0], S, w->F[0], w->s[0][0], w->s[0][1]);
    for (int n = 0; n < NV; ++n)
      for (int k = 0; k < K; ++k) c[n * ks + k] = c[n * ks + k] + dt * w->r[0][n * S + k];
    if (colacc) for (int n = 0; n < NV; ++n) ACC_COL(colacc, ncol, n, dt * w->F[0][n]);
    if (cellacc)
      for (int k = 0; k < K; ++k) {
        ACC_CELL(cellacc, ncell, 0, k, dt * w->s[0][0][k]);
        ACC_CELL(cellacc, ncell, 1, k, dt * w->s[0][1][k]);
      }
    memcpy(F1out, w->F[0], sizeof(double) * NV);
  } else if (method == OC_RK4 || method == OC_RK4_38) {
    const int rk38 = method == OC_RK4_38;
    /* solver_library.F90:142-163 (RK4), :164-185 (3/8 rule) */
    col_rhs(p, v, c, ks, w->r[0], S, w->F[0], w->s[0][0], w->s[0][1]);
    for (int n = 0; n < NV; ++n)
      for (int k = 0; k < K; ++k) {
        double k1 = w->r[0][n * S + k];
        c1[n * S + k] = rk38 ? c[n * ks + k] + (third * dt) * k1
                             : c[n * ks + k] + (0.5 * dt) * k1;
      }
    col_rhs(p, v, c1, S, w->r[1], S, w->F[1], w->s[1][0], w->s[1][1]);
    for (int n = 0; n < NV; ++n)
      for (int k = 0; k < K; ++k) {
        double k1 = w->r[0][n * S + k], k2 = w->r[1][n * S + k];
        c1[n * S + k] = rk38 ? c[n * ks + k] + dt * (k2 - third * k1)
                             : c[n * ks + k] + (0.5 * dt) * k2;
      }
    col_rhs(p, v, c1, S, w->r[2], S, w->F[2], w->s[2][0], w->s[2][1]);
    for (int n = 0; n < NV; ++n)
      for (int k = 0; k < K; ++k) {
        double k1 = w->r[0][n * S + k], k2 = w->r[1][n * S + k], k3 = w->r[2][n * S + k];
        c1[n * S + k] = rk38 ? c[n * ks + k] + dt * ((k1 - k2) + k3)
                             : c[n * ks + k] + dt * k3;
      }
    col_rhs(p, v, c1, S, w->r[3], S, w->F[3], w->s[3][0], w->s[3][1]);
    const double wdt = rk38 ? (dt * 1.0) / 8.0 : dt * third;
#define COMB(a0, a1, a2, a3) \
  (rk38 ? (((a0) + 3.0 * (a1)) + 3.0 * (a2)) + (a3) : (((0.5 * (a0) + (a1)) + (a2)) + 0.5 * (a3)))
    for (int n = 0; n < NV; ++n)
      for (int k = 0; k < K; ++k) {
        long i = n * S + k;
        c[n * ks + k] = c[n * ks + k] + wdt * COMB(w->r[0][i], w->r[1][i], w->r[2][i], w->r[3][i]);
      }
    if (colacc)
      for (int n = 0; n < NV; ++n)
        ACC_COL(colacc, ncol, n, wdt * COMB(w->F[0][n], w->F[1][n], w->F[2][n], w->F[3][n]));
    if (cellacc)
      for (int a = 0; a < 2; ++a)
        for (int k = 0; k < K; ++k)
          ACC_CELL(cellacc, ncell, a, k, wdt * COMB(w->s[0][a][k], w->s[1][a][k], w->s[2][a][k], w->s[3][a][k]));
#undef COMB
    memcpy(F1out, w->F[3], sizeof(double) * NV);
  } else { /* EMP-1 / EMP-2, fixed step */
    col_rhs(p, v, c, ks, w->r[0], S, w->F[0], w->s[0][0], w->s[0][1]);
    double p1 = col_pfactor(p->emp_weights, c, ks, w->r[0], S, NULL, 0, dt, K, &dg->zero_sink, &dg->newton);
    if (p1 < dg->p_min) dg->p_min = p1;
    if (method == OC_EMP1) {
      for (int n = 0; n < NV; ++n)
        for (int k = 0; k < K; ++k) c[n * ks + k] = c[n * ks + k] + (dt * p1) * w->r[0][n * S + k];
      if (colacc) for (int n = 0; n < NV; ++n) ACC_COL(colacc, ncol, n, (dt * p1) * w->F[0][n]);
      if (cellacc)
        for (int k = 0; k < K; ++k) {
          ACC_CELL(cellacc, ncell, 0, k, (dt * p1) * w->s[0][0][k]);
          ACC_CELL(cellacc, ncell, 1, k, (dt * p1) * w->s[0][1][k]);
        }
      memcpy(F1out, w->F[0], sizeof(double) * NV);
    } else {
      for (int n = 0; n < NV; ++n)
        for (int k = 0; k < K; ++k) {
          double y = c[n * ks + k] + (dt * p1) * w->r[0][n * S + k];
          c1[n * S + k] = y < 0.0 ? 0.0 : y;
        }
      col_rhs(p, v, c1, S, w->r[1], S, w->F[1], w->s[1][0], w->s[1][1]);
      for (int n = 0; n < NV; ++n)
        for (int k = 0; k < K; ++k)
          w->r[2][n * S + k] = (w->r[0][n * S + k] + w->r[1][n * S + k]) * 0.5;
      double p2 = col_pfactor(p->emp_weights, c, ks, w->r[2], S, c1, S, dt, K, &dg->zero_sink, &dg->newton);
      if (p2 < dg->p_min) dg->p_min = p2;
      for (int n = 0; n < NV; ++n)
        for (int k = 0; k < K; ++k) c[n * ks + k] = c[n * ks + k] + (dt * p2) * w->r[2][n * S + k];
      if (colacc)
        for (int n = 0; n < NV; ++n) ACC_COL(colacc, ncol, n, (dt * p2) * ((w->F[0][n] + w->F[1][n]) * 0.5));
      if (cellacc)
        for (int a = 0; a < 2; ++a)
          for (int k = 0; k < K; ++k)
            ACC_CELL(cellacc, ncell, a, k, (dt * p2) * ((w->s[0][a][k] + w->s[1][a][k]) * 0.5));
      memcpy(F1out, w->F[1], sizeof(double) * NV);
    }
    /* Patankar keeps components non-negative up to rounding */
    clamp_col(v, c, ks, cellacc, ncell);
  }
  dg->accepted += 1;
  dg->dt_last = dt;
  if (dt < dg->dt_minacc) dg->dt_minacc = dt;
  for (int n = 0; n < NV; ++n)
    for (int k = 0; k < K; ++k)
      if (c[n * ks + k] != c[n * ks + k]) return 1;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* adaptive EMP for one column                                               */
/* ------------------------------------------------------------------------ */

/* Integrates one column over t_span with the embedded EMP-1/EMP-2 pair.
 * err = max_j |c2_j - c1_j| / (atol + rtol max(c2_j, c_j)); accept when
 * err <= 1 or the step is already at dt_min; next step dt * clamp(0.9
 * err^-1/2, 0.2, 5) within [dt_min, dt_max]; a step clipped to the end of the
 * span keeps the larger of the old and the proposed size. */
static int col_emp_adaptive(const oc_params *p, const colview *v, double *c,
                            long ks, double t_span, double *F1out,
                            double *cellacc, long ncell, double *colacc,
                            long ncol, oc_coldiag *dg, int64_t max_attempts,
                            colwork *w)
{
  const int K = v->K;
  const long S = OC_KMAX;
  double *c1 = w->c1;
  double *g = w->r[2];
  double *cn = w->r[3];
  double rem = t_span;
  double dt = dg->dt_next;
  if (!(dt > 0.0)) dt = p->emp_dt_max;
  int64_t attempts = 0;
  while (rem > 0.0) {
    if (attempts >= max_attempts) { dg->status = 2; break; }
    ++attempts;
    int clipped = !(dt < rem);
    double h = clipped ? rem : dt;
    col_rhs(p, v, c, ks, w->r[0], S, w->F[0], w->s[0][0], w->s[0][1]);
    double p1 = col_pfactor(p->emp_weights, c, ks, w->r[0], S, NULL, 0, h, K, &dg->zero_sink, &dg->newton);
    for (int n = 0; n < NV; ++n)
      for (int k = 0; k < K; ++k) {
        double y = c[n * ks + k] + (h * p1) * w->r[0][n * S + k];
        c1[n * S + k] = y < 0.0 ? 0.0 : y;
      }
    col_rhs(p, v, c1, S, w->r[1], S, w->F[1], w->s[1][0], w->s[1][1]);
    for (int n = 0; n < NV; ++n)
      for (int k = 0; k < K; ++k)
        g[n * S + k] = (w->r[0][n * S + k] + w->r[1][n * S + k]) * 0.5;
    double p2 = col_pfactor(p->emp_weights, c, ks, g, S, c1, S, h, K, &dg->zero_sink, &dg->newton);
    /* err = max_j |c2_j - c1_j| / sc_j: per layer the largest ratio is found
     * by cross-multiplication over the variables (in order) and divided
     * once, then the maximum is taken over the layers */
    double err = 0.0;
    int bad = 0;
    for (int k = 0; k < K; ++k) {
      double bn = 0.0, bd = 1.0;
      for (int n = 0; n < NV; ++n) {
        double y = c[n * ks + k] + (h * p2) * g[n * S + k];
        cn[n * S + k] = y;
        double yp = y < 0.0 ? 0.0 : y;
        double c0 = c[n * ks + k];
        double sc = p->emp_atol + p->emp_rtol * (yp > c0 ? yp : c0);
        double dv = fabs(yp - c1[n * S + k]);
        if (dv != dv || sc != sc) bad = 1;
        if (dv * bd > bn * sc) { bn = dv; bd = sc; }
      }
      double e = bn / bd;
      if (e != e) bad = 1;
      if (e > err) err = e;
    }
    if (bad) { dg->status = 1; break; }
    int accept = (err <= 1.0) || !(h > p->emp_dt_min);
    double fac = err > 0.0 ? 0.9 * oc_rsqrt(err) : 5.0;
    if (fac < 0.2) fac = 0.2;
    if (fac > 5.0) fac = 5.0;
    if (p1 < dg->p_min) dg->p_min = p1;
    if (p2 < dg->p_min) dg->p_min = p2;
    if (accept) {
      for (int n = 0; n < NV; ++n) {
        for (int k = 0; k < K; ++k) {
          double y = cn[n * S + k];
          if (y < 0.0) {
            if (cellacc && oc_elem[n] >= 0) ACC_CELL(cellacc, ncell, 2 + oc_elem[n], k, 0.0 - y);
            y = 0.0;
          }
          c[n * ks + k] = y;
        }
      }
      if (colacc)
        for (int n = 0; n < NV; ++n) ACC_COL(colacc, ncol, n, (h * p2) * ((w->F[0][n] + w->F[1][n]) * 0.5));
      if (cellacc)
        for (int a = 0; a < 2; ++a)
          for (int k = 0; k < K; ++k)
            ACC_CELL(cellacc, ncell, a, k, (h * p2) * ((w->s[0][a][k] + w->s[1][a][k]) * 0.5));
      memcpy(F1out, w->F[1], sizeof(double) * NV);
      rem = rem - h;
      dg->accepted += 1;
      if (err > 1.0) dg->forced += 1;
      dg->dt_last = h;
      if (h < dg->dt_minacc) dg->dt_minacc = h;
      double prop = h * fac;
      dt = (clipped && dt > prop) ? dt : prop;
    } else {
      dg->rejected += 1;
      dt = h * fac;
    }
    if (dt < p->emp_dt_min) dt = p->emp_dt_min;
    if (dt > p->emp_dt_max) dt = p->emp_dt_max;
  }
  dg->dt_next = dt;
  return dg->status;
}

/* ------------------------------------------------------------------------ */
/* linearly implicit Rosenbrock stepper, method 7 (project definition)       */
/* ------------------------------------------------------------------------ */

/* ROS2 (Verwer, Spee, Blom and Hundsdorfer 1999) as a W-method:
 *   W k1 = f(c),  W k2 = f(c + h k1) - 2 k1,  c+ = c + h (3/2 k1 + 1/2 k2),
 * gamma = 1 + 1/sqrt(2), with the first-order c + h k1 as the embedded
 * solution (error estimate h (k1 + k2)/2).  W approximates I - gamma h J by
 * the product of the per-species tridiagonal transport factor
 * B = I - gamma h J_T and the per-cell reaction factor A = I - gamma h J_R,
 * plus one correction with the full I - gamma h (J_T + J_R):
 *   k0 = A^-1 B^-1 r,  k = k0 + A^-1 B^-1 (r - (I - gamma h (J_T + J_R)) k0).
 * J_R is the analytic Jacobian of the local rates (oc_kinetics, reference
 * form, at the clamped state), J_T that of the column transport (with the
 * state dependence of a Dirichlet top flux).  A conserves P, B every
 * variable's column inventory up to its Dirichlet top flux, so the top
 * fluxes and the C and N sinks of a step follow exactly from the stage
 * rates, the linearised top flux and the gradients of Cprod and 0.8 Denit
 * (see col_ros2_adaptive).  Positivity: a step whose new state has a
 * component below -1e-3 atol is rejected; smaller negative values are set
 * to 0 and their mass booked as clamp mass. */
#define ROS_GAMMA 1.7071067811865475

typedef struct {
  double cl0, cl1, cs0, cs1;      /* dCpL/d(ldC,sdC), dCpS/d(ldC,sdC) */
  double nl0, nl1, nl2;           /* dNpL/d(ldC,sdC,ldN) */
  double ns0, ns1, ns3;           /* dNpS/d(ldC,sdC,sdN) */
  double pp0, pp1, pp4;           /* dPprod/d(ldC,sdC,detP) */
  double ox0, ox1, ox5, ox6;      /* dOxicMin/d(ldC,sdC,O2,NO3) */
  double dn0, dn1, dn5, dn6;      /* dDenit/d(...) */
  double an0, an1, an5, an6;      /* dAnoxMin/d(...) */
  double ni5, ni7;                /* dNitri/d(O2,NH4) */
  double od5, od8;                /* dOduOx/d(O2,ODU) */
  double pa5, pa9;                /* dPads/d(O2,PO4) */
  double pd8, pd10;               /* dPdes/d(ODU,adsP) */
  double gc0, gc1;                /* dCprod/d(ldC,sdC) */
} oc_kjac;

/* Jacobian of the local rates at c (evaluated at max(c, 0)) */
void oc_kin_jac(const oc_params *p, double fT, const double *c, oc_kjac *J)
{
  const double ldC = nn(c[0]), sdC = nn(c[1]), ldN = nn(c[2]), sdN = nn(c[3]);
  const double dP = nn(c[4]), O2 = nn(c[5]), NO3 = nn(c[6]), NH4 = nn(c[7]);
  const double ODU = nn(c[8]), PO4 = nn(c[9]), aP = nn(c[10]);
  const double kL = fT * p->rL, kS = fT * p->rS;
  const double dL = kL * ldC, dS = kS * sdC;
  const double den = p->CprodMax + (dL + dS);
  const double lim = p->CprodMax / den;
  const double q = lim / den;
  J->cl0 = kL * (lim - dL * q);
  J->cl1 = -((dL * kS) * q);
  J->cs0 = -((dS * kL) * q);
  J->cs1 = kS * (lim - dS * q);
  J->gc0 = J->cl0 + J->cs0;
  J->gc1 = J->cl1 + J->cs1;
  const double aN = kL * ldN, bN = kS * sdN, aP4 = kL * dP;
  J->nl0 = -((aN * kL) * q);
  J->nl1 = -((aN * kS) * q);
  J->nl2 = kL * lim;
  J->ns0 = -((bN * kL) * q);
  J->ns1 = -((bN * kS) * q);
  J->ns3 = kS * lim;
  J->pp0 = -((aP4 * kL) * q);
  J->pp1 = -((aP4 * kS) * q);
  J->pp4 = kL * lim;
  const double Cprod = dL * lim + dS * lim;
  /* partition weights and their derivatives */
  const double qo = O2 + p->ksO2oxic;
  const double oxl = O2 / qo;
  const double oxl5 = p->ksO2oxic / (qo * qo);
  const double qd = O2 + p->kinO2denit, qn = NO3 + p->ksNO3denit;
  const double wd = p->kinO2denit / qd, mn = NO3 / qn;
  const double dnl = wd * mn;
  const double dnl5 = -(dnl / qd);
  const double dnl6 = wd * (p->ksNO3denit / (qn * qn));
  const double qa = O2 + p->kinO2anox, qb = NO3 + p->kinNO3anox;
  const double anl = (p->kinO2anox / qa) * (p->kinNO3anox / qb);
  const double anl5 = -(anl / qa);
  const double anl6 = -(anl / qb);
  const double rs = 1.0 / ((oxl + dnl) + anl);
  const double S5 = (oxl5 + dnl5) + anl5, S6 = dnl6 + anl6;
  const double shox = oxl * rs, shdn = dnl * rs, shan = anl * rs;
  J->ox0 = J->gc0 * shox;
  J->ox1 = J->gc1 * shox;
  J->ox5 = Cprod * ((oxl5 - shox * S5) * rs);
  J->ox6 = Cprod * ((0.0 - shox * S6) * rs);
  J->dn0 = J->gc0 * shdn;
  J->dn1 = J->gc1 * shdn;
  J->dn5 = Cprod * ((dnl5 - shdn * S5) * rs);
  J->dn6 = Cprod * ((dnl6 - shdn * S6) * rs);
  J->an0 = J->gc0 * shan;
  J->an1 = J->gc1 * shan;
  J->an5 = Cprod * ((anl5 - shan * S5) * rs);
  J->an6 = Cprod * ((anl6 - shan * S6) * rs);
  const double rn = fT * p->rnit, qi = O2 + p->ksO2nitri;
  J->ni7 = rn * (O2 / qi);
  J->ni5 = (rn * NH4) * (p->ksO2nitri / (qi * qi));
  const double ro = fT * p->rODUox, qu = O2 + p->ksO2oduox;
  const double o2u = O2 / qu, o2u5 = p->ksO2oduox / (qu * qu);
  J->od8 = ro * o2u;
  J->od5 = (ro * ODU) * o2u5;
  const double kP = fT * p->rPads, qp = ODU + p->PAdsODU;
  J->pa9 = kP * o2u;
  J->pa5 = (kP * PO4) * o2u5;
  J->pd10 = kP * (ODU / qp);
  J->pd8 = (kP * aP) * (p->PAdsODU / (qp * qp));
}

/* out = J_R z */
static void kjac_apply(const oc_kjac *J, const double *z, double *o)
{
  o[0] = -(J->cl0 * z[0] + J->cl1 * z[1]);
  o[1] = -(J->cs0 * z[0] + J->cs1 * z[1]);
  o[2] = -((J->nl0 * z[0] + J->nl1 * z[1]) + J->nl2 * z[2]);
  o[3] = -((J->ns0 * z[0] + J->ns1 * z[1]) + J->ns3 * z[3]);
  o[4] = -((J->pp0 * z[0] + J->pp1 * z[1]) + J->pp4 * z[4]);
  const double ox = (J->ox0 * z[0] + J->ox1 * z[1]) + (J->ox5 * z[5] + J->ox6 * z[6]);
  const double dn = (J->dn0 * z[0] + J->dn1 * z[1]) + (J->dn5 * z[5] + J->dn6 * z[6]);
  const double an = (J->an0 * z[0] + J->an1 * z[1]) + (J->an5 * z[5] + J->an6 * z[6]);
  const double ni = J->ni5 * z[5] + J->ni7 * z[7];
  const double od = J->od5 * z[5] + J->od8 * z[8];
  const double pa = J->pa5 * z[5] + J->pa9 * z[9];
  const double pd = J->pd8 * z[8] + J->pd10 * z[10];
  o[5] = -((ox + 2.0 * ni) + od);
  o[6] = ni - 0.8 * dn;
  o[7] = (0.0 - (o[2] + o[3])) - ni;
  o[8] = an - od;
  o[9] = ((0.0 - o[4]) - pa) + pd;
  o[10] = pa - pd;
}

/* gradients of the C sink (Cprod) and N sink (0.8 Denit) applied to z */
static inline double kjac_sinkC(const oc_kjac *J, const double *z) { return J->gc0 * z[0] + J->gc1 * z[1]; }
static inline double kjac_sinkN(const oc_kjac *J, const double *z)
{
  return 0.8 * ((J->dn0 * z[0] + J->dn1 * z[1]) + (J->dn5 * z[5] + J->dn6 * z[6]));
}

/* (I - g J_R) w = z by the block structure of J_R: detritus C (2x2), then
 * detritus N and P, then O2/NO3/NH4/ODU (NH4 and ODU eliminated, 2x2 in O2
 * and NO3), then PO4/adsP (2x2).  Returns 1 if a pivot is not positive. */
static int kjac_solve(const oc_kjac *J, double g, const double *z, double *w)
{
  const double a00 = 1.0 + g * J->cl0, a01 = g * J->cl1, a10 = g * J->cs0, a11 = 1.0 + g * J->cs1;
  const double det = a00 * a11 - a01 * a10;
  if (!(det > 0.0)) return 1;
  const double idet = 1.0 / det;
  w[0] = (a11 * z[0] - a01 * z[1]) * idet;
  w[1] = (a00 * z[1] - a10 * z[0]) * idet;
  w[2] = (z[2] - g * (J->nl0 * w[0] + J->nl1 * w[1])) / (1.0 + g * J->nl2);
  w[3] = (z[3] - g * (J->ns0 * w[0] + J->ns1 * w[1])) / (1.0 + g * J->ns3);
  w[4] = (z[4] - g * (J->pp0 * w[0] + J->pp1 * w[1])) / (1.0 + g * J->pp4);
  const double r5 = z[5] - g * (J->ox0 * w[0] + J->ox1 * w[1]);
  const double r6 = z[6] - g * (0.8 * (J->dn0 * w[0] + J->dn1 * w[1]));
  const double r7 = z[7] + g * (((J->nl0 + J->ns0) * w[0] + (J->nl1 + J->ns1) * w[1]) + (J->nl2 * w[2] + J->ns3 * w[3]));
  const double r8 = z[8] + g * (J->an0 * w[0] + J->an1 * w[1]);
  const double A55 = 1.0 + g * ((J->ox5 + 2.0 * J->ni5) + J->od5), A56 = g * J->ox6;
  const double A57 = g * (2.0 * J->ni7), A58 = g * J->od8;
  const double A65 = 0.0 - g * (J->ni5 - 0.8 * J->dn5), A66 = 1.0 + g * (0.8 * J->dn6), A67 = 0.0 - g * J->ni7;
  const double A75 = g * J->ni5, A77 = 1.0 + g * J->ni7;
  const double A85 = 0.0 - g * (J->an5 - J->od5), A86 = 0.0 - g * J->an6, A88 = 1.0 + g * J->od8;
  const double i77 = 1.0 / A77, i88 = 1.0 / A88;
  const double m75 = A75 * i77, r7s = r7 * i77, r8s = r8 * i88;
  const double b55 = (A55 - A57 * m75) - A58 * (A85 * i88);
  const double b56 = A56 - A58 * (A86 * i88);
  const double b65 = A65 - A67 * m75;
  const double rr5 = (r5 - A57 * r7s) - A58 * r8s;
  const double rr6 = r6 - A67 * r7s;
  const double det2 = b55 * A66 - b56 * b65;
  if (!(det2 > 0.0)) return 1;
  const double id2 = 1.0 / det2;
  w[5] = (A66 * rr5 - b56 * rr6) * id2;
  w[6] = (b55 * rr6 - b65 * rr5) * id2;
  w[7] = (r7 - A75 * w[5]) * i77;
  w[8] = ((r8 - A85 * w[5]) - A86 * w[6]) * i88;
  const double r9 = z[9] + g * ((((J->pp0 * w[0] + J->pp1 * w[1]) + J->pp4 * w[4]) - J->pa5 * w[5]) + J->pd8 * w[8]);
  const double r10 = z[10] + g * (J->pa5 * w[5] - J->pd8 * w[8]);
  const double A99 = 1.0 + g * J->pa9, A9a = 0.0 - g * J->pd10, Aa9 = 0.0 - g * J->pa9, Aaa = 1.0 + g * J->pd10;
  const double det3 = A99 * Aaa - A9a * Aa9;
  if (!(det3 > 0.0)) return 1;
  const double id3 = 1.0 / det3;
  w[9] = (Aaa * r9 - A9a * r10) * id3;
  w[10] = (A99 * r10 - Aa9 * r9) * id3;
  return 0;
}

/* For the tests: Jz = J_R z, w = (I - g J_R)^-1 z and gs = (grad Cprod . z,
 * grad 0.8 Denit . z) at state c; returns 1 if the solve fails. */
int oc_kin_jac_ops(const oc_params *p, double fT, const double *c, double g, const double *z, double *Jz,
                   double *w, double *gs)
{
  oc_kjac J;
  oc_kin_jac(p, fT, c, &J);
  kjac_apply(&J, z, Jz);
  gs[0] = kjac_sinkC(&J, z);
  gs[1] = kjac_sinkN(&J, z);
  return kjac_solve(&J, g, z, w);
}

/* per-species tridiagonal factor B = I - g J_T: sub-diagonal -a_k,
 * diagonal b_k, super-diagonal -c_k with
 *   a_k = g W_k G_k q_k-1,  c_k = g W_k G_k+1 q_k+1,
 *   b_k = 1 + g W_k (G_k [k > 0 or Dirichlet top] + G_k+1 [k < K-1]) q_k,
 * q = phi (particulate) or 1; Thomas elimination, inv_k = 1/pivot_k */
static inline double tri_a(double g, const double *G, const double *W, const double *q, int k)
{
  return k > 0 ? g * ((W[k] * G[k]) * q[k - 1]) : 0.0;
}
static inline double tri_c(double g, const double *G, const double *W, const double *q, int k, int K)
{
  return k < K - 1 ? g * ((W[k] * G[k + 1]) * q[k + 1]) : 0.0;
}
static void tri_factor(double g, const double *G, const double *W, const double *q, int K, int dir, double *inv)
{
  double e = 0.0;
  for (int k = 0; k < K; ++k) {
    const double s = ((k > 0 || dir) ? G[k] : 0.0) + (k < K - 1 ? G[k + 1] : 0.0);
    const double b = 1.0 + g * (W[k] * (s * q[k]));
    const double piv = k > 0 ? b - tri_a(g, G, W, q, k) * e : b;
    inv[k] = 1.0 / piv;
    e = tri_c(g, G, W, q, k, K) * inv[k];
  }
}
static void tri_solve(double g, const double *G, const double *W, const double *q, int K, const double *inv,
                      const double *r, double *z)
{
  double d = 0.0;
  for (int k = 0; k < K; ++k) {
    d = k > 0 ? (r[k] + tri_a(g, G, W, q, k) * d) * inv[k] : r[k] * inv[k];
    z[k] = d;
  }
  for (int k = K - 2; k >= 0; --k) z[k] = z[k] + (tri_c(g, G, W, q, k, K) * inv[k]) * z[k + 1];
}
/* out = J_T z for one species: the interface fluxes of z, with the
 * linearised Dirichlet top flux -G_0 q_0 z_0 when dir, zero at the bottom */
static void tri_apply(const double *G, const double *W, const double *q, int K, int dir, const double *z, double *o)
{
  double Fz = dir ? -(G[0] * (z[0] * q[0])) : 0.0;  /* flux through the upper interface of layer k */
  for (int k = 0; k < K; ++k) {
    const double Fd = k < K - 1 ? -(G[k + 1] * (z[k + 1] * q[k + 1] - z[k] * q[k])) : 0.0;
    o[k] = (Fz - Fd) * W[k];
    Fz = Fd;
  }
}

/* Dirichlet top boundary for variable n */
static inline int ros_dir(const oc_params *p, int n) { return (oc_part[n] ? p->bcup_part : p->bcup_dis) == 2; }

typedef struct {
  double Gb[OC_KMAX], Gd[OC_KMAX], W[OC_KMAX], one[OC_KMAX];
  oc_kjac J[OC_KMAX];
  double inv[NV][OC_KMAX];
  double v[NV][OC_KMAX], k0[NV][OC_KMAX], t[NV][OC_KMAX];
  double f0[NV][OC_KMAX], f1[NV][OC_KMAX], k1[NV][OC_KMAX], k2[NV][OC_KMAX], c1[NV][OC_KMAX];
  double s[2][2][OC_KMAX], F[2][NV], top[2][NV];
} roswork;

/* k = W^-1 r (AMF and one correction); top[n] receives the linearised
 * Dirichlet top-flux terms -G_0 q_0 (k0 + v2) of the two transport solves
 * (0 for other species).  Returns 1 on a failed local solve. */
static int ros_wsolve(const oc_params *p, const colview *v, roswork *w, double g, double (*r)[OC_KMAX],
                      double (*k)[OC_KMAX], double *top)
{
  const int K = v->K;
  double a[NV], b[NV];
  for (int n = 0; n < NV; ++n) {
    const double *G = oc_part[n] ? w->Gb : w->Gd, *q = oc_part[n] ? v->phi : w->one;
    tri_solve(g, G, w->W, q, K, w->inv[n], r[n], w->v[n]);
  }
  for (int l = 0; l < K; ++l) {
    for (int n = 0; n < NV; ++n) a[n] = w->v[n][l];
    if (kjac_solve(&w->J[l], g, a, b)) return 1;
    for (int n = 0; n < NV; ++n) w->k0[n][l] = b[n];
  }
  for (int n = 0; n < NV; ++n) top[n] = w->k0[n][0];
  /* residual r - (I - g (J_T + J_R)) k0 = (r - k0) + g (J_T k0 + J_R k0) */
  for (int n = 0; n < NV; ++n) {
    const double *G = oc_part[n] ? w->Gb : w->Gd, *q = oc_part[n] ? v->phi : w->one;
    tri_apply(G, w->W, q, K, ros_dir(p, n), w->k0[n], w->t[n]);
  }
  for (int l = 0; l < K; ++l) {
    for (int n = 0; n < NV; ++n) a[n] = w->k0[n][l];
    kjac_apply(&w->J[l], a, b);
    for (int n = 0; n < NV; ++n) w->t[n][l] = (r[n][l] - w->k0[n][l]) + g * (w->t[n][l] + b[n]);
  }
  for (int n = 0; n < NV; ++n) {
    const double *G = oc_part[n] ? w->Gb : w->Gd, *q = oc_part[n] ? v->phi : w->one;
    tri_solve(g, G, w->W, q, K, w->inv[n], w->t[n], w->v[n]);
    top[n] = ros_dir(p, n) ? 0.0 - ((G[0] * q[0]) * (top[n] + w->v[n][0])) : 0.0;
  }
  for (int l = 0; l < K; ++l) {
    for (int n = 0; n < NV; ++n) a[n] = w->v[n][l];
    if (kjac_solve(&w->J[l], g, a, b)) return 1;
    for (int n = 0; n < NV; ++n) k[n][l] = w->k0[n][l] + b[n];
  }
  return 0;
}

/* Adaptive ROS2 over t_span for one column (same interface, controller and
 * budget layout as col_emp_adaptive).  Budget: the top fluxes integrate as
 *   h/2 [(F(c) + F(c + h k1)) + g (T_1 + T_2)]
 * with T_i the linearised Dirichlet top-flux terms of stage i's solve, and
 * the per-cell sinks as
 *   h/2 [(s(c) + s(c + h k1)) + g (grad s . k1 + grad s . k2)]
 * for s = Cprod (DIC) and 0.8 Denit (N2): exactly what the update moves,
 * since B changes a column inventory only through its top flux and A only
 * through these sinks.
 * Diagnostics: zero_sink counts rejections for negativity, newton rejected
 * local solves. */
static int col_ros2_adaptive(const oc_params *p, const colview *v, double *c, long ks, double t_span,
                             double *F1out, double *cellacc, long ncell, double *colacc, long ncol,
                             oc_coldiag *dg, int64_t max_attempts, roswork *w)
{
  const int K = v->K;
  const long S = OC_KMAX;
  if (v->zw || p->bcup_part == 4) { dg->status = 3; return 3; }
  {
    double B = v->B, fT = oc_fT(v->T);
    for (int k = 0; k < K; ++k) {
      const double Dbio = ((((B * fT) / 86400.0) / 10000.0) * (1.0 - v->phiI[k])) * v->bfac[k];
      const double Ddis = Dbio + (((p->D0 + v->T * 0.035) * v->phiI[k]) / 86400.0) / 10000.0;
      const double len = k == 0 ? v->dz[k] : v->dzc[k];
      w->Gb[k] = Dbio / len;
      w->Gd[k] = Ddis / len;
      w->W[k] = 1.0 / (v->phi[k] * v->dz[k]);
      w->one[k] = 1.0;
    }
  }
  const double fT = oc_fT(v->T);
  const double negtol = 1.0e-3 * p->emp_atol;
  double rem = t_span;
  double dt = dg->dt_next;
  if (!(dt > 0.0)) dt = p->emp_dt_max;
  int64_t attempts = 0;
  while (rem > 0.0) {
    if (attempts >= max_attempts) { dg->status = 2; break; }
    ++attempts;
    int clipped = !(dt < rem);
    double h = clipped ? rem : dt;
    const double g = ROS_GAMMA * h;
    col_rhs(p, v, c, ks, &w->f0[0][0], S, w->F[0], w->s[0][0], w->s[0][1]);
    for (int l = 0; l < K; ++l) {
      double a[NV];
      for (int n = 0; n < NV; ++n) a[n] = c[n * ks + l];
      oc_kin_jac(p, fT, a, &w->J[l]);
    }
    for (int n = 0; n < NV; ++n) {
      const double *G = oc_part[n] ? w->Gb : w->Gd, *q = oc_part[n] ? v->phi : w->one;
      tri_factor(g, G, w->W, q, K, ros_dir(p, n), w->inv[n]);
    }
    int fail = ros_wsolve(p, v, w, g, w->f0, w->k1, w->top[0]);
    if (!fail) {
      for (int n = 0; n < NV; ++n)
        for (int l = 0; l < K; ++l) {
          double y = c[n * ks + l] + h * w->k1[n][l];
          w->c1[n][l] = y < 0.0 ? 0.0 : y;
        }
      col_rhs(p, v, &w->c1[0][0], S, &w->f1[0][0], S, w->F[1], w->s[1][0], w->s[1][1]);
      for (int n = 0; n < NV; ++n)
        for (int l = 0; l < K; ++l) w->f1[n][l] = w->f1[n][l] - 2.0 * w->k1[n][l];
      fail = ros_wsolve(p, v, w, g, w->f1, w->k2, w->top[1]);
    }
    double err = 0.0;
    int bad = 0, neg = 0;
    if (!fail) {
      for (int l = 0; l < K; ++l) {
        double bn = 0.0, bd = 1.0;
        for (int n = 0; n < NV; ++n) {
          const double k1 = w->k1[n][l], k2 = w->k2[n][l], c0 = c[n * ks + l];
          const double y = c0 + h * (1.5 * k1 + 0.5 * k2);
          w->c1[n][l] = y;
          const double yp = y < 0.0 ? 0.0 : y;
          const double sc = p->emp_atol + p->emp_rtol * (yp > c0 ? yp : c0);
          const double dv = fabs(h * (0.5 * k1 + 0.5 * k2));
          if (dv != dv || sc != sc) bad = 1;
          if (y < 0.0 - negtol) neg = 1;
          if (dv * bd > bn * sc) { bn = dv; bd = sc; }
        }
        const double e = bn / bd;
        if (e != e) bad = 1;
        if (e > err) err = e;
      }
    }
    if (bad) { dg->status = 1; break; }
    if (fail && !(h > p->emp_dt_min)) { dg->status = 1; break; }
    const int accept = !fail && ((err <= 1.0 && !neg) || !(h > p->emp_dt_min));
    double fac = err > 0.0 ? 0.9 * oc_rsqrt(err) : 5.0;
    if (fac < 0.2) fac = 0.2;
    if (fac > 5.0) fac = 5.0;
    if (accept) {
      for (int n = 0; n < NV; ++n)
        for (int l = 0; l < K; ++l) {
          double y = w->c1[n][l];
          if (y < 0.0) {
            if (cellacc && oc_elem[n] >= 0) ACC_CELL(cellacc, ncell, 2 + oc_elem[n], l, 0.0 - y);
            y = 0.0;
          }
          c[n * ks + l] = y;
        }
      if (colacc)
        for (int n = 0; n < NV; ++n)
          ACC_COL(colacc, ncol, n, (h * 0.5) * ((w->F[0][n] + w->F[1][n]) + g * (w->top[0][n] + w->top[1][n])));
      if (cellacc)
        for (int l = 0; l < K; ++l) {
          double a1[NV], a2[NV];
          for (int n = 0; n < NV; ++n) { a1[n] = w->k1[n][l]; a2[n] = w->k2[n][l]; }
          const double gC = kjac_sinkC(&w->J[l], a1) + kjac_sinkC(&w->J[l], a2);
          const double gN = kjac_sinkN(&w->J[l], a1) + kjac_sinkN(&w->J[l], a2);
          ACC_CELL(cellacc, ncell, 0, l, (h * 0.5) * ((w->s[0][0][l] + w->s[1][0][l]) + g * gC));
          ACC_CELL(cellacc, ncell, 1, l, (h * 0.5) * ((w->s[0][1][l] + w->s[1][1][l]) + g * gN));
        }
      memcpy(F1out, w->F[1], sizeof(double) * NV);
      rem = rem - h;
      dg->accepted += 1;
      if (err > 1.0 || neg) dg->forced += 1;
      dg->dt_last = h;
      if (h < dg->dt_minacc) dg->dt_minacc = h;
      double prop = h * fac;
      dt = (clipped && dt > prop) ? dt : prop;
    } else {
      dg->rejected += 1;
      if (fail) {
        dg->newton += 1;
        dt = h * 0.25;
      } else if (err <= 1.0) {
        dg->zero_sink += 1;
        dt = h * 0.5;
      } else {
        dt = h * fac;
      }
    }
    if (dt < p->emp_dt_min) dt = p->emp_dt_min;
    if (dt > p->emp_dt_max) dt = p->emp_dt_max;
  }
  dg->dt_next = dt;
  return dg->status;
}

/* ------------------------------------------------------------------------ */
/* domain-level entry points                                                 */
/* ------------------------------------------------------------------------ */

static int check_domain(const oc_domain *d)
{
  if (d->ncol < 0) return -1;
  for (int c = 0; c < d->ncol; ++c) {
    int K = d->off[c + 1] - d->off[c];
    if (K < 1 || K > OC_KMAX) return -2;
  }
  return 0;
}

/* get_rhs for the whole domain; masked columns get rhs = 0 and flux = 0 */
int oc_rhs(const oc_params *p, const oc_domain *d, const double *conc,
           double *rhs, double *flux_out)
{
  if (check_domain(d)) return -1;
  const long ncell = d->ncell;
  for (int c = 0; c < d->ncol; ++c) {
    int o = d->off[c];
    int K = d->off[c + 1] - o;
    double F1[NV], sC[OC_KMAX], sN[OC_KMAX];
    if (d->mask[c]) {
      for (int n = 0; n < NV; ++n) {
        for (int k = 0; k < K; ++k) rhs[n * ncell + o + k] = 0.0;
        if (flux_out) flux_out[(long)n * d->ncol + c] = 0.0;
      }
      continue;
    }
    colview v;
    make_view(d, c, &v);
    col_rhs(p, &v, conc + o, ncell, rhs + o, ncell, F1, sC, sN);
    if (flux_out)
      for (int n = 0; n < NV; ++n) flux_out[(long)n * d->ncol + c] = F1[n];
  }
  return 0;
}

static void diag_init(oc_coldiag *dg)
{
  if (dg->dt_minacc == 0.0) dg->dt_minacc = INFINITY;
  if (dg->p_min == 0.0) dg->p_min = 1.0;
}

/* One call of ode_solver(rhs_driver, dt, method) on the whole domain
 * (solver_library.F90:80-189), methods 0..5; 6 integrates dt adaptively per
 * column.  ADAPTIVE_EULER is global: any cell anywhere with
 * c1 - (1 + relative_change_min) c < 0 quarters the step for every column
 * (:119-129).  last_min_dt (may be NULL) follows :130-135. */
int oc_ode_solver(const oc_params *p, const oc_domain *d, double *conc,
                  double dt, int method, double *flux_out, double *cellacc,
                  double *colacc, oc_coldiag *diag, double *last_min_dt)
{
  if (check_domain(d)) return -1;
  const long ncell = d->ncell, ncol = d->ncol;
  int rc = 0;
  if (method == OC_ADAPTIVE_EULER) {
    double *r = (double *)malloc(sizeof(double) * NV * ncell);
    double *c1 = (double *)malloc(sizeof(double) * NV * ncell);
    double *F = (double *)calloc((size_t)NV * ncol, sizeof(double));
    double *s = (double *)malloc(sizeof(double) * 2 * ncell);
    double dt_int = 0.0, dt_red = dt;
    while (dt_int < dt) {
      for (int c = 0; c < ncol; ++c) {
        int o = d->off[c], K = d->off[c + 1] - o;
        if (d->mask[c]) {
          for (int n = 0; n < NV; ++n) for (int k = 0; k < K; ++k) r[n * ncell + o + k] = 0.0;
          continue;
        }
        colview v;
        make_view(d, c, &v);
        double F1[NV];
        col_rhs(p, &v, conc + o, ncell, r + o, ncell, F1, s + o, s + ncell + o);
        for (int n = 0; n < NV; ++n) F[n * ncol + c] = F1[n];
      }
      int viol = 0;
      for (long i = 0; i < NV * ncell; ++i) {
        c1[i] = conc[i] + dt_red * r[i];
        if ((c1[i] - (1.0 + p->relative_change_min) * conc[i]) < 0.0) viol = 1;
      }
      /* masked columns: rhs = 0, so c1 = c and the test is not triggered
       * for their (positive) missing values */
      if (viol && dt_red > p->dt_min) {
        dt_red = dt_red * 0.25;
      } else {
        if (last_min_dt && dt_red < *last_min_dt) *last_min_dt = dt_red;
        for (int c = 0; c < ncol; ++c) {
          if (d->mask[c]) continue;
          int o = d->off[c], K = d->off[c + 1] - o;
          for (int n = 0; n < NV; ++n)
            for (int k = 0; k < K; ++k) conc[n * ncell + o + k] = c1[n * ncell + o + k];
          if (colacc) for (int n = 0; n < NV; ++n) ACC_COL(colacc + c, ncol, n, dt_red * F[n * ncol + c]);
          if (cellacc)
            for (int k = 0; k < K; ++k) {
              ACC_CELL(cellacc + o, ncell, 0, k, dt_red * s[o + k]);
              ACC_CELL(cellacc + o, ncell, 1, k, dt_red * s[ncell + o + k]);
            }
          if (diag) {
            diag[c].accepted += 1;
            diag[c].dt_last = dt_red;
            diag_init(&diag[c]);
            if (dt_red < diag[c].dt_minacc) diag[c].dt_minacc = dt_red;
          }
        }
        dt_int = dt_int + dt_red;
      }
    }
    if (flux_out) memcpy(flux_out, F, sizeof(double) * NV * ncol);
    if (p->clamp)
      for (int c = 0; c < ncol; ++c) {
        if (d->mask[c]) continue;
        colview v;
        make_view(d, c, &v);
        int o = d->off[c];
        clamp_col(&v, conc + o, ncell, cellacc ? cellacc + o : NULL, ncell);
      }
    for (long i = 0; i < NV * ncell; ++i)
      if (conc[i] != conc[i]) rc = 1;
    free(r); free(c1); free(F); free(s);
    return rc;
  }
  colwork *w = (colwork *)malloc(sizeof(colwork));
  for (int c = 0; c < ncol; ++c) {
    if (d->mask[c]) continue;
    colview v;
    make_view(d, c, &v);
    int o = d->off[c];
    double F1[NV];
    oc_coldiag dloc;
    memset(&dloc, 0, sizeof dloc);
    oc_coldiag *dg = diag ? &diag[c] : &dloc;
    diag_init(dg);
    int r;
    if (method == OC_EMP_ADAPTIVE)
      r = col_emp_adaptive(p, &v, conc + o, ncell, dt, F1, cellacc ? cellacc + o : NULL, ncell,
                           colacc ? colacc + c : NULL, ncol, dg, INT64_MAX, w);
    else if (method == OC_ROS2_ADAPTIVE) {
      roswork *rw = (roswork *)malloc(sizeof(roswork));
      r = col_ros2_adaptive(p, &v, conc + o, ncell, dt, F1, cellacc ? cellacc + o : NULL, ncell,
                            colacc ? colacc + c : NULL, ncol, dg, INT64_MAX, rw);
      free(rw);
    }
    else
      r = col_step(p, &v, method, dt, conc + o, ncell, F1, cellacc ? cellacc + o : NULL, ncell,
                   colacc ? colacc + c : NULL, ncol, dg, w);
    if (p->clamp && method <= OC_RK4_38)
      clamp_col(&v, conc + o, ncell, cellacc ? cellacc + o : NULL, ncell);
    if (flux_out) for (int n = 0; n < NV; ++n) flux_out[(long)n * ncol + c] = F1[n];
    if (r) rc = 1;
  }
  free(w);
  return rc;
}

/* nsteps consecutive ode_solver calls (plus the caller's clamp when
 * p->clamp), columns in parallel for the column-local methods. */
int oc_run(const oc_params *p, const oc_domain *d, double *conc, double dt,
           int nsteps, int method, double *flux_out, double *cellacc,
           double *colacc, oc_coldiag *diag, int nthreads)
{
  if (check_domain(d)) return -1;
  if (method == OC_ADAPTIVE_EULER) {
    int rc = 0;
    for (int s = 0; s < nsteps; ++s)
      rc |= oc_ode_solver(p, d, conc, dt, method, flux_out, cellacc, colacc, diag, NULL);
    return rc;
  }
  const long ncell = d->ncell, ncol = d->ncol;
  int rc = 0;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(| : rc)
#endif
  {
    colwork *w = (colwork *)malloc(sizeof(colwork));
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int c = 0; c < ncol; ++c) {
      if (d->mask[c]) continue;
      colview v;
      make_view(d, c, &v);
      int o = d->off[c];
      double F1[NV];
      oc_coldiag dloc;
      memset(&dloc, 0, sizeof dloc);
      oc_coldiag *dg = diag ? &diag[c] : &dloc;
      diag_init(dg);
      for (int s = 0; s < nsteps; ++s) {
        int r = col_step(p, &v, method, dt, conc + o, ncell, F1, cellacc ? cellacc + o : NULL, ncell,
                         colacc ? colacc + c : NULL, ncol, dg, w);
        if (p->clamp && method <= OC_RK4_38)
          clamp_col(&v, conc + o, ncell, cellacc ? cellacc + o : NULL, ncell);
        if (r) { rc |= 1; dg->status = 1; break; }
      }
      if (flux_out) for (int n = 0; n < NV; ++n) flux_out[(long)n * ncol + c] = F1[n];
    }
    free(w);
  }
  return rc;
}

/* Adaptive EMP over t_span for every wet column (columns in parallel). */
int oc_emp_integrate(const oc_params *p, const oc_domain *d, double *conc,
                     double t_span, double *flux_out, double *cellacc,
                     double *colacc, oc_coldiag *diag, int64_t max_attempts,
                     int nthreads)
{
  if (check_domain(d)) return -1;
  const long ncell = d->ncell, ncol = d->ncol;
  int rc = 0;
#ifdef _OPENMP
  if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads) reduction(| : rc)
#endif
  {
    colwork *w = (colwork *)malloc(sizeof(colwork));
    roswork *rw = p->adaptive_method == OC_ROS2_ADAPTIVE ? (roswork *)malloc(sizeof(roswork)) : NULL;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
    for (int c = 0; c < ncol; ++c) {
      if (d->mask[c]) continue;
      colview v;
      make_view(d, c, &v);
      int o = d->off[c];
      double F1[NV];
      for (int n = 0; n < NV; ++n) F1[n] = flux_out ? flux_out[(long)n * ncol + c] : 0.0;
      diag_init(&diag[c]);
      int r = rw ? col_ros2_adaptive(p, &v, conc + o, ncell, t_span, F1, cellacc ? cellacc + o : NULL,
                                     ncell, colacc ? colacc + c : NULL, ncol, &diag[c], max_attempts, rw)
                 : col_emp_adaptive(p, &v, conc + o, ncell, t_span, F1, cellacc ? cellacc + o : NULL,
                                    ncell, colacc ? colacc + c : NULL, ncol, &diag[c], max_attempts, w);
      if (flux_out) for (int n = 0; n < NV; ++n) flux_out[(long)n * ncol + c] = F1[n];
      if (r) rc |= r;
    }
    free(w);
    free(rw);
  }
  return rc;
}

/* budget[c*OC_NBUD + i]: i<11 integrated top flux per variable, 11 sinkC,
 * 12 sinkN, 13..15 clamp C/N/P; cell terms weighted by phi*dz and summed
 * over layers with the pairwise tree. */
int oc_budget(const oc_domain *d, const double *cellacc, const double *colacc,
              double *budget)
{
  const long ncell = d->ncell, ncol = d->ncol;
  for (int c = 0; c < ncol; ++c) {
    int o = d->off[c], K = d->off[c + 1] - o;
    double *b = budget + (long)c * OC_NBUD;
    for (int n = 0; n < NV; ++n)
      b[n] = colacc ? colacc[(long)n * ncol + c] + colacc[(long)(NV + n) * ncol + c] : 0.0;
    for (int a = 0; a < 5; ++a) {
      double t[OC_KMAX];
      for (int k = 0; k < K; ++k)
        t[k] = (d->phi[o + k] * d->dz[o + k]) * (cellacc[a * ncell + o + k] + cellacc[(5 + a) * ncell + o + k]);
      b[NV + a] = tree_sum(t, K);
    }
  }
  return 0;
}

/* inv[c*NV + n] = sum_k phi_k dz_k c_{k,n} (pairwise tree) */
int oc_inventory(const oc_domain *d, const double *conc, double *inv)
{
  const long ncell = d->ncell;
  for (int c = 0; c < d->ncol; ++c) {
    int o = d->off[c], K = d->off[c + 1] - o;
    for (int n = 0; n < NV; ++n) {
      double t[OC_KMAX];
      for (int k = 0; k < K; ++k) t[k] = (d->phi[o + k] * d->dz[o + k]) * conc[n * ncell + o + k];
      inv[(long)c * NV + n] = tree_sum(t, K);
    }
  }
  return 0;
}

/* resid[c] = max_{k,n} |rhs_{k,n}| / max(|c_{k,n}|, cfloor[n])  (s-1) */
void oc_assemble_forcing(const oc_params *p, const oc_domain *d, const double *conc,
                         const double *temp_surface, const double *c_surface,
                         const double *w_surface, const int32_t *present,
                         double *temp, double *cbw, double *fluxup)
{
  for (int c = 0; c < d->ncol; ++c) {
    if (d->mask && d->mask[c]) continue;
    if (temp_surface) temp[c] = temp_surface[c]; /* :1903-1939 */
    const long top = d->off[c];
    for (int n = 0; n < NV; ++n) {
      if (!present[n]) continue; /* field absent: skipped (:1963-1967) */
      const long i = (long)n * d->ncol + c;
      if (oc_part[n]) {
        fluxup[i] = -(c_surface[i] * w_surface[i]); /* :1978-1986, downward positive */
      } else {
        cbw[i] = c_surface[i]; /* :1989-2000 */
        if (p->bcup_dis == 1) {
          /* :2012-2014, in its operation order */
          double x = (conc[(long)n * d->ncell + top] - cbw[i]) / d->dz[top];
          x = x * ((d->biot[c] + p->D0) + temp[c] * 0.035);
          x = x * d->phi[top];
          x = x / 86400.0;
          x = x / 10000.0;
          fluxup[i] = -x;
        } else {
          fluxup[i] = 0.0; /* :2021 */
        }
      }
    }
  }
}

int oc_residual(const oc_params *p, const oc_domain *d, const double *conc,
                const double *cfloor, double *resid)
{
  if (check_domain(d)) return -1;
  const long ncell = d->ncell;
  double *r = (double *)malloc(sizeof(double) * NV * ncell);
  oc_rhs(p, d, conc, r, NULL);
  for (int c = 0; c < d->ncol; ++c) {
    int o = d->off[c], K = d->off[c + 1] - o;
    double m = 0.0;
    if (!d->mask[c])
      for (int n = 0; n < NV; ++n)
        for (int k = 0; k < K; ++k) {
          double x = fabs(conc[n * ncell + o + k]);
          double den = x > cfloor[n] ? x : cfloor[n];
          double e = fabs(r[n * ncell + o + k]) / den;
          if (e > m || e != e) m = e;
        }
    resid[c] = m;
  }
  free(r);
  return 0;
}
