/*
 * kt_oracle.c -- CPU FP64 restatement of gpu-vh's Kurganov-Tadmor viscous
 * hydrodynamics step.  TEST INFRASTRUCTURE ONLY (see kt_oracle.h).
 *
 * Each function names the reference file:line whose behaviour it restates
 * (paths relative to the reference root, rhic/src unless stated).  The
 * structure deliberately mirrors the reference's kernel split (Source, X, Y,
 * Z Euler kernels, inferred variables, regulation, ghost cells, convex
 * average) so that a reader can check it side by side; loops over cells are
 * plain OpenMP loops.
 *
 * Built with -ffp-contract=off: every operation rounds as written.
 */
#include "kt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NG 2          /* ghost layers, LatticeParameters.h:13-15          */
#define NVMAX 15      /* conserved variables with PI, DynamicalVariables.cuh:41 */
#define MAX_ITERS 1000000 /* EnergyMomentumTensor.cu:15                      */
#define THETA 2.0     /* FluxLimiter.cu:15                                  */
#define EOS_FACTOR 13.8997 /* EquationOfState.cuh:24                        */

/* transport constants, TransportCoefficients.cuh:14-18 (PRECISION=double) */
static const double delta_pipi = 1.33333;
static const double tau_pipi = 1.42857;
static const double lambda_piPi = 1.2;
static const double delta_PiPi = 0.666667;

/* float literals used by SourceTerms.cu:47-57, kept at float value */
#define F_0P2 ((double)0.2f)
#define F_THIRD ((double)0.333333f)
#define F_1P6 ((double)1.6f)

typedef struct kto_ctx {
  kto_params P;
  int ncx, ncy, ncz;
  long G;
  int nv;                       /* 14 (shear only) or 15 (shear + bulk)   */
  double *qa[NVMAX], *qb[NVMAX], *qs[NVMAX]; /* q^n, Q, qS buffers          */
  double **q, **Q, **qS;        /* current role of each buffer             */
  double *ua[4], *ub[4], *uc[4];
  double **u, **up, **uS;
  double *e, *p, *reg;
} kto_ctx;

static __thread int g_last_iters;

/* ------------------------------------------------------------------ EoS */

/* pseudo-log index of the tabulated EoS: nodes at 2^E * (1 + j/M) */
static void tab_locate(const kto_params *P, double e, int *k, double *frac) {
  const int M = P->tab_per_octave, N = P->tab_n;
#define NODE_E(kk) ldexp(1.0 + (double)((kk) % M) / M, P->tab_emin_exp + (kk) / M)
  const double e_first = NODE_E(0), e_last = NODE_E(N - 1);
  int kk;
  double fr;
  if (e >= e_first && e < e_last) {
    int E;
    double m = frexp(e, &E); /* e = m 2^E, m in [0.5,1) */
    double f = (2.0 * m - 1.0) * (double)M;
    int j = (int)f;
    fr = f - (double)j;
    kk = (E - 1 - P->tab_emin_exp) * M + j;
    if (kk > N - 2) { /* f rounded up onto the last node */
      kk = N - 2;
      fr += 1.0;
    }
  } else if (e >= e_last) { /* beyond the table (and +inf): extrapolate */
    kk = N - 2;
    fr = (e - NODE_E(N - 2)) / (e_last - NODE_E(N - 2));
  } else { /* below the first node, or NaN: extrapolate the first interval */
    kk = 0;
    fr = (e - e_first) / (NODE_E(1) - e_first);
  }
#undef NODE_E
  *k = kk;
  *frac = fr;
}

static double tab_interp(const double *y, int k, double fr) {
  return y[k] + fr * (y[k + 1] - y[k]);
}

/* equilibriumPressure, EquationOfState.cu:31-80 */
static double eos_p(const kto_params *P, double e) {
  if (P->eos == KTO_EOS_CONFORMAL) return e / 3;
  if (P->eos == KTO_EOS_TABLE) {
    int k; double fr;
    tab_locate(P, e, &k, &fr);
    return tab_interp(P->tab_p, k, fr);
  }
  double e1 = e, e2 = e * e, e3 = e2 * e, e4 = e3 * e, e5 = e4 * e, e6 = e5 * e;
  double e7 = e6 * e, e8 = e7 * e, e9 = e8 * e, e10 = e9 * e, e11 = e10 * e,
         e12 = e11 * e;
  static const double a[13] = {
      -0.25181736420168666, 9737.845799644809, 1.077580993288114e6,
      3.1729694865420084e6, 1.6357487344679043e6, 334334.4309240126,
      41913.439282708554, 6340.448389300905, 141.5073484468774,
      0.7158279081255019, 0.0009417586777847889, 3.1188455176941583e-7,
      1.9531729608963267e-11};
  static const double b[13] = {
      45829.44617893836, 4.0574329080826794e6, 2.0931169138134286e7,
      1.3512402226067686e7, 1.7851642641834426e6, 278581.2989342773,
      26452.34905933697, 499.04919730607065, 2.3405487982094204,
      0.002962497695527404, 9.601103399348206e-7, 5.928138360995685e-11,
      3.2581066229887368e-18};
  const double ep[13] = {1, e1, e2, e3, e4, e5, e6, e7, e8, e9, e10, e11, e12};
  /* nested fma from the highest power inwards: fma(a12,e12,fma(a11,...)) */
  double na = a[0], nb = b[0];
  na = fma(a[1], ep[1], a[0]);
  nb = fma(b[1], ep[1], b[0]);
  for (int i = 2; i <= 12; ++i) {
    na = fma(a[i], ep[i], na);
    nb = fma(b[i], ep[i], nb);
  }
  return na / nb;
}

/* speedOfSoundSquared, EquationOfState.cu:83-116 (conformal returns the
 * integer quotient 1/3 == 0, :114) */
static double eos_cs2(const kto_params *P, double e) {
  if (P->eos == KTO_EOS_CONFORMAL) return 0.0;
  if (P->eos == KTO_EOS_TABLE) {
    int k; double fr;
    tab_locate(P, e, &k, &fr);
    return tab_interp(P->tab_cs2, k, fr);
  }
  double e1 = e, e2 = e * e1, e3 = e2 * e1, e4 = e3 * e1, e5 = e4 * e1,
         e6 = e5 * e1, e7 = e6 * e1, e8 = e7 * e1, e9 = e8 * e1, e10 = e9 * e1,
         e11 = e10 * e1, e12 = e11 * e1, e13 = e12 * e1;
  double num = 5.191934309650155e-32 + 4.123605749683891e-23 * e +
               3.1955868410879504e-16 * e2 + 1.4170364808063119e-10 * e3 +
               6.087136671592452e-6 * e4 + 0.02969737949090831 * e5 +
               15.382615282179595 * e6 + 460.6487249985994 * e7 +
               1612.4245252438795 * e8 + 275.0492627924299 * e9 +
               58.60283714484669 * e10 + 6.504847576502024 * e11 +
               0.03009027913262399 * e12 + 8.189430244031285e-6 * e13;
  double den = 1.4637868900982493e-30 + 6.716598285341542e-22 * e +
               3.5477700458515908e-15 * e2 + 1.1225580509306008e-9 * e3 +
               0.00003551782901018317 * e4 + 0.13653226327408863 * e5 +
               60.85769171450653 * e6 + 1800.5461219450308 * e7 +
               15190.225535036281 * e8 + 590.2572000057821 * e9 +
               293.99144775704605 * e10 + 21.461303090563028 * e11 +
               0.09301685073435291 * e12 + 0.000024810902623582917 * e13;
  return num / den;
}

/* effectiveTemperature, EquationOfState.cu:119-148 */
static double eos_T(const kto_params *P, double e) {
  if (P->eos == KTO_EOS_CONFORMAL) return pow(e / EOS_FACTOR, 0.25);
  if (P->eos == KTO_EOS_TABLE) {
    int k; double fr;
    tab_locate(P, e, &k, &fr);
    return tab_interp(P->tab_T, k, fr);
  }
  double e1 = e, e2 = e * e1, e3 = e2 * e1, e4 = e3 * e1, e5 = e4 * e1,
         e6 = e5 * e1, e7 = e6 * e1, e8 = e7 * e1, e9 = e8 * e1, e10 = e9 * e1,
         e11 = e10 * e1;
  double num = 1.510073201405604e-29 + 8.014062800678687e-18 * e +
               2.4954778310451065e-10 * e2 + 0.000063810382643387 * e3 +
               0.4873490574161924 * e4 + 207.48582344326206 * e5 +
               6686.07424325115 * e6 + 14109.766109389702 * e7 +
               1471.6180520527757 * e8 + 14.055788949565482 * e9 +
               0.015421252394182246 * e10 + 1.5780479034557783e-6 * e11;
  double den = 7.558667139355393e-28 + 1.3686372302041508e-16 * e +
               2.998130743142826e-9 * e2 + 0.0005036835870305458 * e3 +
               2.316902328874072 * e4 + 578.0778724946719 * e5 +
               11179.193315394154 * e6 + 17965.67607192861 * e7 +
               1051.0730543534657 * e8 + 5.916312075925817 * e9 +
               0.003778342768228011 * e10 + 1.8472801679382593e-7 * e11;
  return num / den;
}

/* equilibriumEnergyDensity, EquationOfState.cu:151-204.  The tabulated form
 * inverts the monotone T(e) table by bisection + linear interpolation. */
double kto_e_of_T(const kto_params *P, double T) {
  if (P->eos == KTO_EOS_CONFORMAL) return EOS_FACTOR * pow(T, 4.0);
  if (P->eos == KTO_EOS_TABLE) {
    const int N = P->tab_n, M = P->tab_per_octave;
    const double *Tt = P->tab_T;
#define NODE_E(kk) ldexp(1.0 + (double)((kk) % M) / M, P->tab_emin_exp + (kk) / M)
    if (T <= Tt[0]) {
      if (T <= 0) return 0.0;
      return NODE_E(0) * pow(T / Tt[0], 4.0);
    }
    if (T >= Tt[N - 1]) return NODE_E(N - 1) * pow(T / Tt[N - 1], 4.0);
    int lo = 0, hi = N - 1;
    while (hi - lo > 1) {
      int mid = (lo + hi) / 2;
      if (Tt[mid] <= T) lo = mid; else hi = mid;
    }
    double f = (T - Tt[lo]) / (Tt[hi] - Tt[lo]);
    return NODE_E(lo) + f * (NODE_E(hi) - NODE_E(lo));
#undef NODE_E
  }
  double T1 = T, T2 = T1 * T1, T3 = T2 * T1, T4 = T3 * T1, T5 = T4 * T1,
         T6 = T5 * T1, T7 = T6 * T1, T8 = T7 * T1, T9 = T8 * T1, T10 = T9 * T1,
         T11 = T10 * T1, T12 = T11 * T1, T13 = T12 * T1, T14 = T13 * T1,
         T15 = T14 * T1, T16 = T15 * T1, T17 = T16 * T1, T18 = T17 * T1,
         T19 = T18 * T1, T20 = T19 * T1, T21 = T20 * T1, T22 = T21 * T1,
         T23 = T22 * T1;
  double num = -0.011958188410851651 + 119.89423098138208 * T -
               3156.9475699248055 * T2 + 32732.86844374939 * T3 -
               187899.8994764422 * T4 + 712537.3610845465 * T5 -
               1.557049803609345e6 * T6 + 1.4852519861308339e6 * T7 +
               532132.6079941876 * T8 - 1.963099445042592e6 * T9 -
               4484.44579242679 * T10 + 1.7984228830058286e6 * T11 +
               119345.25619517374 * T12 - 1.3499773937058165e6 * T13 -
               207838.4995663606 * T14 + 654970.2138652403 * T15 -
               78643.00334616247 * T16 + 40274.00078068926 * T17 +
               422619.58977657766 * T18 - 409688.07836393174 * T19 -
               62005.75915066359 * T20 + 46788.14270090656 * T21 +
               40784.330477857235 * T22 - 12589.47744840392 * T23;
  double den = 31630.074365558292 - 127100.88940643385 * T +
               173528.1225422275 * T2 - 39403.297956865215 * T3 -
               85582.57873541754 * T4 + 9320.560804233442 * T5 +
               50882.74198960172 * T6 + 20335.926473421183 * T7 -
               14897.725710713818 * T8 - 23836.484117457 * T9 -
               13726.013896090335 * T10 + 4517.908673107615 * T11 +
               18056.19917986404 * T12 + 14954.82860467155 * T13 +
               2569.623976952738 * T14 - 9304.046211514986 * T15 -
               15606.429173842751 * T16 + 8383.710735812094 * T17 +
               1591.3177623932843 * T18 - 678.748230997762 * T19 -
               33.58687934953277 * T20 + 3.2520554133126285 * T21 -
               0.19647288043440464 * T22 + 0.005443394551264717 * T23;
  return num / den;
}

void kto_eos_p_batch(const kto_params *P, long n, const double *e, double *p) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n; ++i) p[i] = eos_p(P, e[i]);
}

void kto_eos(const kto_params *P, double e, double *p, double *cs2, double *T) {
  *p = eos_p(P, e);
  *cs2 = eos_cs2(P, e);
  *T = eos_T(P, e);
}

/* bulkViscosityToEntropyDensity, TransportCoefficients.cu:27-36, with the
 * temperature scale as a parameter (reference value 1.01355 fm^-1). */
double kto_zeta_over_s(const kto_params *P, double T) {
  double x = T / P->zeta_tpeak;
  if (x > 1.05)
    return 0.9 * exp(-(x - 1) / 0.025) + 0.25 * exp(-(x - 1) / 0.13) + 0.001;
  else if (x < 0.995)
    return 0.9 * exp((x - 1) / 0.0025) + 0.22 * exp((x - 1) / 0.022) + 0.03;
  else
    return -13.77 * x * x + 27.55 * x - 13.45;
}

/* --------------------------------------------------------- flux limiter */

/* sign / minmod / minmod3 / approximateDerivative, FluxLimiter.cu:17-39 */
static double sgn(double x) { return (x < 0) ? -1.0 : 1.0; }

double kto_minmod(double x, double y) {
  return (sgn(x) + sgn(y)) * fmin(fabs(x), fabs(y)) / 2;
}

double kto_approx_deriv(double x, double y, double z) {
  double left = THETA * (y - x);
  double ctr = (z - x) / 2;
  double right = THETA * (z - y);
  return kto_minmod(left, kto_minmod(ctr, right));
}

/* --------------------------------------------------- inferred variables */

/* energyDensityFromConservedVariables, EnergyMomentumTensor.cu:19-48 */
static double energy_from_conserved(const kto_params *P, double ePrev,
                                    double M0, double M, double Pi) {
  if (P->eos == KTO_EOS_CONFORMAL) {
    g_last_iters = 0;
    return fabs(sqrt(fabs(4 * M0 * M0 - 3 * M)) - M0);
  }
  double e0 = ePrev;
  for (int j = 0; j < MAX_ITERS; ++j) {
    double p = eos_p(P, e0);
    double cs2 = eos_cs2(P, e0);
    double cst2 = p / e0;
    double A = fma(M0, 1 - cst2, Pi);
    double B = fma(M0, M0 + Pi, -M);
    double H = sqrt(fabs(A * A + 4 * cst2 * B));
    double D = (A - H) / (2 * cst2);
    double f = e0 + D;
    double fp = 1 - ((cs2 - cst2) *
                     (B + D * H - ((cs2 - cst2) * cst2 * D * M0) / e0)) /
                        (cst2 * e0 * H);
    double e;
    if (j < MAX_ITERS / 2) e = e0 - f / fp;
    else e = e0 - 0.9 * f / fp;
    if (fabs(e - e0) <= 0.001 * fabs(e)) {
      g_last_iters = j + 1;
      return e;
    }
    e0 = e;
    /* a NaN iterate stays NaN to MAX_ITERS and is returned: same value */
    if (isnan(e0)) break;
  }
  g_last_iters = MAX_ITERS;
  return e0;
}

int kto_last_newton_iters(void) { return g_last_iters; }

/* getInferredVariables, EnergyMomentumTensor.cu:51-108.  The Pi consistency
 * fix at :82-85 is guarded by '#ifdef Pi' (a variable, not a macro) and is
 * therefore never compiled; it is omitted here too. */
static void inferred(const kto_params *P, double t, const double *q,
                     double ePrev, double *e, double *p, double *ut,
                     double *ux, double *uy, double *un) {
  double Pi = P->bulk ? q[14] : 0.0;
  double M0 = q[0] - q[4];
  double M1 = q[1] - q[5];
  double M2 = q[2] - q[6];
  double M3 = q[3] - q[7];
  double M = M1 * M1 + M2 * M2 + t * t * M3 * M3;
  double ee;
  if (ePrev <= 0.1) {
    ee = M0 - M / M0;
    g_last_iters = 0;
  } else {
    ee = energy_from_conserved(P, ePrev, M0, M, Pi);
  }
  double pp = eos_p(P, ee);
  if (ee < 1.e-7) {
    ee = 1.e-7;
    pp = 1.e-7;
  }
  double PP = pp + Pi;
  double E = 1 / (ee + PP);
  double uT = sqrt(fabs((M0 + PP) * E));
  double E2 = E / uT;
  *e = ee;
  *p = pp;
  *ut = uT;
  *ux = M1 * E2;
  *uy = M2 * E2;
  *un = M3 * E2;
}

void kto_inferred(const kto_params *P, double t, const double *q, double ePrev,
                  double *out) {
  inferred(P, t, q, ePrev, out, out + 1, out + 2, out + 3, out + 4, out + 5);
}

/* ------------------------------------------------------------ KT flux */

/* half-site extrapolations, HalfSiteExtrapolation.cu:16-29 */
static double right_fwd(const double *d) { return d[3] - kto_approx_deriv(d[2], d[3], d[4]) / 2; }
static double left_fwd(const double *d) { return d[2] + kto_approx_deriv(d[1], d[2], d[3]) / 2; }
static double right_bwd(const double *d) { return d[2] - kto_approx_deriv(d[1], d[2], d[3]) / 2; }
static double left_bwd(const double *d) { return d[1] + kto_approx_deriv(d[0], d[1], d[2]) / 2; }

/* flux, SemiDiscreteKurganovTadmorScheme.cu:16-57, with the spectral
 * radius |u_dir/u^tau| (SpectralRadius.cu:18-30), local speed
 * max(left,right) (LocalPropagationSpeed.cu:16-24) and F_dir = u_dir q/u^tau
 * (FluxFunctions.cu:19-31). */
static void flux(const kto_params *P, int nv, int dir, const double *data,
                 double t, double ePrev, int fwd, double *H) {
  double qR[NVMAX] = {0}, qL[NVMAX] = {0};
  for (int n = 0; n < nv; ++n) {
    const double *d = data + 5 * n;
    qR[n] = fwd ? right_fwd(d) : right_bwd(d);
    qL[n] = fwd ? left_fwd(d) : left_bwd(d);
  }
  double eR, pR, uR[4], eL, pL, uL[4];
  inferred(P, t, qR, ePrev, &eR, &pR, &uR[0], &uR[1], &uR[2], &uR[3]);
  inferred(P, t, qL, ePrev, &eL, &pL, &uL[0], &uL[1], &uL[2], &uL[3]);
  double rhoL = fabs(uL[1 + dir] / uL[0]);
  double rhoR = fabs(uR[1 + dir] / uR[0]);
  double a = fmax(rhoL, rhoR);
  for (int n = 0; n < nv; ++n) {
    double FqR = uR[1 + dir] * qR[n] / uR[0];
    double FqL = uL[1 + dir] * qL[n] / uL[0];
    double res = FqR + FqL - a * (qR[n] - qL[n]);
    res /= 2;
    H[n] = res;
  }
}

void kto_flux(const kto_params *P, int dir, const double *data, double t,
              double ePrev, int fwd, double *H) {
  flux(P, P->bulk ? 15 : 14, dir, data, t, ePrev, fwd, H);
}

/* --------------------------------------------------------- source terms */

/* setPimunuSourceTerms, SourceTerms.cu:30-256 */
static void pimunu_sources(const kto_params *P, double *rhs, double t,
                           double e, double p, double ut, double ux, double uy,
                           double un, double utp, double uxp, double uyp,
                           double unp, const double *pi, double Pi,
                           const double du[3][4], double dkvk) {
  const double dt = P->dt;
  double pitt = pi[0], pitx = pi[1], pity = pi[2], pitn = pi[3], pixx = pi[4],
         pixy = pi[5], pixn = pi[6], piyy = pi[7], piyn = pi[8], pinn = pi[9];
  double dxut = du[0][0], dxux = du[0][1], dxuy = du[0][2], dxun = du[0][3];
  double dyut = du[1][0], dyux = du[1][1], dyuy = du[1][2], dyun = du[1][3];
  double dnut = du[2][0], dnux = du[2][1], dnuy = du[2][2], dnun = du[2][3];

  /* shear and bulk coefficients, :46-60 */
  double T = eos_T(P, e);
  double taupiInv = F_0P2 * (T / P->etabar);
  double beta_pi = F_0P2 * (e + p);
  double cs2 = eos_cs2(P, e);
  double a = F_THIRD - cs2;
  double a2 = a * a;
  double beta_Pi = 15.0 * a2 * (e + p);
  double lambda_Pipi = F_1P6 * a;
  double zetabar = kto_zeta_over_s(P, T);
  double tauPiInv = 15 * a2 * (T / zetabar);

  double ut2 = ut * ut, un2 = un * un, t2 = t * t, t3 = t * t2;

  /* backward-difference time derivatives, :67-71 */
  double dtut = (ut - utp) / dt;
  double dtux = (ux - uxp) / dt;
  double dtuy = (uy - uyp) / dt;
  double dtun = (un - unp) / dt;

  /* acceleration with lowered spatial indices, :76-86 */
  double Dut = ut * dtut + ux * dxut + uy * dyut + un * dnut + t * un * un;
  double dut = Dut - t * un * un;
  double dux = ut * dtux + ux * dxux + uy * dyux + un * dnux;
  double Dux = -dux;
  double duy = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
  double Duy = -duy;
  double dun = ut * dtun + ux * dxun + uy * dyun + un * dnun;
  double Dun = -t2 * dun - 2.0 * t * ut * un;

  double theta = ut / t + dtut + dxux + dyuy + dnun; /* :91 */

  /* shear tensor sigma^{mu nu}, :96-106 */
  double theta3 = theta / 3.0;
  double stt = -t * ut * un2 + (dtut - ut * dut) + (ut2 - 1.0) * theta3;
  double stx = -(t * un2 * ux) / 2.0 + (dtux - dxut) / 2.0 -
               (ux * dut + ut * dux) / 2.0 + ut * ux * theta3;
  double sty = -(t * un2 * uy) / 2.0 + (dtuy - dyut) / 2.0 -
               (uy * dut + ut * duy) / 2.0 + ut * uy * theta3;
  double stn = -un * (2.0 * ut2 + t2 * un2) / (2.0 * t) +
               (dtun - dnut / t2) / 2.0 - (un * dut + ut * dun) / 2.0 +
               ut * un * theta3;
  double sxx = -(dxux + ux * dux) + (1.0 + ux * ux) * theta3;
  double sxy = -(dxuy + dyux) / 2.0 - (uy * dux + ux * duy) / 2.0 +
               ux * uy * theta3;
  double sxn = -ut * ux * un / t - (dxun + dnux / t2) / 2.0 -
               (un * dux + ux * dun) / 2.0 + ux * un * theta3;
  double syy = -(dyuy + uy * duy) + (1.0 + uy * uy) * theta3;
  double syn = -ut * uy * un / t - (dyun + dnuy / t2) / 2.0 -
               (un * duy + uy * dun) / 2.0 + uy * un * theta3;
  double snn = -ut * (1.0 + 2.0 * t2 * un2) / t3 - dnun / t2 - un * dun +
               (1.0 / t2 + un2) * theta3;

  /* vorticity omega^mu_nu (McNelis-Du signs), :123-136 */
  double wtx = (-dtux - dxut + ut * dux - ux * dut - t * un2 * ux) / 2.0;
  double wty = (-dtuy - dyut + ut * duy - uy * dut - t * un2 * uy) / 2.0;
  double wtn = (-t2 * dtun - 2.0 * t * un - dnut - ut * Dun - t2 * un * dut -
                t3 * un * un2) / 2.0;
  double wxy = (dxuy - dyux + ux * duy - uy * dux) / 2.0;
  double wxn = (t2 * dxun - dnux - ux * Dun - t2 * un * dux) / 2.0;
  double wyn = (t2 * dyun - dnuy - uy * Dun - t2 * un * duy) / 2.0;
  double wxt = wtx, wyt = wty, wnt = wtn / t2, wyx = -wxy, wnx = -wxn / t2,
         wny = -wyn / t2;

  /* I1 = (u^mu pi^{nu l} + u^nu pi^{mu l}) D u_l, :141-150 */
  double I1tt = 2.0 * ut * (pitt * Dut + pitx * Dux + pity * Duy + pitn * Dun);
  double I1tx = (pitt * ux + pitx * ut) * Dut + (pitx * ux + pixx * ut) * Dux +
                (pity * ux + pixy * ut) * Duy + (pitn * ux + pixn * ut) * Dun;
  double I1ty = (pitt * uy + pity * ut) * Dut + (pitx * uy + pixy * ut) * Dux +
                (pity * uy + piyy * ut) * Duy + (pitn * uy + piyn * ut) * Dun;
  double I1tn = (pitt * un + pitn * ut) * Dut + (pitx * un + pixn * ut) * Dux +
                (pity * un + piyn * ut) * Duy + (pitn * un + pinn * ut) * Dun;
  double I1xx = 2.0 * ux * (pitx * Dut + pixx * Dux + pixy * Duy + pixn * Dun);
  double I1xy = (pitx * uy + pity * ux) * Dut + (pixx * uy + pixy * ux) * Dux +
                (pixy * uy + piyy * ux) * Duy + (pixn * uy + piyn * ux) * Dun;
  double I1xn = (pitx * un + pitn * ux) * Dut + (pixx * un + pixn * ux) * Dux +
                (pixy * un + piyn * ux) * Duy + (pixn * un + pinn * ux) * Dun;
  double I1yy = 2.0 * uy * (pity * Dut + pixy * Dux + piyy * Duy + piyn * Dun);
  double I1yn = (pity * un + pitn * uy) * Dut + (pixy * un + pixn * uy) * Dux +
                (piyy * un + piyn * uy) * Duy + (piyn * un + pinn * uy) * Dun;
  double I1nn = 2.0 * un * (pitn * Dut + pixn * Dux + piyn * Duy + pinn * Dun);

  /* I2 = theta pi, :155-164 */
  double I2tt = theta * pitt, I2tx = theta * pitx, I2ty = theta * pity,
         I2tn = theta * pitn, I2xx = theta * pixx, I2xy = theta * pixy,
         I2xn = theta * pixn, I2yy = theta * piyy, I2yn = theta * piyn,
         I2nn = theta * pinn;

  /* I3 = pi^{mu l} w^nu_l + pi^{nu l} w^mu_l, :169-178 */
  double I3tt = 2.0 * (pitx * wtx + pity * wty + pitn * wtn);
  double I3tx = pitt * wxt + pity * wxy + pitn * wxn + pixx * wtx + pixy * wty +
                pixn * wtn;
  double I3ty = pitt * wyt + pitx * wyx + pitn * wyn + pixy * wtx + piyy * wty +
                piyn * wtn;
  double I3tn = pitt * wnt + pitx * wnx + pity * wny + pixn * wtx + piyn * wty +
                pinn * wtn;
  double I3xx = 2.0 * (pitx * wxt + pixy * wxy + pixn * wxn);
  double I3xy = pitx * wyt + pity * wxt + pixx * wyx + piyy * wxy + pixn * wyn +
                piyn * wxn;
  double I3xn = pitx * wnt + pitn * wxt + pixx * wnx + pixy * wny + piyn * wxy +
                pinn * wxn;
  double I3yy = 2.0 * (pity * wyt + pixy * wyx + piyn * wyn);
  double I3yn = pity * wnt + pitn * wyt + pixy * wnx + pixn * wyx + piyy * wny +
                pinn * wyn;
  double I3nn = 2.0 * (pitn * wnt + pixn * wnx + piyn * wny);

  /* I4 = pi^{l<mu} sigma^{nu>}_l, :183-203 */
  double ux2 = ux * ux, uy2 = uy * uy;
  double ps = pitt * stt - 2.0 * pitx * stx - 2.0 * pity * sty + pixx * sxx +
              2.0 * pixy * sxy + piyy * syy - 2.0 * pitn * stn * t2 +
              2.0 * pixn * sxn * t2 + 2.0 * piyn * syn * t2 +
              pinn * snn * t2 * t2;
  double ps3 = ps / 3.0;
  double I4tt = (pitt * stt - pitx * stx - pity * sty - t2 * pitn * stn) -
                (1 - ut2) * ps3;
  double I4tx = (pitt * stx + pitx * stt) / 2.0 - (pitx * sxx + pixx * stx) / 2.0 -
                (pity * sxy + pixy * sty) / 2.0 -
                t2 * (pitn * sxn + pixn * stn) / 2.0 + (ut * ux) * ps3;
  double I4ty = (pitt * sty + pity * stt) / 2.0 - (pitx * sxy + pixy * stx) / 2.0 -
                (pity * syy + piyy * sty) / 2.0 -
                t2 * (pitn * syn + piyn * stn) / 2.0 + (ut * uy) * ps3;
  double I4tn = (pitt * stn + pitn * stt) / 2.0 - (pitx * sxn + pixn * stx) / 2.0 -
                (pity * syn + piyn * sty) / 2.0 -
                t2 * (pitn * snn + pinn * stn) / 2.0 + (ut * un) * ps3;
  double I4xx = (pitx * stx - pixx * sxx - pixy * sxy - t2 * pixn * sxn) +
                (1.0 + ux2) * ps3;
  double I4xy = (pitx * sty + pity * stx) / 2.0 - (pixx * sxy + pixy * sxx) / 2.0 -
                (pixy * syy + piyy * sxy) / 2.0 -
                t2 * (pixn * syn + piyn * sxn) / 2.0 + (ux * uy) * ps3;
  double I4xn = (pitx * stn + pitn * stx) / 2.0 - (pixx * sxn + pixn * sxx) / 2.0 -
                (pixy * syn + piyn * sxy) / 2.0 -
                t2 * (pixn * snn + pinn * sxn) / 2.0 + (ux * un) * ps3;
  double I4yy = (pity * sty - pixy * sxy - piyy * syy - t2 * piyn * syn) +
                (1.0 + uy2) * ps3;
  double I4yn = (pity * stn + pitn * sty) / 2.0 - (pixy * sxn + pixn * sxy) / 2.0 -
                (piyy * syn + piyn * syy) / 2.0 -
                t2 * (piyn * snn + pinn * syn) / 2.0 + (uy * un) * ps3;
  double I4nn = (pitn * stn - pixn * sxn - piyn * syn - t2 * pinn * snn) +
                (1.0 / t2 + un2) * ps3;

  /* I = I1 + d I2 - I3 + t I4 - l Pi sigma, :209-218 */
  double Itt = I1tt + delta_pipi * I2tt - I3tt + tau_pipi * I4tt - lambda_piPi * Pi * stt;
  double Itx = I1tx + delta_pipi * I2tx - I3tx + tau_pipi * I4tx - lambda_piPi * Pi * stx;
  double Ity = I1ty + delta_pipi * I2ty - I3ty + tau_pipi * I4ty - lambda_piPi * Pi * sty;
  double Itn = I1tn + delta_pipi * I2tn - I3tn + tau_pipi * I4tn - lambda_piPi * Pi * stn;
  double Ixx = I1xx + delta_pipi * I2xx - I3xx + tau_pipi * I4xx - lambda_piPi * Pi * sxx;
  double Ixy = I1xy + delta_pipi * I2xy - I3xy + tau_pipi * I4xy - lambda_piPi * Pi * sxy;
  double Ixn = I1xn + delta_pipi * I2xn - I3xn + tau_pipi * I4xn - lambda_piPi * Pi * sxn;
  double Iyy = I1yy + delta_pipi * I2yy - I3yy + tau_pipi * I4yy - lambda_piPi * Pi * syy;
  double Iyn = I1yn + delta_pipi * I2yn - I3yn + tau_pipi * I4yn - lambda_piPi * Pi * syn;
  double Inn = I1nn + delta_pipi * I2nn - I3nn + tau_pipi * I4nn - lambda_piPi * Pi * snn;

  /* relaxation right-hand sides with Milne Christoffel terms, :224-233 */
  double dpitt = 2.0 * beta_pi * stt - pitt * taupiInv - Itt - 2.0 * un * t * pitn;
  double dpitx = 2.0 * beta_pi * stx - pitx * taupiInv - Itx - un * t * pixn;
  double dpity = 2.0 * beta_pi * sty - pity * taupiInv - Ity - un * t * piyn;
  double dpitn = 2.0 * beta_pi * stn - pitn * taupiInv - Itn - un * t * pinn -
                 (ut * pitn + un * pitt) / t;
  double dpixx = 2.0 * beta_pi * sxx - pixx * taupiInv - Ixx;
  double dpixy = 2.0 * beta_pi * sxy - pixy * taupiInv - Ixy;
  double dpixn = 2.0 * beta_pi * sxn - pixn * taupiInv - Ixn - (ut * pixn + un * pitx) / t;
  double dpiyy = 2.0 * beta_pi * syy - piyy * taupiInv - Iyy;
  double dpiyn = 2.0 * beta_pi * syn - piyn * taupiInv - Iyn - (ut * piyn + un * pity) / t;
  double dpinn = 2.0 * beta_pi * snn - pinn * taupiInv - Inn - 2.0 * (ut * pinn + un * pitn) / t;

  /* bulk, :238 */
  double dPi = -beta_Pi * theta - Pi * tauPiInv - delta_PiPi * Pi * theta + lambda_Pipi * ps;

  /* :243-255 */
  rhs[0] = dpitt / ut + pitt * dkvk;
  rhs[1] = dpitx / ut + pitx * dkvk;
  rhs[2] = dpity / ut + pity * dkvk;
  rhs[3] = dpitn / ut + pitn * dkvk;
  rhs[4] = dpixx / ut + pixx * dkvk;
  rhs[5] = dpixy / ut + pixy * dkvk;
  rhs[6] = dpixn / ut + pixn * dkvk;
  rhs[7] = dpiyy / ut + piyy * dkvk;
  rhs[8] = dpiyn / ut + piyn * dkvk;
  rhs[9] = dpinn / ut + pinn * dkvk;
  rhs[10] = P->bulk ? (dPi / ut + Pi * dkvk) : 0.0;
}

/* ------------------------------------------------------- Euler kernels */

static inline long idx(const kto_ctx *c, int i, int j, int k) {
  return i + (long)c->ncx * (j + (long)c->ncy * k); /* DynamicalVariables.cu:33-35 */
}

/* loadSourceTerms2, SourceTerms.cu:570-755, for one cell given its stencil
 * derivatives: du[d][m] = d_d u^m and dp[d] = d_d p (d = x, y, eta; the
 * USE_APPROX_DERIVATIVE minmod derivatives divided by the spacing).
 * S[0..3]: T^{tau mu} geometric and pressure sources; S[4..14]: pi, Pi RHS. */
void kto_source_cell(const kto_params *P, double t, const double *Q, double e,
                     double ps, const double *u, const double *up,
                     const double du[3][4], const double *dp, double *S) {
  const int nv = P->bulk ? 15 : 14;
  double Pi = P->bulk ? Q[14] : 0.0;
  double ut = u[0], ux = u[1], uy = u[2], un = u[3];
  /* T^{tau mu} sources, SourceTerms.cu:726-737 */
  double pinn = Q[13];
  double PP = ps + Pi;
  double tnn = (e + PP) * un * un + PP / t / t + pinn; /* Tnn, EnergyMomentumTensor.cu:205-207 */
  double vx = ux / ut, vy = uy / ut, vn = un / ut;
  double dxvx = (du[0][1] - vx * du[0][0]) / ut;
  double dyvy = (du[1][2] - vy * du[1][0]) / ut;
  double dnvn = (du[2][3] - vn * du[2][0]) / ut;
  double dkvk = dxvx + dyvy + dnvn;
  S[0] = -(Q[0] / t + t * tnn) + dkvk * (Q[4] - ps - Pi) - vx * dp[0] - vy * dp[1] - vn * dp[2];
  S[1] = -Q[1] / t - dp[0] + dkvk * Q[5];
  S[2] = -Q[2] / t - dp[1] + dkvk * Q[6];
  S[3] = -3 * Q[3] / t - dp[2] / (t * t) + dkvk * Q[7];
  double rhs[11];
  pimunu_sources(P, rhs, t, e, ps, ut, ux, uy, un, up[0], up[1], up[2], up[3], &Q[4], Pi, du,
                 dkvk);
  for (int n = 4; n < nv; ++n) S[n] = rhs[n - 4];
  if (nv == 14) S[14] = 0.0;
}

/* eulerStepKernelSource_1D, EulerStep.cu:947-1009: q' = q + dt S with the
 * u and p derivatives of SourceTerms.cu:662-719 (USE_APPROX_DERIVATIVE). */
static void euler_source(kto_ctx *c, double t, double **cur, double **upd,
                         double **u, double **up) {
  const kto_params *P = &c->P;
  const int nv = c->nv;
  const long sx = 1, sy = c->ncx, sz = (long)c->ncx * c->ncy;
  const double *p = c->p, *e = c->e;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = NG; k < P->nz + NG; ++k)
    for (int j = NG; j < P->ny + NG; ++j)
      for (int i = NG; i < P->nx + NG; ++i) {
        long s = idx(c, i, j, k);
        double Q[NVMAX], S[NVMAX];
        for (int n = 0; n < NVMAX; ++n) Q[n] = (n < nv) ? cur[n][s] : 0.0;
        const long st[3] = {sx, sy, sz};
        const double h[3] = {P->dx, P->dy, P->dz};
        double du[3][4], dp[3];
        for (int d = 0; d < 3; ++d) {
          for (int m = 0; m < 4; ++m)
            du[d][m] = kto_approx_deriv(u[m][s - st[d]], u[m][s], u[m][s + st[d]]) / h[d];
          dp[d] = kto_approx_deriv(p[s - st[d]], p[s], p[s + st[d]]) / h[d];
        }
        const double uc[4] = {u[0][s], u[1][s], u[2][s], u[3][s]};
        const double upc[4] = {up[0][s], up[1][s], up[2][s], up[3][s]};
        kto_source_cell(P, t, Q, e[s], p[s], uc, upc, (const double(*)[4])du, dp, S);
        for (int n = 0; n < nv; ++n) upd[n][s] = Q[n] + P->dt * S[n];
      }
}

/* eulerStepKernel{X,Y,Z}_1D + loadSourceTerms{X,Y,Z}, EulerStep.cu:1012-1277
 * and SourceTerms.cu:444-567 */
static void euler_dir(kto_ctx *c, int dir, double t, double **cur, double **upd,
                      double **u) {
  const kto_params *P = &c->P;
  const int nv = c->nv;
  const long st = dir == 0 ? 1 : (dir == 1 ? c->ncx : (long)c->ncx * c->ncy);
  const double h = dir == 0 ? P->dx : (dir == 1 ? P->dy : P->dz);
  const double *e = c->e;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = NG; k < P->nz + NG; ++k)
    for (int j = NG; j < P->ny + NG; ++j)
      for (int i = NG; i < P->nx + NG; ++i) {
        long s = idx(c, i, j, k);
        double I[5 * NVMAX] = {0}, H[NVMAX], r[NVMAX];
        for (int n = 0; n < nv; ++n)
          for (int m = 0; m < 5; ++m) I[5 * n + m] = cur[n][s + (m - 2) * st];
        flux(P, nv, dir, I, t, e[s], 1, H);
        for (int n = 0; n < nv; ++n) r[n] = -H[n];
        flux(P, nv, dir, I, t, e[s], 0, H);
        for (int n = 0; n < nv; ++n) {
          r[n] += H[n];
          r[n] /= h;
        }
        /* directional pi / Pi gradient sources: central differences */
        double fac = 1 / h / 2;
#define CD(n) ((I[5 * (n) + 3] - I[5 * (n) + 1]) * fac)
        double ut = u[0][s], udir = u[1 + dir][s];
        double v = udir / ut;
        double Sd[4];
        if (dir == 0) {
          double dxpitt = CD(4), dxpitx = CD(5), dxpity = CD(6), dxpitn = CD(7),
                 dxpixx = CD(8), dxpixy = CD(9), dxpixn = CD(10);
          if (P->bulk) {
            double dxPi = CD(14);
            Sd[0] = dxpitt * v - dxpitx - v * dxPi;
            Sd[1] = dxpitx * v - dxpixx - dxPi;
          } else {
            Sd[0] = dxpitt * v - dxpitx;
            Sd[1] = dxpitx * v - dxpixx;
          }
          Sd[2] = dxpity * v - dxpixy;
          Sd[3] = dxpitn * v - dxpixn;
        } else if (dir == 1) {
          double dypitt = CD(4), dypitx = CD(5), dypity = CD(6), dypitn = CD(7),
                 dypixy = CD(9), dypiyy = CD(11), dypiyn = CD(12);
          if (P->bulk) {
            double dyPi = CD(14);
            Sd[0] = dypitt * v - dypity - v * dyPi;
            Sd[2] = dypity * v - dypiyy - dyPi;
          } else {
            Sd[0] = dypitt * v - dypity;
            Sd[2] = dypity * v - dypiyy;
          }
          Sd[1] = dypitx * v - dypixy;
          Sd[3] = dypitn * v - dypiyn;
        } else {
          double dnpitt = CD(4), dnpitx = CD(5), dnpity = CD(6), dnpitn = CD(7),
                 dnpixn = CD(10), dnpiyn = CD(12), dnpinn = CD(13);
          if (P->bulk) {
            double dnPi = CD(14);
            Sd[0] = dnpitt * v - dnpitn - v * dnPi;
            Sd[3] = dnpitn * v - dnpinn - dnPi / (t * t);
          } else {
            Sd[0] = dnpitt * v - dnpitn;
            Sd[3] = dnpitn * v - dnpinn;
          }
          Sd[1] = dnpitx * v - dnpixn;
          Sd[2] = dnpity * v - dnpiyn;
        }
#undef CD
        for (int n = 0; n < 4; ++n) {
          r[n] += Sd[n];
          r[n] *= P->dt;
        }
        for (int n = 4; n < nv; ++n) r[n] *= P->dt;
        for (int n = 0; n < nv; ++n) upd[n][s] += r[n];
      }
}

/* eulerStep with EULER_STEP_SPLIT_1D, FullyDiscreteKurganovTadmorScheme.cu:35-58 */
static void euler_step(kto_ctx *c, double t, double **cur, double **upd,
                       double **u, double **up) {
  euler_source(c, t, cur, upd, u, up);
  euler_dir(c, 0, t, cur, upd, u);
  euler_dir(c, 1, t, cur, upd, u);
  euler_dir(c, 2, t, cur, upd, u);
}

/* setInferredVariablesKernel, EnergyMomentumTensor.cu:110-154 */
static void set_inferred(kto_ctx *c, double **q, double **u, double t) {
  const kto_params *P = &c->P;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = NG; k < P->nz + NG; ++k)
    for (int j = NG; j < P->ny + NG; ++j)
      for (int i = NG; i < P->nx + NG; ++i) {
        long s = idx(c, i, j, k);
        double qs[NVMAX] = {0};
        for (int n = 0; n < 8; ++n) qs[n] = q[n][s];
        if (P->bulk) qs[14] = q[14][s];
        double ee, pp, a, b, cc, d;
        inferred(P, t, qs, c->e[s], &ee, &pp, &a, &b, &cc, &d);
        c->e[s] = ee;
        c->p[s] = pp;
        u[0][s] = a;
        u[1][s] = b;
        u[2][s] = cc;
        u[3][s] = d;
      }
}

/* Shear regulation factor of one cell, RegulateDissipativeCurrents.cu:69-112
 * (xi0 = 0.1, rho_max = 1; the trace and u.pi ratios are signed, :88-92):
 * tanh(rho)/rho when e <= e_reg, else 1.  pi[10] = tt tx ty tn xx xy xn yy yn nn. */
double kto_regulate_factor(double t, const double *pi, double e_s, double p_s,
                           const double *u, double e_reg) {
  double pitt = pi[0], pitx = pi[1], pity = pi[2], pitn = pi[3], pixx = pi[4],
         pixy = pi[5], pixn = pi[6], piyy = pi[7], piyn = pi[8], pinn = pi[9];
  double ut = u[0], ux = u[1], uy = u[2], un = u[3];
  double xi0 = 0.1, rhomax = 1.0;
  double t2 = t * t;
  double pipi = pitt * pitt - 2 * pitx * pitx - 2 * pity * pity + pixx * pixx +
                2 * pixy * pixy + piyy * piyy - 2 * pitn * pitn * t2 +
                2 * pixn * pixn * t2 + 2 * piyn * piyn * t2 + pinn * pinn * t2 * t2;
  double spipi = sqrt(fabs(pipi));
  double pimumu = pitt - pixx - piyy - pinn * t * t;
  double piu0 = -pitn * t2 * un + pitt * ut - pitx * ux - pity * uy;
  double piu1 = -pixn * t2 * un + pitx * ut - pixx * ux - pixy * uy;
  double piu2 = -piyn * t2 * un + pity * ut - pixy * ux - piyy * uy;
  double piu3 = -pinn * t2 * un + pitn * ut - pixn * ux - piyn * uy;
  double a1 = (spipi / rhomax) * (1 / sqrt(e_s * e_s + 3 * p_s * p_s));
  double den = xi0 * rhomax * spipi;
  double a2 = pimumu / den, a3 = piu0 / den, a4 = piu1 / den, a5 = piu2 / den,
         a6 = piu3 / den;
  double rho = fmax(a1, fmax(a2, fmax(a3, fmax(a4, fmax(a5, a6)))));
  double facShear = tanh(rho) / rho;
  if (fabs(rho) < 1.e-7) facShear = 1.0;
  if (e_s > e_reg) facShear = 1.0;
  return facShear;
}

/* regulateDissipativeCurrents, RegulateDissipativeCurrents.cu:16-134.  The
 * Pi read is guarded by '#ifdef Pi' (never defined), so the bulk factor is
 * tanh(0)/0 -> replaced by 1 unless p is 0 or NaN, where it stays NaN. */
static void regulate(kto_ctx *c, double t, double **q, double **u) {
  const kto_params *P = &c->P;
  const double e_reg = kto_e_of_T(P, P->T_reg);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = NG; k < P->nz + NG; ++k)
    for (int j = NG; j < P->ny + NG; ++j)
      for (int i = NG; i < P->nx + NG; ++i) {
        long s = idx(c, i, j, k);
        double pi[10], uc[4];
        for (int m = 0; m < 10; ++m) pi[m] = q[4 + m][s];
        for (int m = 0; m < 4; ++m) uc[m] = u[m][s];
        double e_s = c->e[s], p_s = c->p[s];
        double facShear = kto_regulate_factor(t, pi, e_s, p_s, uc, e_reg);
        double rhoBulk = 0.0 / p_s;
        double facBulk = tanh(rhoBulk) / rhoBulk;
        if (fabs(rhoBulk) < 1.e-7) facBulk = 1.0;
        if (e_s > e_reg) facBulk = 1.0;
        for (int n = 4; n < 14; ++n) q[n][s] *= facShear;
        if (P->bulk) q[14][s] *= facBulk;
        c->reg[s] = facShear;
      }
}

/* ------------------------------------------------ validity diagnostics */

/* checkValidityKernel, HydrodynamicValidity.cu:29-227, for one cell: e, p,
 * u, u_prev of the cell, du[d][m] = d_d u^m the central differences of
 * :50-68 (d = x, y, eta), pi[10] = tt tx ty tn xx xy xn yy yn nn, Pi the
 * bulk pressure (0 without bulk).  The vorticity is the older convention
 * this kernel keeps (:116-121), not the one of the source terms. */
void kto_validity_cell(const kto_params *P, double t, double e_s, double p_s,
                       const double *u, const double *up, const double du[3][4],
                       const double *pi, double Pi, double *out) {
  const double etabar = P->etabar;
  double ut = u[0], ux = u[1], uy = u[2], un = u[3];
  double utp = up[0], uxp = up[1], uyp = up[2], unp = up[3];
  double dxut = du[0][0], dxux = du[0][1], dxuy = du[0][2], dxun = du[0][3];
  double dyut = du[1][0], dyux = du[1][1], dyuy = du[1][2], dyun = du[1][3];
  double dnut = du[2][0], dnux = du[2][1], dnuy = du[2][2], dnun = du[2][3];
  double ut2 = ut * ut, un2 = un * un, t2 = t * t, t3 = t * t2;

  /* :76-83 */
  double T = eos_T(P, e_s);
  double cs2 = eos_cs2(P, e_s);
  double a = F_THIRD - cs2;
  double a2 = a * a;
  double lambda_Pipi = F_1P6 * a;
  double zetabar = kto_zeta_over_s(P, T);
  double tauPiInv = 15 * a2 * (T / zetabar);

  /* :86-100 */
  double dtut = (ut - utp) / P->dt, dtux = (ux - uxp) / P->dt;
  double dtuy = (uy - uyp) / P->dt, dtun = (un - unp) / P->dt;
  double Dut = ut * dtut + ux * dxut + uy * dyut + un * dnut + t * un * un;
  double dut = Dut - t * un * un;
  double dux = ut * dtux + ux * dxux + uy * dyux + un * dnux;
  double duy = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
  double dun = ut * dtun + ux * dxun + uy * dyun + un * dnun;
  double Dun = -t2 * dun - 2 * t * ut * un;
  double theta = ut / t + dtut + dxux + dyuy + dnun;

  /* shear tensor sigma^{mu nu}, :103-113 */
  double theta3 = theta / 3;
  double stt = -t * ut * un2 + (dtut - ut * dut) + (ut2 - 1) * theta3;
  double stx = -(t * un2 * ux) / 2 + (dtux - dxut) / 2 - (ux * dut + ut * dux) / 2 + ut * ux * theta3;
  double sty = -(t * un2 * uy) / 2 + (dtuy - dyut) / 2 - (uy * dut + ut * duy) / 2 + ut * uy * theta3;
  double stn = -un * (2 * ut2 + t2 * un2) / (2 * t) + (dtun - dnut / t2) / 2 -
               (un * dut + ut * dun) / 2 + ut * un * theta3;
  double sxx = -(dxux + ux * dux) + (1 + ux * ux) * theta3;
  double sxy = -(dxuy + dyux) / 2 - (uy * dux + ux * duy) / 2 + ux * uy * theta3;
  double sxn = -ut * ux * un / t - (dxun + dnux / t2) / 2 - (un * dux + ux * dun) / 2 + ux * un * theta3;
  double syy = -(dyuy + uy * duy) + (1 + uy * uy) * theta3;
  double syn = -ut * uy * un / t - (dyun + dnuy / t2) / 2 - (un * duy + uy * dun) / 2 + uy * un * theta3;
  double snn = -ut * (1 + 2 * t2 * un2) / t3 - dnun / t2 - un * dun + (1 / t2 + un2) * theta3;

  /* vorticity, :116-128 */
  double wtx = (dtux + dxut) / 2 + (ux * dut - ut * dux) / 2 + t * un2 * ux / 2;
  double wty = (dtuy + dyut) / 2 + (uy * dut - ut * duy) / 2 + t * un2 * uy / 2;
  double wtn = (t2 * dtun + 2 * t * un + dnut) / 2 + (t2 * un * dut - ut * Dun) + t3 * un * un2 / 2;
  double wxy = (dyux - dxuy) / 2 + (uy * dux - ux * duy) / 2;
  double wxn = (dnux - t2 * dxun) / 2 + (t2 * un * dux - ux * Dun) / 2;
  double wyn = (dnuy - t2 * dyun) / 2 + (t2 * un * duy - uy * Dun) / 2;
  double wxt = wtx, wyt = wty, wnt = wtn / t2, wyx = -wxy, wnx = -wxn / t2, wny = -wyn / t2;

  double pitt = pi[0], pitx = pi[1], pity = pi[2], pitn = pi[3], pixx = pi[4];
  double pixy = pi[5], pixn = pi[6], piyy = pi[7], piyn = pi[8], pinn = pi[9];

  /* I2, I3, I4, :148-192 */
  double I2tt = theta * pitt, I2tx = theta * pitx, I2ty = theta * pity, I2tn = theta * pitn;
  double I2xx = theta * pixx, I2xy = theta * pixy, I2xn = theta * pixn, I2yy = theta * piyy;
  double I2yn = theta * piyn, I2nn = theta * pinn;
  double I3tt = 2 * (pitx * wtx + pity * wty + pitn * wtn);
  double I3tx = pitt * wxt + pity * wxy + pitn * wxn + pixx * wtx + pixy * wty + pixn * wtn;
  double I3ty = pitt * wyt + pitx * wyx + pitn * wyn + pixy * wtx + piyy * wty + piyn * wtn;
  double I3tn = pitt * wnt + pitx * wnx + pity * wny + pixn * wtx + piyn * wty + pinn * wtn;
  double I3xx = 2 * (pitx * wxt + pixy * wxy + pixn * wxn);
  double I3xy = pitx * wyt + pity * wxt + pixx * wyx + piyy * wxy + pixn * wyn + piyn * wxn;
  double I3xn = pitx * wnt + pitn * wxt + pixx * wnx + pixy * wny + piyn * wxy + pinn * wxn;
  double I3yy = 2 * (pity * wyt + pixy * wyx + piyn * wyn);
  double I3yn = pity * wnt + pitn * wyt + pixy * wnx + pixn * wyx + piyy * wny + pinn * wyn;
  double I3nn = 2 * (pitn * wnt + pixn * wnx + piyn * wny);
  double ux2 = ux * ux, uy2 = uy * uy;
  double ps = pitt * stt - 2 * pitx * stx - 2 * pity * sty + pixx * sxx + 2 * pixy * sxy +
              piyy * syy - 2 * pitn * stn * t2 + 2 * pixn * sxn * t2 + 2 * piyn * syn * t2 +
              pinn * snn * t2 * t2;
  double ps3 = ps / 3;
  double I4tt = (pitt * stt - pitx * stx - pity * sty - t2 * pitn * stn) - (1 - ut2) * ps3;
  double I4tx = (pitt * stx + pitx * stt) / 2 - (pitx * sxx + pixx * stx) / 2 -
                (pity * sxy + pixy * sty) / 2 - t2 * (pitn * sxn + pixn * stn) / 2 + (ut * ux) * ps3;
  double I4ty = (pitt * sty + pity * stt) / 2 - (pitx * sxy + pixy * stx) / 2 -
                (pity * syy + piyy * sty) / 2 - t2 * (pitn * syn + piyn * stn) / 2 + (ut * uy) * ps3;
  double I4tn = (pitt * stn + pitn * stt) / 2 - (pitx * sxn + pixn * stx) / 2 -
                (pity * syn + piyn * sty) / 2 - t2 * (pitn * snn + pinn * stn) / 2 + (ut * un) * ps3;
  double I4xx = (pitx * stx - pixx * sxx - pixy * sxy - t2 * pixn * sxn) + (1 + ux2) * ps3;
  double I4xy = (pitx * sty + pity * stx) / 2 - (pixx * sxy + pixy * sxx) / 2 -
                (pixy * syy + piyy * sxy) / 2 - t2 * (pixn * syn + piyn * sxn) / 2 + (ux * uy) * ps3;
  double I4xn = (pitx * stn + pitn * stx) / 2 - (pixx * sxn + pixn * sxx) / 2 -
                (pixy * syn + piyn * sxy) / 2 - t2 * (pixn * snn + pinn * sxn) / 2 + (ux * un) * ps3;
  double I4yy = (pity * sty - pixy * sxy - piyy * syy - t2 * piyn * syn) + (1 + uy2) * ps3;
  double I4yn = (pity * stn + pitn * sty) / 2 - (pixy * sxn + pixn * sxy) / 2 -
                (piyy * syn + piyn * syy) / 2 - t2 * (piyn * snn + pinn * syn) / 2 + (uy * un) * ps3;
  double I4nn = (pitn * stn - pixn * sxn - piyn * syn - t2 * pinn * snn) + (1 / t2 + un2) * ps3;

  /* :194-212 */
  double pipi = pitt * pitt - 2 * pitx * pitx - 2 * pity * pity + pixx * pixx + 2 * pixy * pixy +
                piyy * piyy - 2 * pitn * pitn * t2 + 2 * pixn * pixn * t2 + 2 * piyn * piyn * t2 +
                pinn * pinn * t2 * t2;
  double ss = stt * stt - 2 * stx * stx - 2 * sty * sty + sxx * sxx + 2 * sxy * sxy + syy * syy -
              2 * stn * stn * t2 + 2 * sxn * sxn * t2 + 2 * syn * syn * t2 + snn * snn * t2 * t2;
  double Jtt = delta_pipi * I2tt - I3tt + tau_pipi * I4tt - lambda_piPi * Pi * stt;
  double Jtx = delta_pipi * I2tx - I3tx + tau_pipi * I4tx - lambda_piPi * Pi * stx;
  double Jty = delta_pipi * I2ty - I3ty + tau_pipi * I4ty - lambda_piPi * Pi * sty;
  double Jtn = delta_pipi * I2tn - I3tn + tau_pipi * I4tn - lambda_piPi * Pi * stn;
  double Jxx = delta_pipi * I2xx - I3xx + tau_pipi * I4xx - lambda_piPi * Pi * sxx;
  double Jxy = delta_pipi * I2xy - I3xy + tau_pipi * I4xy - lambda_piPi * Pi * sxy;
  double Jxn = delta_pipi * I2xn - I3xn + tau_pipi * I4xn - lambda_piPi * Pi * sxn;
  double Jyy = delta_pipi * I2yy - I3yy + tau_pipi * I4yy - lambda_piPi * Pi * syy;
  double Jyn = delta_pipi * I2yn - I3yn + tau_pipi * I4yn - lambda_piPi * Pi * syn;
  double Jnn = delta_pipi * I2nn - I3nn + tau_pipi * I4nn - lambda_piPi * Pi * snn;
  double J = -delta_PiPi * Pi * theta + lambda_Pipi * ps;
  double JJ = Jtt * Jtt - 2 * Jtx * Jtx - 2 * Jty * Jty + Jxx * Jxx + 2 * Jxy * Jxy + Jyy * Jyy -
              2 * Jtn * Jtn * t2 + 2 * Jxn * Jxn * t2 + 2 * Jyn * Jyn * t2 + Jnn * Jnn * t2 * t2;

  /* :214-227 */
  out[KTO_V_RE_PI] = sqrt(fabs(pipi)) / p_s;
  out[KTO_V_RE2_PI] = T / 2 / etabar / (e_s + p_s) * sqrt(fabs(JJ / ss));
  out[KTO_V_RE_BULK] = fabs(Pi) / p_s;
  out[KTO_V_RE2_BULK] = fabs(J / zetabar / theta) * T / (e_s + p_s);
  out[KTO_V_KN_TAUPI] = 5 * etabar * theta / T;
  out[KTO_V_KN_TAUBULK] = theta / tauPiInv;
  out[KTO_V_TAUPI] = 5 * etabar / T;
  out[KTO_V_DXUX] = dxux;
  out[KTO_V_DYUY] = dyuy;
  out[KTO_V_THETA] = theta;
}

/* checkValidity (HydrodynamicValidity.cu:231-235) over the interior.  The
 * reference calls it at the end of twoStepRungeKutta with d_q, which the run
 * loop swaps only afterwards (FullyDiscreteKurganovTadmorScheme.cu:137,
 * HydroPlugin.cpp:422): pi and Pi are then those of q^n, the buffer that is
 * Q after kto_step.  prev_q selects that buffer; prev_q == 0 reads the
 * current q (the call before the first step, HydroPlugin.cpp:222).  out
 * holds KTO_V_N ghosted arrays; interior cells are written, the last field
 * with the regulation factors of the last regulation. */
void kto_validity(void *h, double t, int prev_q, double *out) {
  kto_ctx *c = (kto_ctx *)h;
  const kto_params *P = &c->P;
  double **q = prev_q ? c->Q : c->q;
  const double fx = 1 / P->dx / 2, fy = 1 / P->dy / 2, fz = 1 / P->dz / 2;
  const long sy = c->ncx, sz = (long)c->ncx * c->ncy;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = NG; k < P->nz + NG; ++k)
    for (int j = NG; j < P->ny + NG; ++j)
      for (int i = NG; i < P->nx + NG; ++i) {
        long s = idx(c, i, j, k);
        double uc[4], upc[4], du[3][4], pi[10], v[KTO_V_N];
        for (int m = 0; m < 4; ++m) {
          const double *um = c->u[m];
          uc[m] = um[s];
          upc[m] = c->up[m][s];
          du[0][m] = (um[s + 1] - um[s - 1]) * fx;
          du[1][m] = (um[s + sy] - um[s - sy]) * fy;
          du[2][m] = (um[s + sz] - um[s - sz]) * fz;
        }
        for (int m = 0; m < 10; ++m) pi[m] = q[4 + m][s];
        double Pi = P->bulk ? q[14][s] : 0.0;
        kto_validity_cell(P, t, c->e[s], c->p[s], uc, upc, du, pi, Pi, v);
        for (int f = 0; f < KTO_V_REG; ++f) out[f * c->G + s] = v[f];
        out[KTO_V_REG * c->G + s] = c->reg[s];
      }
}

/* setGhostCells + kernels I/J/K, GhostCells.cu:15-144: two layers per face
 * copy the nearest interior plane; edges and corners are left untouched. */
static void ghost_copy(kto_ctx *c, long s, long sBC, double **q, double **u) {
  c->e[s] = c->e[sBC];
  c->p[s] = c->p[sBC];
  for (int m = 0; m < 4; ++m) u[m][s] = u[m][sBC];
  for (int n = 0; n < c->nv; ++n) q[n][s] = q[n][sBC];
}

static void set_ghost_cells(kto_ctx *c, double **q, double **u) {
  const int nx = c->P.nx, ny = c->P.ny, nz = c->P.nz;
  for (int k = NG; k < nz + NG; ++k)
    for (int j = NG; j < ny + NG; ++j) {
      for (int i = 0; i <= 1; ++i) ghost_copy(c, idx(c, i, j, k), idx(c, 2, j, k), q, u);
      for (int i = nx + 2; i <= nx + 3; ++i)
        ghost_copy(c, idx(c, i, j, k), idx(c, nx + 1, j, k), q, u);
    }
  for (int k = NG; k < nz + NG; ++k)
    for (int i = NG; i < nx + NG; ++i) {
      for (int j = 0; j <= 1; ++j) ghost_copy(c, idx(c, i, j, k), idx(c, i, 2, k), q, u);
      for (int j = ny + 2; j <= ny + 3; ++j)
        ghost_copy(c, idx(c, i, j, k), idx(c, i, ny + 1, k), q, u);
    }
  for (int j = NG; j < ny + NG; ++j)
    for (int i = NG; i < nx + NG; ++i) {
      for (int k = 0; k <= 1; ++k) ghost_copy(c, idx(c, i, j, k), idx(c, i, j, 2), q, u);
      for (int k = nz + 2; k <= nz + 3; ++k)
        ghost_copy(c, idx(c, i, j, k), idx(c, i, j, nz + 1), q, u);
    }
}

/* convexCombinationEulerStepKernel, FullyDiscreteKurganovTadmorScheme.cu:60-104 */
static void convex(kto_ctx *c, double **q, double **Q) {
  const kto_params *P = &c->P;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = NG; k < P->nz + NG; ++k)
    for (int j = NG; j < P->ny + NG; ++j)
      for (int i = NG; i < P->nx + NG; ++i) {
        long s = idx(c, i, j, k);
        for (int n = 0; n < c->nv; ++n) {
          Q[n][s] += q[n][s];
          Q[n][s] /= 2;
        }
      }
}

/* twoStepRungeKutta, FullyDiscreteKurganovTadmorScheme.cu:109-139, followed
 * by setCurrentConservedVariables (HydroPlugin.cpp:422).  The leading
 * regulation of the stale qS buffer (:112) only rescales values that the
 * predictor then overwrites, so it is left out. */
void kto_step(void *h, double t) {
  kto_ctx *c = (kto_ctx *)h;
  const double dt = c->P.dt;
  euler_step(c, t, c->q, c->qS, c->u, c->up);
  t += dt;
  set_inferred(c, c->qS, c->uS, t);
  regulate(c, t, c->qS, c->uS);
  set_ghost_cells(c, c->qS, c->uS);

  euler_step(c, t, c->qS, c->Q, c->uS, c->u);
  convex(c, c->q, c->Q);
  double **tmp = c->up; /* swapFluidVelocity(&d_up, &d_u) */
  c->up = c->u;
  c->u = tmp;
  set_inferred(c, c->Q, c->u, t);
  regulate(c, t, c->Q, c->u);
  set_ghost_cells(c, c->Q, c->u);

  double **tq = c->q; /* setCurrentConservedVariables */
  c->q = c->Q;
  c->Q = tq;
}

void kto_steps(void *h, double t0, int n_begin, int nsteps) {
  kto_ctx *c = (kto_ctx *)h;
  for (int n = n_begin; n < n_begin + nsteps; ++n) kto_step(h, t0 + n * c->P.dt);
}

void kto_euler_only(void *h, double t) {
  kto_ctx *c = (kto_ctx *)h;
  euler_step(c, t, c->q, c->qS, c->u, c->up);
}

/* ------------------------------------------------------------- set-up */

void *kto_create(const kto_params *P) {
  kto_ctx *c = (kto_ctx *)calloc(1, sizeof(kto_ctx));
  c->P = *P;
  c->ncx = P->nx + 4;
  c->ncy = P->ny + 4;
  c->ncz = P->nz + 4;
  c->G = (long)c->ncx * c->ncy * c->ncz;
  c->nv = P->bulk ? 15 : 14;
  for (int n = 0; n < NVMAX; ++n) {
    c->qa[n] = (double *)calloc(c->G, sizeof(double));
    c->qb[n] = (double *)calloc(c->G, sizeof(double));
    c->qs[n] = (double *)calloc(c->G, sizeof(double));
  }
  for (int m = 0; m < 4; ++m) {
    c->ua[m] = (double *)calloc(c->G, sizeof(double));
    c->ub[m] = (double *)calloc(c->G, sizeof(double));
    c->uc[m] = (double *)calloc(c->G, sizeof(double));
  }
  c->e = (double *)calloc(c->G, sizeof(double));
  c->p = (double *)calloc(c->G, sizeof(double));
  c->reg = (double *)malloc(sizeof(double) * c->G);
  for (long s = 0; s < c->G; ++s) c->reg[s] = 1.0; /* DynamicalVariables.cu:81 */
  c->q = c->qa;
  c->Q = c->qb;
  c->qS = c->qs;
  c->u = c->ua;
  c->up = c->ub;
  c->uS = c->uc;
  return c;
}

void kto_destroy(void *h) {
  kto_ctx *c = (kto_ctx *)h;
  if (!c) return;
  for (int n = 0; n < NVMAX; ++n) {
    free(c->qa[n]);
    free(c->qb[n]);
    free(c->qs[n]);
  }
  for (int m = 0; m < 4; ++m) {
    free(c->ua[m]);
    free(c->ub[m]);
    free(c->uc[m]);
  }
  free(c->e);
  free(c->p);
  free(c->reg);
  free(c);
}

long kto_ghosted_size(void *h) { return ((kto_ctx *)h)->G; }

/* setConservedVariables, DynamicalVariables.cu:349-392, with T^{tau mu}
 * from Ttt..Ttn (EnergyMomentumTensor.cu:159-177); copyHostToDeviceMemory
 * sets u_prev = u (DynamicalVariables.cu:280-284) and run() fills ghosts
 * (HydroPlugin.cpp:220). */
void kto_set_primitives(void *h, double t0, const double *e, const double *p,
                        const double *ut, const double *ux, const double *uy,
                        const double *un, const double *pi10, const double *Pi) {
  kto_ctx *c = (kto_ctx *)h;
  const kto_params *P = &c->P;
  const long nint = (long)P->nx * P->ny * P->nz;
  (void)t0;
  for (int k = 0; k < P->nz; ++k)
    for (int j = 0; j < P->ny; ++j)
      for (int i = 0; i < P->nx; ++i) {
        long a = i + (long)P->nx * (j + (long)P->ny * k);
        long s = idx(c, i + NG, j + NG, k + NG);
        double es = e[a], psv = p[a];
        double Pis = (Pi && P->bulk) ? Pi[a] : 0.0;
        c->e[s] = es;
        c->p[s] = psv;
        c->u[0][s] = ut[a];
        c->u[1][s] = ux[a];
        c->u[2][s] = uy[a];
        c->u[3][s] = un[a];
        for (int m = 0; m < 10; ++m) c->q[4 + m][s] = pi10 ? pi10[m * nint + a] : 0.0;
        if (P->bulk) c->q[14][s] = Pis;
        double PP = psv + Pis;
        c->q[0][s] = (es + PP) * ut[a] * ut[a] - PP + c->q[4][s];
        c->q[1][s] = (es + PP) * ut[a] * ux[a] + c->q[5][s];
        c->q[2][s] = (es + PP) * ut[a] * uy[a] + c->q[6][s];
        c->q[3][s] = (es + PP) * ut[a] * un[a] + c->q[7][s];
      }
  set_ghost_cells(c, c->q, c->u);
  /* u_prev gets the host copy of u, whose ghost cells are still zero */
  for (long s = 0; s < c->G; ++s)
    for (int m = 0; m < 4; ++m) c->up[m][s] = 0.0;
  for (int k = 0; k < P->nz; ++k)
    for (int j = 0; j < P->ny; ++j)
      for (int i = 0; i < P->nx; ++i) {
        long s = idx(c, i + NG, j + NG, k + NG);
        for (int m = 0; m < 4; ++m) c->up[m][s] = c->u[m][s];
      }
}

static double *field_ptr(kto_ctx *c, int f, int from_stage) {
  if (f >= 0 && f < NVMAX) return from_stage ? c->qS[f] : c->q[f];
  switch (f) {
    case KTO_E: return c->e;
    case KTO_P: return c->p;
    case KTO_UT: case KTO_UX: case KTO_UY: case KTO_UN: return c->u[f - KTO_UT];
    case KTO_UTP: case KTO_UXP: case KTO_UYP: case KTO_UNP: return c->up[f - KTO_UTP];
    case KTO_REG: return c->reg;
  }
  return NULL;
}

void kto_get(void *h, int field, int ghosted, int from_stage, double *out) {
  kto_ctx *c = (kto_ctx *)h;
  double *src = field_ptr(c, field, from_stage);
  if (!src) return;
  if (ghosted) {
    memcpy(out, src, sizeof(double) * c->G);
    return;
  }
  const kto_params *P = &c->P;
  for (int k = 0; k < P->nz; ++k)
    for (int j = 0; j < P->ny; ++j)
      for (int i = 0; i < P->nx; ++i)
        out[i + (long)P->nx * (j + (long)P->ny * k)] = src[idx(c, i + NG, j + NG, k + NG)];
}

void kto_set_field(void *h, int field, const double *in_ghosted) {
  kto_ctx *c = (kto_ctx *)h;
  double *dst = field_ptr(c, field, 0);
  if (dst) memcpy(dst, in_ghosted, sizeof(double) * c->G);
}
