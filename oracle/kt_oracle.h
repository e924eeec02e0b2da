/*
 * kt_oracle.h -- CPU FP64 restatement of gpu-vh's Kurganov-Tadmor step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the correctness oracle that the HIP
 * product path is checked against.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and only as the checker or as the
 * timed CPU baseline.  The product library never links or calls it.
 *
 * Semantics follow the reference with PRECISION=double and every float-typed
 * libm call mapped to its double counterpart ("ref64-clean"):
 *   fminf->fmin, fmaxf->fmax, fabsf->fabs, sqrtf->sqrt, fmaf->fma,
 *   powf->pow, tanhf->tanh, rsqrtf->1/sqrt, fdividef->'/'.
 * Float-suffixed literals of the reference keep their float value
 * ((double)0.2f etc.), truncated decimal constants are kept as written.
 */
#ifndef KT_ORACLE_H
#define KT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* equation of state kinds */
#define KTO_EOS_CONFORMAL 0 /* p=e/3 (EquationOfState.cuh:16 CONFORMAL_EOS)      */
#define KTO_EOS_WB 1        /* Wuppertal-Budapest rational fits (EquationOfState.cu) */
#define KTO_EOS_TABLE 2     /* tabulated p, cs2, T on a pseudo-log grid in e       */

/* field ids for kto_get / kto_set_field */
enum {
  KTO_TTT = 0, KTO_TTX, KTO_TTY, KTO_TTN,
  KTO_PITT, KTO_PITX, KTO_PITY, KTO_PITN, KTO_PIXX, KTO_PIXY, KTO_PIXN,
  KTO_PIYY, KTO_PIYN, KTO_PINN, KTO_PI,
  KTO_E = 15, KTO_P = 16, KTO_UT = 17, KTO_UX, KTO_UY, KTO_UN,
  KTO_UTP = 21, KTO_UXP, KTO_UYP, KTO_UNP,
  KTO_REG = 25,
  KTO_NFIELDS = 26
};

typedef struct kto_params {
  int nx, ny, nz;        /* interior lattice points                       */
  double dx, dy, dz, dt; /* lattice spacings                              */
  double etabar;         /* eta/s                                         */
  int bulk;              /* 1: evolve Pi (reference #define PI)           */
  int eos;               /* KTO_EOS_*                                     */
  double zeta_tpeak;     /* x = T/zeta_tpeak in zeta/s (ref: 1.01355)     */
  double T_reg;          /* regulation temperature [fm^-1]                */
  /* tabulated EoS (eos == KTO_EOS_TABLE) */
  int tab_n;             /* number of nodes                               */
  int tab_per_octave;    /* nodes per power of two in e                   */
  int tab_emin_exp;      /* first node is e = 2^tab_emin_exp              */
  const double *tab_p, *tab_cs2, *tab_T;
} kto_params;

void *kto_create(const kto_params *P);
void kto_destroy(void *h);
/* Initial state from primitive fields on the interior (nx*ny*nz, i fastest):
 * computes T^{tau mu} like setConservedVariables, sets u_prev = u and fills
 * ghost cells.  pi10 holds 10 arrays in the order tt tx ty tn xx xy xn yy yn nn;
 * Pi may be NULL (zero). */
void kto_set_primitives(void *h, double t0, const double *e, const double *p,
                        const double *ut, const double *ux, const double *uy,
                        const double *un, const double *pi10, const double *Pi);
/* One KT step at proper time t (twoStepRungeKutta + the conserved-variable
 * buffer swap of the run loop). */
void kto_step(void *h, double t);
void kto_steps(void *h, double t0, int n_begin, int nsteps);
/* Only the predictor Euler step q -> qS at time t (for unit pinning). */
void kto_euler_only(void *h, double t);
/* Copy a field out (interior only when ghosted == 0, full ghosted grid else).
 * For kto_euler_only the conserved fields 0..14 are read from qS when
 * from_stage == 1. */
void kto_get(void *h, int field, int ghosted, int from_stage, double *out);
void kto_set_field(void *h, int field, const double *in_ghosted);
long kto_ghosted_size(void *h);

/* unit-level exports for the oracle's own pinning tests */
double kto_minmod(double x, double y);
double kto_approx_deriv(double x, double y, double z);
void kto_eos(const kto_params *P, double e, double *p, double *cs2, double *T);
double kto_e_of_T(const kto_params *P, double T);
void kto_eos_p_batch(const kto_params *P, long n, const double *e, double *p);
double kto_zeta_over_s(const kto_params *P, double T);
/* q[15] -> out[6] = e, p, ut, ux, uy, un */
void kto_inferred(const kto_params *P, double t, const double *q, double ePrev,
                  double *out);
/* KT flux over a 5-point stencil data[15*5] (var-major), dir 0/1/2,
 * forward (i+1/2) when fwd != 0 else backward (i-1/2); H[15]. */
void kto_flux(const kto_params *P, int dir, const double *data, double t,
              double ePrev, int fwd, double *H);
/* loadSourceTerms2 (T^{tau mu} sources and pi, Pi relaxation RHS) of one
 * cell: Q[15] conserved, e, p, u[4], u_prev[4], du[3][4] = d_i u^m,
 * dp[3] = d_i p; S[15] out. */
void kto_source_cell(const kto_params *P, double t, const double *Q, double e,
                     double p, const double *u, const double *up,
                     const double du[3][4], const double *dp, double *S);
/* shear regulation factor of one cell (1 when e > e_reg) */
double kto_regulate_factor(double t, const double *pi, double e, double p,
                           const double *u, double e_reg);
/* validity diagnostics (HydrodynamicValidity.cu), output field order */
enum {
  KTO_V_KN_TAUPI = 0, /* knudsenNumberTaupi            */
  KTO_V_KN_TAUBULK,   /* knudsenNumberTauPi            */
  KTO_V_RE_PI,        /* inverseReynoldsNumberPimunu   */
  KTO_V_RE2_PI,       /* inverseReynoldsNumberTilde2Pimunu */
  KTO_V_RE_BULK,      /* inverseReynoldsNumberPi       */
  KTO_V_RE2_BULK,     /* inverseReynoldsNumberTilde2Pi */
  KTO_V_TAUPI, KTO_V_DXUX, KTO_V_DYUY, KTO_V_THETA,
  KTO_V_REG,          /* regulations (RegulateDissipativeCurrents.cu:132) */
  KTO_V_N
};
/* checkValidityKernel of one cell: du[d][m] = d_d u^m central differences,
 * pi[10] tt..nn, out[KTO_V_REG] (all but the regulation factor). */
void kto_validity_cell(const kto_params *P, double t, double e, double p,
                       const double *u, const double *up, const double du[3][4],
                       const double *pi, double Pi, double *out);
/* The validity arrays of the lattice (KTO_V_N ghosted arrays, interior
 * written); prev_q != 0 takes pi, Pi from q^n of the last step. */
void kto_validity(void *h, double t, int prev_q, double *out);
/* Newton iteration count of the last kto_inferred call (non-conformal). */
int kto_last_newton_iters(void);

#ifdef __cplusplus
}
#endif
#endif
