/*
 * gpuvh.h -- C ABI of the MI355X-native gpu-vh Kurganov-Tadmor step.
 *
 * Plain pointers and sizes only.  Every entry point returns GVH_OK (0) or a
 * negative GVH_E* code; gvh_last_error() gives the message of the calling
 * thread's last failure.  No entry point exits the process.
 *
 * Reference paths are relative to the derekeverett/gpu-vh root.
 */
#ifndef GPUVH_H
#define GPUVH_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GVH_OK 0
#define GVH_EINVAL (-1)   /* bad argument                              */
#define GVH_EHIP (-2)     /* HIP runtime failure                       */
#define GVH_EIO (-3)      /* file missing or malformed                 */
#define GVH_ENOMEM (-4)   /* device allocation failed                  */

/* Parameter blocks: same fields in the same order as the reference's POD
 * structs, so a caller holding one can pass it field for field. */

/* LatticeParameters, rhic/include/LatticeParameters.h:17-31 */
typedef struct gvh_lattice_parameters {
  int numLatticePointsX;
  int numLatticePointsY;
  int numLatticePointsRapidity;
  int numComputationalLatticePointsX;
  int numComputationalLatticePointsY;
  int numComputationalLatticePointsRapidity;
  int numProperTimePoints;
  double latticeSpacingX;
  double latticeSpacingY;
  double latticeSpacingRapidity;
  double latticeSpacingProperTime;
} gvh_lattice_parameters;

/* InitialConditionParameters, rhic/include/InitialConditionParameters.h:13-26 */
typedef struct gvh_ic_parameters {
  int initialConditionType;
  int numberOfNucleonsPerNuclei;
  int numberOfSourceFiles;
  double initialEnergyDensity;
  double scatteringCrossSectionNN;
  double impactParameter;
  double fractionOfBinaryCollisions;
  double rapidityVariance;
  double rapidityMean;
  int sourceType;
} gvh_ic_parameters;

/* HydroParameters, rhic/include/HydroParameters.h:15-22 */
typedef struct gvh_hydro_parameters {
  double initialProperTimePoint;
  double shearViscosityToEntropyDensity;
  double freezeoutTemperatureGeV;
  int initializePimunuNavierStokes;
  double regulationTemperatureGeV;
} gvh_hydro_parameters;

/* Positional .properties readers (keys ignored, line order significant):
 * readLatticeParameters (rhic/src/LatticeParameters.cpp:65-138),
 * readHydroParameters (rhic/src/HydroParameters.cpp:49-98),
 * readInitialConditionParameters (rhic/src/InitialConditionParameters.cpp:65-149).
 * A missing file is GVH_EIO (the reference prints and continues). */
int gvh_read_lattice_parameters(const char *config_dir, gvh_lattice_parameters *out);
int gvh_read_hydro_parameters(const char *config_dir, gvh_hydro_parameters *out);
int gvh_read_ic_parameters(const char *config_dir, gvh_ic_parameters *out);

/* Equation of state.  The reference selects it at compile time
 * (CONFORMAL_EOS, rhic/include/EquationOfState.cuh:16); here it is a
 * context option. */
#define GVH_EOS_CONFORMAL 0 /* p = e/3, T = (e/13.8997)^(1/4)                 */
#define GVH_EOS_WB 1        /* Wuppertal-Budapest fits, EquationOfState.cu:31-204 */
#define GVH_EOS_TABLE 2     /* tabulated p, c_s^2, T on a pseudo-log grid in e */

#define GVH_FP64 0  /* plain FP64 everywhere                                   */
#define GVH_MIXED 1 /* FP32 stencil/flux/source and FP32 pi, Pi storage;
                       FP64 T^{tau mu} accumulation, root-find and EoS       */

typedef struct gvh_options {
  int eos;             /* GVH_EOS_*                                         */
  int bulk;            /* 1: evolve Pi (reference #define PI,
                          rhic/include/DynamicalVariables.cuh:14)           */
  int precision;       /* GVH_FP64 or GVH_MIXED                             */
  double zeta_tpeak;   /* temperature scale of zeta/s [fm^-1];
                          reference 1.01355 (TransportCoefficients.cu:28)   */
  /* tabulated EoS, used when eos == GVH_EOS_TABLE: node k sits at
   * e_k = 2^(emin_exp + k/per_octave) * (1 + (k % per_octave)/per_octave) */
  int tab_n;
  int tab_per_octave;
  int tab_emin_exp;
  const double *tab_p;
  const double *tab_cs2;
  const double *tab_T;
  void *stream;        /* hipStream_t to run on; NULL = the context's own */
} gvh_options;

/* Field ids for state transfer, in the reference's SoA order:
 * CONSERVED_VARIABLES (DynamicalVariables.cuh:50-71), then e, p,
 * FLUID_VELOCITY u (:74-80) and the previous-step velocity up. */
enum {
  GVH_TTT = 0, GVH_TTX, GVH_TTY, GVH_TTN,
  GVH_PITT, GVH_PITX, GVH_PITY, GVH_PITN, GVH_PIXX, GVH_PIXY, GVH_PIXN,
  GVH_PIYY, GVH_PIYN, GVH_PINN, GVH_PI,
  GVH_E = 15, GVH_P, GVH_UT, GVH_UX, GVH_UY, GVH_UN,
  GVH_UTP, GVH_UXP, GVH_UYP, GVH_UNP,
  GVH_NFIELDS
};

typedef struct gvh_context gvh_context;

/* Device-side failure counters (the reference only printf()s these,
 * EnergyMomentumTensor.cu:43,92-94,96-99). */
typedef struct gvh_diagnostics {
  unsigned long long newton_not_converged; /* MAX_ITERS reached           */
  unsigned long long e_nan;                /* inferred e is NaN           */
  unsigned long long e_clamped;            /* e < 1e-7 clamped            */
} gvh_diagnostics;

/* Allocate a context: device state for one lattice (all buffers of
 * allocateDeviceMemory, DynamicalVariables.cu:147-269) plus the step
 * constants of initializeCUDAConstantParameters (CudaConfiguration.cu:209-248).
 * The regulation temperature comes from hydro->regulationTemperatureGeV. */
int gvh_create(const gvh_lattice_parameters *lattice,
               const gvh_hydro_parameters *hydro, const gvh_options *opts,
               gvh_context **out);
int gvh_destroy(gvh_context *ctx);

/* Initial state from primitive fields on the interior lattice
 * (nx*ny*nz values each, x fastest): T^{tau mu} from
 * setConservedVariables (DynamicalVariables.cu:349-392), u_prev = u
 * (copyHostToDeviceMemory, :271-308) and ghost cells (HydroPlugin.cpp:220).
 * pi10 holds 10 consecutive arrays tt tx ty tn xx xy xn yy yn nn; pi10 and
 * Pi may be NULL for zero. */
int gvh_set_initial_state(gvh_context *ctx, double t0, const double *e,
                          const double *p, const double *ut, const double *ux,
                          const double *uy, const double *un,
                          const double *pi10, const double *Pi);

/* Replace one field of the current state from a host array on the full
 * ghosted grid ((nx+4)(ny+4)(nz+4) values). */
int gvh_set_field(gvh_context *ctx, int field, const double *host_ghosted);

/* One KT step at proper time t: twoStepRungeKutta
 * (FullyDiscreteKurganovTadmorScheme.cu:109-139) followed by the
 * conserved-variable swap of the run loop (HydroPlugin.cpp:422).
 * Asynchronous on the context's stream. */
int gvh_step(gvh_context *ctx, double t);
/* nsteps steps at t = t0 + n*dt, n = n_begin .. n_begin+nsteps-1, as the
 * run loop computes them (HydroPlugin.cpp:415,424). */
int gvh_steps(gvh_context *ctx, double t0, int n_begin, int nsteps);
/* One step with a HIP event pair around each of the two stage kernels;
 * blocks until done and returns their durations in milliseconds. */
int gvh_step_timed(gvh_context *ctx, double t, float *ms_stage1, float *ms_stage2);
int gvh_synchronize(gvh_context *ctx);

/* Copy a field of the current state to the host: interior only
 * (ghosted == 0, nx*ny*nz values) or the full ghosted grid. */
int gvh_get_field(gvh_context *ctx, int field, int ghosted, double *host_out);

/* Number of interior cells with e > e_threshold (the run loop's
 * termination scan, HydroPlugin.cpp:396-407), reduced on the device. */
int gvh_count_above(gvh_context *ctx, double e_threshold, long long *count);

int gvh_get_diagnostics(gvh_context *ctx, gvh_diagnostics *out);

/* Equation-of-state helpers of the context (host evaluation):
 * equilibriumEnergyDensity(T) and effectiveTemperature(e). */
int gvh_energy_density_of_T(gvh_context *ctx, double T, double *e);
/* equilibriumPressure (EquationOfState.cu:31-80) of n energy densities, for
 * initial conditions: the reference sets p = p(e) when it reads e.dat
 * (InitialConditions.cpp:127). */
int gvh_equilibrium_pressure(gvh_context *ctx, long long n, const double *e, double *p);

/* Lattice geometry of the context. */
int gvh_dims(gvh_context *ctx, int *nx, int *ny, int *nz);

const char *gvh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
