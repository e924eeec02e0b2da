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
#define GVH_ENOCONV (-5)  /* a batched event's root-find reached its
                             Newton iteration budget (event retired)  */

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
                       FP32 half-site root-finds (tabulated and conformal
                       EoS); FP64 T^{tau mu} accumulation, the cell's own
                       root-find and its EoS                                  */
#define GVH_MIXED_F64_ROOTS 2 /* as GVH_MIXED with every root-find and its
                                 EoS lookups in FP64 (the 12 half-site ones
                                 included): the split of FP32 flux, limiter,
                                 source and stencil arithmetic and FP32 pi, Pi
                                 storage against FP64 accumulation, root-find
                                 and EoS inversion.  Single contexts only.   */

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
/* nsteps steps from run-loop index n_begin (as gvh_steps) with HIP events on
 * the context's stream around every stage kernel, waited for once at the end:
 * the summed stage-1 and stage-2 kernel times and the GPU time of the whole
 * sequence, in ms.  A measurement helper with no counterpart in the
 * reference (its only timer is clock() around the step,
 * HydroPlugin.cpp:88-105,414-426). */
int gvh_steps_timed(gvh_context *ctx, double t0, int n_begin, int nsteps, float *ms_stage1,
                    float *ms_stage2, float *ms_total);
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

/* ------------------------------------------------------------ freeze-out
 * The run loop's freeze-out finder (HydroPlugin.cpp:152-393 with
 * freezeout/FreezeOut.cpp and cornelius), searched on the device: a ring of
 * FOFREQ+1 time slices of (u^x, u^y, u^eta, e, pi^xx, pi^xy, pi^xeta,
 * pi^yy, pi^yeta, Pi) in HBM, and surface rows of GVH_FO_COLUMNS doubles:
 * tau x y eta, t*dsigma_{tau,x,y,eta}, u^x u^y u^eta, e T P,
 * pi^xx pi^xy pi^xeta pi^yy pi^yeta, Pi (HydroPlugin.cpp:317-386). */
#define GVH_FO_COLUMNS 20
#define GVH_FO_MAX_FREQ 64

/* Flags of gvh_freezeout_init.  0 reproduces the reference's rows. */
#define GVH_FO_EXACT_TAU 1      /* cell tau of every window = its slice's time;
                                   reference: one step late after the first
                                   window (HydroPlugin.cpp:304-305)          */
#define GVH_FO_NORMAL_TAU 2     /* normal scaled by the element's tau;
                                   reference: by the current t (:328-331)    */
#define GVH_FO_CURRENT_FIELDS 4 /* slices hold the current u and pi;
                                   reference: the previous step's after an odd
                                   number of steps, as its host copy reads the
                                   first-allocated buffers
                                   (DynamicalVariables.cu:310-331, 400-408,
                                   FullyDiscreteKurganovTadmorScheme.cu:132) */

/* Allocate the history ring of a context for windows of fofreq steps
 * (FOFREQ, HydroPlugin.cpp:40) at iso-energy density e_fo.  The lattice must
 * be 2+1D (nz == 1) or 3+1D (HydroPlugin.cpp:157-182). */
int gvh_freezeout_init(gvh_context *ctx, int fofreq, double e_fo, int flags);
/* Record the current state for run-loop index n >= 1: slice n % fofreq + 1,
 * or, when n % fofreq == 0, slice 1 after slice fofreq moved to slice 0
 * (swapAndSetHydroVariables / setHydroVariables, FreezeOut.cpp:49-126). */
int gvh_freezeout_record(gvh_context *ctx, int n);
/* Search the window of run-loop index n (the run loop calls it when
 * n % fofreq == fofreq - 1, HydroPlugin.cpp:276) at current time t; returns
 * the number of surface rows.  Slices from 2 on for n <= fofreq, else from 0. */
int gvh_freezeout_find(gvh_context *ctx, int n, double t, long long *nelements);
/* Copy the first nrows rows of the last search (nrows * GVH_FO_COLUMNS). */
int gvh_freezeout_rows(gvh_context *ctx, double *rows, long long nrows);
/* Copy variable var (0..9, order above) of logical slice `logical`. */
int gvh_freezeout_slot(gvh_context *ctx, int logical, int var, double *host_out);
/* Surface elements of independent cells (Cornelius::find_surface_3d/4d,
 * cornelius.cpp:2161-2287) on the device: corners[ncells][2^dim] with the
 * last axis fastest, dx[dim], iso-value v0.  counts[ncells]; centroids and
 * normals [ncells][10][dim]. */
int gvh_cornelius_cells(int dim, long long ncells, const double *corners, const double *dx,
                        double v0, int *counts, double *centroids, double *normals);
/* gvh_cornelius_cells without allocation.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * that a HIP graph may capture, tight serving loops, harnesses that forbid
 * hipMalloc).  corners, counts, centroids and normals are DEVICE pointers in
 * gvh_cornelius_cells' layouts; dx (dim values) is a host array.  The work
 * runs on `stream` (a hipStream_t; NULL: the null stream) and the call
 * returns once it has finished.  Its scratch (the element records and an
 * overflow counter) is a static device array of about 10 MiB that each GPU
 * holds until the process ends, so one call takes at most
 * GVH_CORNELIUS_NOALLOC_MAX_CELLS cells (more: GVH_EINVAL).  Every call on a
 * GPU uses that GPU's one array: calls that overlap on one GPU race. */
#define GVH_CORNELIUS_NOALLOC_MAX_CELLS 16384
int gvh_cornelius_cells_noalloc(int dim, long long ncells, const double *corners,
                                const double *dx, double v0, int *counts, double *centroids,
                                double *normals, void *stream);

/* ------------------------------------------------- validity diagnostics
 * Knudsen and inverse Reynolds numbers per cell: checkValidityKernel
 * (HydrodynamicValidity.cu:19-229) and the shear regulation factor that
 * regulateDissipativeCurrents stores beside them
 * (RegulateDissipativeCurrents.cu:132), i.e. the reference's VALIDITY_DOMAIN
 * (DynamicalVariables.cuh:83-98; the two arrays it allocates but never
 * writes are left out).  The reference fills them every step and never reads
 * them back; here they are computed on request. */
enum {
  GVH_V_KN_TAUPI = 0, /* knudsenNumberTaupi  = 5 (eta/s) theta / T          */
  GVH_V_KN_TAUBULK,   /* knudsenNumberTauPi  = theta / (1/tau_Pi)           */
  GVH_V_RE_PI,        /* inverseReynoldsNumberPimunu = sqrt|pi:pi| / p      */
  GVH_V_RE2_PI,       /* inverseReynoldsNumberTilde2Pimunu                  */
  GVH_V_RE_BULK,      /* inverseReynoldsNumberPi = |Pi| / p                 */
  GVH_V_RE2_BULK,     /* inverseReynoldsNumberTilde2Pi                      */
  GVH_V_TAUPI,        /* taupi = 5 (eta/s) / T                              */
  GVH_V_DXUX,         /* dxux, central difference                           */
  GVH_V_DYUY,         /* dyuy                                               */
  GVH_V_THETA,        /* theta, expansion rate                              */
  GVH_V_REGULATIONS,  /* regulations: shear factor of the last step's end   */
  GVH_VALIDITY_FIELDS
};
/* take pi and Pi from the current state instead of q^n of the last step */
#define GVH_VALIDITY_CURRENT_PI 1

/* Allocate the validity arrays (GVH_VALIDITY_FIELDS blocks of the ghosted
 * grid; regulations start at 1 as DynamicalVariables.cu:81) and record the
 * regulation factor in every later step. */
int gvh_validity_enable(gvh_context *ctx);
/* checkValidity (HydrodynamicValidity.cu:231-235) at proper time t from the
 * context's state, as the reference calls it at the end of a step
 * (FullyDiscreteKurganovTadmorScheme.cu:137) and before the first
 * (HydroPlugin.cpp:222): e, p, u and u_prev of the current state; pi and Pi
 * of q^n, the conserved state before the last step (the reference passes
 * d_q, which its run loop swaps only after the step), unless flags has
 * GVH_VALIDITY_CURRENT_PI or no step has been taken.  Enables the arrays on
 * first use.  Asynchronous on the context's stream. */
int gvh_check_validity(gvh_context *ctx, double t, int flags);
/* Copy validity field f: interior cells (ghosted == 0, nx*ny*nz values, x
 * fastest) or the full ghosted grid. */
int gvh_get_validity(gvh_context *ctx, int field, int ghosted, double *host_out);

/* -------------------------------------------------------------- run loop */
typedef struct gvh_run_options {
  int freq;      /* snapshot cadence FREQ (HydroPlugin.cpp:39); reference 10   */
  int fofreq;    /* freeze-out window FOFREQ (:40); reference 10               */
  int fo_flags;  /* GVH_FO_*; reference 0                                      */
  int snapshots; /* write the FREQ snapshots (outputDynamicalQuantities, :44) */
  int surface;   /* write <root_dir>/output/surface.dat                        */
  int verbose;   /* progress lines on stdout                                   */
  int validity;  /* with snapshots, also write the validity fields (the
                    reference's commented-out outputs, HydroPlugin.cpp:72-83)  */
} gvh_run_options;

typedef struct gvh_run_stats {
  int steps;              /* KT steps taken                                 */
  int frozen;             /* 1 if ended by the all-cells-frozen rule        */
  double t_final;         /* proper time at the end                         */
  long long surface_rows; /* rows written to surface.dat                    */
  double ms_steps;        /* wall time in the step calls                    */
  double ms_freezeout;    /* wall time recording and searching the surface  */
  double ms_total;
} gvh_run_stats;

void gvh_run_options_default(gvh_run_options *ro);

/* One hydrodynamic event: run() of HydroPlugin.cpp:107-434.  Initial
 * condition types 10 (input/e.dat, ut.dat, ... under root_dir; p from the
 * EoS) and 11 (input/Tmunu.dat) are read (InitialConditions.cpp:29-489);
 * snapshots go to output_dir as <name>_<t>.dat (FileIO.cpp:15-60), the
 * surface to <root_dir>/output/surface.dat (the reference opens
 * "output/surface.dat" in the working directory, which its driver passes as
 * the root directory).  ro and stats may be NULL. */
int gvh_run(const gvh_lattice_parameters *lattice, const gvh_ic_parameters *ic,
            const gvh_hydro_parameters *hydro, const char *root_dir, const char *output_dir,
            const gvh_options *opts, const gvh_run_options *ro, gvh_run_stats *stats);

/* HYDRO::run_hydro (HydroWrapper.cpp:17-87): options -c/--config DIR,
 * -o/--output DIR, -h/--hydro (run), -t/--test; parameters from the config
 * directory, root directory = working directory. */
int gvh_run_hydro(int argc, char **argv, const gvh_options *opts, const gvh_run_options *ro);

/* ------------------------------------------------------------- batches
 * Many events on one GPU with continuous batching.  No counterpart in the
 * reference, which runs one event per process (HydroPlugin.cpp:107-434):
 * each event here follows that run loop (nt + 1 steps at most, the
 * all-frozen stop of :396-413, freeze-out windows of :270-393) while all
 * running events advance together, one launch per kernel per step, from two
 * captured HIP graphs.  Events share the batch's gvh_options (EoS, bulk,
 * precision); each has its own lattice, time step and hydro parameters. */
typedef struct gvh_batch gvh_batch;

typedef struct gvh_batch_options {
  unsigned long long hbm_bytes; /* admission budget; 0 = 80% of free device memory */
  int max_events;               /* event slots (concurrently running or kept)      */
  int fofreq;                   /* freeze-out window FOFREQ, reference 10           */
  int autotune;                 /* time kernel arrangements per lattice shape       */
} gvh_batch_options;

typedef struct gvh_event {
  gvh_lattice_parameters lattice; /* numProperTimePoints = the run loop's nt */
  gvh_hydro_parameters hydro;     /* tau0, eta/s, T_fo, T_reg                */
  const double *e;                /* initial energy density, nx*ny*nz, x fastest;
                                     at rest, pi = Pi = 0, p from the EoS     */
  int surface;                    /* find the freeze-out surface              */
  int fo_flags;                   /* GVH_FO_* for the surface rows            */
  int keep_state;                 /* keep the final state for gvh_batch_field */
  int newton_budget;              /* 0: the reference's 10^6 Newton iterations per
                                     root-find (EnergyMomentumTensor.cu:15,22);
                                     n > 0: at most n, and the event is retired
                                     with status GVH_ENOCONV after the first step
                                     in which a root-find reached n without
                                     converging, so that an unstable event does
                                     not hold its co-batched events to 10^6-
                                     iteration chains                         */
} gvh_event;

typedef struct gvh_event_result {
  long long id;
  int steps;                /* KT steps taken                              */
  int frozen;               /* 1 if ended by the all-frozen rule           */
  double t_final;           /* proper time at the end                      */
  long long surface_rows;   /* rows found (gvh_batch_rows)                 */
  double t_submit, t_admit, t_done; /* seconds on the batch's clock        */
  int status;               /* GVH_OK; GVH_ENOCONV (iteration budget
                               reached); or the error code of a failed
                               admission (the event never ran)         */
  unsigned long long newton_not_converged; /* root-finds that reached the
                                              iteration limit            */
} gvh_event_result;

void gvh_batch_options_default(gvh_batch_options *bo);
int gvh_batch_create(const gvh_options *opts, const gvh_batch_options *bo, gvh_batch **out);
int gvh_batch_destroy(gvh_batch *b);
/* Queue an event (its arrays are copied).  The lattice and hydro
 * parameters are checked as gvh_create checks them.  An event that fails at
 * admission (device set-up) is finished with that error in its result's
 * status; the other events go on. */
int gvh_batch_submit(gvh_batch *b, const gvh_event *ev, long long *id);
/* Admit queued events that fit, advance every running event by one step,
 * move the freeze-out window searches on, retire finished events.  A
 * window's search runs beside the event's next steps (its history ring has
 * two halves, alternating by window); its rows are in gvh_batch_rows once the
 * event has finished.  *nrunning = running + queued events afterwards. */
int gvh_batch_step(gvh_batch *b, int *nrunning);
/* gvh_batch_step until nothing is queued or running. */
int gvh_batch_run(gvh_batch *b);
/* Up to max results of events finished since the last poll. */
int gvh_batch_poll(gvh_batch *b, gvh_event_result *out, int max, int *n);
/* Surface rows (GVH_FO_COLUMNS doubles each) of a finished event. */
int gvh_batch_rows(gvh_batch *b, long long id, double *rows, long long nrows);
/* Interior field of a finished event submitted with keep_state. */
int gvh_batch_field(gvh_batch *b, long long id, int field, double *host_out);
/* Forget a finished event (frees a kept state). */
int gvh_batch_release(gvh_batch *b, long long id);
int gvh_batch_stats(gvh_batch *b, long long *queued, long long *running,
                    unsigned long long *hbm_used, unsigned long long *hbm_budget);

/* Time the stage-kernel arrangements and traversal strips on the context's
 * current state (restored afterwards) and keep the fastest; no counterpart
 * in the reference (fixed launch shapes, CudaConfiguration.cu:27-28). */
int gvh_autotune(gvh_context *ctx, int steps, int *variant, int *strip, float *ms_per_step);

const char *gvh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
