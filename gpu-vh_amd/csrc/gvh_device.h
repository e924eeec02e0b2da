// gvh_device.h -- device math of the KT step for gfx950 (FP64 form).
//
// Reference files are cited per function (paths relative to rhic/src).
// All state arithmetic is double.  Divisions that the reference performs by
// a per-launch constant (dt, dx, t) are replaced by a multiplication with a
// host-computed reciprocal; that, FMA contraction and the flux written as
// (u_dir/u^tau)*q change results at the rounding level only.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace gvh {

constexpr int NG = 2;             // ghost layers (LatticeParameters.h:13-15)
constexpr int MAX_ITERS = 1000000; // EnergyMomentumTensor.cu:15
constexpr double THETA = 2.0;     // FluxLimiter.cu:15
constexpr double EOS_FACTOR = 13.8997; // EquationOfState.cuh:24

enum { EOS_CONFORMAL = 0, EOS_WB = 1, EOS_TABLE = 2 };

// Threads per block of the stage kernels, and cells per tile of a batch.
// One wave per block: a block holds its CU resources until its slowest wave
// ends, and the waves of a block differ in their root-find iteration counts,
// so smaller blocks leave fewer idle slots (the 201x201x101 step: 2.40 ms
// with 64, 2.51 ms with 128, 2.57 ms with 256 threads).
#ifndef GVH_BLOCK
#define GVH_BLOCK 64
#endif

// counters of failure events, indexes into StageArgs::diag
enum { DIAG_NOT_CONVERGED = 0, DIAG_E_NAN = 1, DIAG_E_CLAMPED = 2, DIAG_N = 4 };

// Tabulated EoS on the device.  Node k sits at
// e_k = 2^(emin_exp + k/M) (1 + (k % M)/M); per node the value and the
// difference to node k+1 are stored so that an interpolation is one load.
struct EosTable {
  const double4 *pc;  // {p_k, p_{k+1}-p_k, cs2_k, cs2_{k+1}-cs2_k}
  const double2 *tt;  // {T_k, T_{k+1}-T_k}
  const float4 *pcf;  // pc rounded to FP32 (the mixed form's flux root-finds)
  int n, M, emin_exp;
  double e_first, e_second;  // e_0, e_1
  double e_pen, e_last;      // e_{n-2}, e_{n-1}
  double inv_first_span;     // 1/(e_1 - e_0)
  double inv_last_span;      // 1/(e_{n-1} - e_{n-2})
  float f_first, f_pen, f_last, f_inv_first_span, f_inv_last_span;  // the same in FP32
};

struct StageArgs {
  const double *__restrict__ qin;   // stage input conserved, 15 x G
  const double *__restrict__ qn;    // q^n for the convex average (stage 2)
  double *__restrict__ qout;        // stage output conserved
  const double *__restrict__ uin;   // u of the stage input, 4 x G
  const double *__restrict__ uprev; // u used for the time derivative
  double *__restrict__ uout;
  const double *__restrict__ ein;
  const double *__restrict__ pin;
  double *__restrict__ eout;
  double *__restrict__ pout;
  unsigned long long *__restrict__ diag;
  double *__restrict__ reg;  // stage 2 stores the shear regulation factor here if set
  long long G;          // stride between variables
  int nx, ny, nz, ncx, ncy;
  double t;             // time of the Euler step
  double tinf;          // time of the inferred-variable solve
  double dt, inv_dt;
  double h[3], inv_h[3], half_inv_h[3];  // dx, dy, deta and reciprocals
  double etabar;        // eta/s
  double zeta_tpeak;    // zeta/s temperature scale
  double e_reg;         // regulation below this energy density
  EosTable tab;
  int strip, nstrips;   // y rows per traversal strip, number of strips
  int variant;          // kernel arrangement of a stage (gvh_stage.hip)
  int max_iters;        // Newton iterations per root-find: MAX_ITERS (the reference's),
                        // or a smaller budget of a batched event
};

// Buffers of one event and their current roles: the reference's pointer
// swaps (FullyDiscreteKurganovTadmorScheme.cu:117-133,
// DynamicalVariables.cu:400-408) as indices.  q[iq] holds q^n, q[1-iq] the
// corrector output Q, q[2] the predictor qS; u[iu] u^n, u[iup] u^{n-1},
// u[3-iu-iup] the predictor u*; (e, p)[ie] the current state.
struct Roles {
  int iq, iu, iup, ie;
};
struct Buffers {
  double *q[3], *u[3], *e[2], *p[2];
};

// State pointers of RK stage 1 or 2 for the given roles.
__host__ __device__ inline void stage_pointers(StageArgs &A, const Buffers &B, const Roles &R,
                                               int stage) {
  const int iqs = 2, iQ = 1 - R.iq, ius = 3 - R.iu - R.iup;
  const int ie_in = stage == 1 ? R.ie : 1 - R.ie;
  A.qin = B.q[stage == 1 ? R.iq : iqs];
  A.qn = B.q[R.iq];
  A.qout = B.q[stage == 1 ? iqs : iQ];
  A.uin = B.u[stage == 1 ? R.iu : ius];
  A.uprev = B.u[stage == 1 ? R.iup : R.iu];
  A.uout = B.u[stage == 1 ? ius : R.iup];
  A.ein = B.e[ie_in];
  A.pin = B.p[ie_in];
  A.eout = B.e[1 - ie_in];
  A.pout = B.p[1 - ie_in];
}

// Times of stage 1 or 2 of the step that starts at t (the predictor's
// inferred variables are solved at t + dt, FullyDiscreteKurganovTadmorScheme.cu:119).
__host__ __device__ inline void stage_times(StageArgs &A, double t, int stage) {
  A.t = stage == 1 ? t : t + A.dt;
  A.tinf = t + A.dt;
}

// Roles after a step: u^{n+1} was written into the u^{n-1} buffer, q^{n+1}
// into Q.
__host__ __device__ inline void rotate_roles(Roles &R) {
  const int old_u = R.iu;
  R.iu = R.iup;
  R.iup = old_u;
  R.iq = 1 - R.iq;
}

// One event of a batch in device memory.
enum { SLOT_FREE = 0, SLOT_RUN = 1, SLOT_DONE = 2 };
struct BatchSlot {
  StageArgs A;        // constants of the event; pointers and times are set per stage
  Buffers B;
  Roles R;
  double t0;
  int steps;          // KT steps done
  int status;         // SLOT_*
  int max_steps;      // the run loop's nt + 1
  int frozen;         // steps seen with no interior cell above e_fo (never reset)
  int fofreq;
  int fo_flags;       // GVH_FO_CURRENT_FIELDS selects the current u, pi for the ring
  int frozen_stop;    // 1 if it ended by the all-frozen rule
  int error;          // 1 if it was retired for a root-find over its iteration budget
  int bulk;           // Pi is evolved
  int tile_first, ntiles;  // its tiles (blocks of cells) in the tile map
  double e_fo;
  unsigned long long above;  // interior cells with e > e_fo in the current state
  double *ring;       // freeze-out history, halves x (fofreq+1) x 10 x nint, or nullptr
  long long nint;
  int halves;         // 2: window m records into ring half m mod 2
};

// ------------------------------------------------------- cell access
// Element s of an array with a uniform (scalar) base: the byte offset is
// formed as a 32-bit product, so a load or store is the scalar-base +
// 32-bit-vector-offset form and needs no 64-bit address per access.  Every
// array of a lattice is below 4 GiB (the context checks the ghosted size).
template <typename T>
__device__ __forceinline__ T ld_cell(const T *base, unsigned s) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                      s * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void st_cell(T *base, unsigned s, T v) {
  *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + s * (unsigned)sizeof(T)) = v;
}

// ------------------------------------------------------------ limiter
// approximateDerivative (FluxLimiter.cu:33-39): the nested generalized
// minmod(theta*dl, minmod(dc, theta*dr)) with sgn(0)=+1 equals
// sign * min(|theta dl|, |dc|, |theta dr|) when all three arguments are in the
// same sign class (x < 0 versus x >= 0) and +0 otherwise -- exactly, since
// min and the factors 2 and 1/2 do not round.  That form needs no products
// of signs.  With a = y - x, c = z - y, b = z - x: min(2|a|, |b|/2, 2|c|)
// = 2 min(|a|, |c|, |b|/4) exactly, and a, c in one class force b into it
// (y >= x and z >= y give z >= x, likewise for <), so the class test needs
// only a and c.
//
// Every use takes half of it (the half-site states q +- D/2) or twice the
// half (the gradients), so the kernels evaluate D/2 = sign min(|a|, |c|,
// |b|/4) directly; the factors 2 and 1/2 cancel exactly.
// FP32: the two-argument minmod is the median of (p, q, 0), so
// D/2 = med3(a, med3(b/4, c, 0), 0), two v_med3_f32.
// FP64 (there is no v_med3_f64): the magnitude from two v_min_f64, the sign
// of a inserted into its high word, zero when a and c differ in sign bit.
// A zero a or c makes the magnitude zero, so classing -0 with the negatives
// changes at most the sign of a zero result.
static_assert(THETA == 2.0, "the half-slope form assumes theta = 2");
template <typename R>
__device__ __forceinline__ R half_slope_t(R x, R y, R z);
template <>
__device__ __forceinline__ float half_slope_t<float>(float x, float y, float z) {
  const float a = y - x, b = z - x, c = z - y;
  return __builtin_amdgcn_fmed3f(a, __builtin_amdgcn_fmed3f(b * 0.25f, c, 0.0f), 0.0f);
}
template <>
__device__ __forceinline__ double half_slope_t<double>(double x, double y, double z) {
  const double a = y - x, b = z - x, c = z - y;
  const double m = fmin(fmin(fabs(a), fabs(c)), fabs(b) * 0.25);
  int sx = __double2hiint(a) ^ __double2hiint(c);
  // an empty register constraint keeps the sign test one 32-bit xor and
  // compare; left visible, the combiner widens it to a 64-bit xor and compare
  asm("" : "+v"(sx));
  return sx >= 0 ? copysign(m, a) : 0.0;
}
template <typename R>
__device__ __forceinline__ R slope_t(R x, R y, R z) {
  const R h = half_slope_t<R>(x, y, z);
  return h + h;
}
__device__ __forceinline__ double slope(double x, double y, double z) {
  return slope_t<double>(x, y, z);
}

// ------------------------------------------------ division, square root
// The IEEE-exact expansions of a/b and sqrt(x) on gfx950 cost ~11 and ~15
// VALU instructions.  These start from the hardware approximations
// (v_rcp_f64, v_rsq_f64: at most 4.6e-8 and 5.3e-8 relative, measured on
// 4M inputs over [1e-6, 1e6]) and refine with FMA steps to 0.5 ulp on
// those inputs (the reciprocal needs both Newton steps: one leaves 10 ulp).
// Operands here are finite, normal and nonzero in any physical state;
// special values are not treated.
__device__ __forceinline__ double rcp_nr(double b) {
  double r = __builtin_amdgcn_rcp(b);
  double e = fma(-b, r, 1.0);
  r = fma(r, e, r);
  e = fma(-b, r, 1.0);
  return fma(r, e, r);
}
__device__ __forceinline__ double div_nr(double a, double b) {
  const double r = rcp_nr(b);
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);  // ~ 1/sqrt(x)
  double g = x * y, h = 0.5 * y;
  double e = fma(-g, h, 0.5);
  g = fma(g, e, g);
  h = fma(h, e, h);
  e = fma(-g, h, 0.5);
  g = fma(g, e, g);
  h = fma(h, e, h);
  const double d = fma(-g, g, x);
  return x == 0.0 ? x : fma(d, h, g);
}

// The same operations in the arithmetic type of a root-find: FP64 as above,
// FP32 from the hardware approximations (about 1 ulp).
__device__ __forceinline__ double rcp_t(double b) { return rcp_nr(b); }
__device__ __forceinline__ double div_t(double a, double b) { return div_nr(a, b); }
__device__ __forceinline__ double sqrt_t(double x) { return sqrt_nr(x); }
__device__ __forceinline__ float rcp_t(float b) { return __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float div_t(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float sqrt_t(float x) { return __builtin_amdgcn_sqrtf(x); }

// a / b in the stencil and source arithmetic of a stage: the IEEE division
// in FP64 (the reference's expression), the hardware reciprocal (about 1 ulp)
// times a in FP32.
__device__ __forceinline__ double div_st(double a, double b) { return a / b; }
__device__ __forceinline__ float div_st(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// ---------------------------------------------------------------- EoS
struct TabPos {
  int k;
  double fr;
};

__device__ __forceinline__ TabPos tab_locate(const EosTable &T, double e) {
  TabPos r;
  if (e >= T.e_first && e < T.e_last) {
    int E;
    const double m = frexp(e, &E);  // e = m 2^E, m in [0.5, 1)
    const double f = (2.0 * m - 1.0) * (double)T.M;
    const int j = (int)f;
    r.fr = f - (double)j;
    r.k = (E - 1 - T.emin_exp) * T.M + j;
    if (r.k > T.n - 2) {  // f rounded up onto the last node
      r.k = T.n - 2;
      r.fr += 1.0;
    }
  } else if (e >= T.e_last) {  // beyond the table, +inf included
    r.k = T.n - 2;
    r.fr = (e - T.e_pen) * T.inv_last_span;
  } else {  // below the first node, or NaN
    r.k = 0;
    r.fr = (e - T.e_first) * T.inv_first_span;
  }
  return r;
}

// equilibriumPressure (EquationOfState.cu:31-80) and speedOfSoundSquared
// (:83-116); the conformal c_s^2 is the reference's integer 1/3 == 0.
template <int EOS>
__device__ __forceinline__ void eos_p_cs2(const EosTable &T, double e, double &p,
                                          double &cs2) {
  if constexpr (EOS == EOS_CONFORMAL) {
    p = e / 3.0;
    cs2 = 0.0;
  } else if constexpr (EOS == EOS_TABLE) {
    TabPos q = tab_locate(T, e);
    const double4 v = ld_cell(T.pc, (unsigned)q.k);
    p = v.x + q.fr * v.y;
    cs2 = v.z + q.fr * v.w;
  } else {
    double e2 = e * e, e3 = e2 * e, e4 = e3 * e, e5 = e4 * e, e6 = e5 * e;
    double e7 = e6 * e, e8 = e7 * e, e9 = e8 * e, e10 = e9 * e, e11 = e10 * e;
    double e12 = e11 * e, e13 = e12 * e;
    double a = fma(1.9531729608963267e-11, e12, fma(3.1188455176941583e-7, e11,
               fma(0.0009417586777847889, e10, fma(0.7158279081255019, e9,
               fma(141.5073484468774, e8, fma(6340.448389300905, e7,
               fma(41913.439282708554, e6, fma(334334.4309240126, e5,
               fma(1.6357487344679043e6, e4, fma(3.1729694865420084e6, e3,
               fma(1.077580993288114e6, e2, fma(9737.845799644809, e,
               -0.25181736420168666))))))))))));
    double b = fma(3.2581066229887368e-18, e12, fma(5.928138360995685e-11, e11,
               fma(9.601103399348206e-7, e10, fma(0.002962497695527404, e9,
               fma(2.3405487982094204, e8, fma(499.04919730607065, e7,
               fma(26452.34905933697, e6, fma(278581.2989342773, e5,
               fma(1.7851642641834426e6, e4, fma(1.3512402226067686e7, e3,
               fma(2.0931169138134286e7, e2, fma(4.0574329080826794e6, e,
               45829.44617893836))))))))))));
    p = a / b;
    double cn = 5.191934309650155e-32 + 4.123605749683891e-23 * e +
                3.1955868410879504e-16 * e2 + 1.4170364808063119e-10 * e3 +
                6.087136671592452e-6 * e4 + 0.02969737949090831 * e5 +
                15.382615282179595 * e6 + 460.6487249985994 * e7 +
                1612.4245252438795 * e8 + 275.0492627924299 * e9 +
                58.60283714484669 * e10 + 6.504847576502024 * e11 +
                0.03009027913262399 * e12 + 8.189430244031285e-6 * e13;
    double cd = 1.4637868900982493e-30 + 6.716598285341542e-22 * e +
                3.5477700458515908e-15 * e2 + 1.1225580509306008e-9 * e3 +
                0.00003551782901018317 * e4 + 0.13653226327408863 * e5 +
                60.85769171450653 * e6 + 1800.5461219450308 * e7 +
                15190.225535036281 * e8 + 590.2572000057821 * e9 +
                293.99144775704605 * e10 + 21.461303090563028 * e11 +
                0.09301685073435291 * e12 + 0.000024810902623582917 * e13;
    cs2 = cn / cd;
  }
}

template <int EOS>
__device__ __forceinline__ double eos_p(const EosTable &T, double e) {
  if constexpr (EOS == EOS_CONFORMAL) {
    return e / 3.0;
  } else if constexpr (EOS == EOS_TABLE) {
    TabPos q = tab_locate(T, e);
    const double4 v = ld_cell(T.pc, (unsigned)q.k);
    return v.x + q.fr * v.y;
  } else {
    double p, c;
    eos_p_cs2<EOS>(T, e, p, c);
    return p;
  }
}

// p and c_s^2 of the tabulated EoS in FP32 (tab_locate with float
// arithmetic, FP32 node values).
__device__ __forceinline__ void tab_p_cs2_f(const EosTable &T, float e, float &p, float &cs2) {
  int k;
  float fr;
  if (e >= T.f_first && e < T.f_last) {
    int E;
    const float m = frexpf(e, &E);
    const float f = (2.0f * m - 1.0f) * (float)T.M;
    const int j = (int)f;
    fr = f - (float)j;
    k = (E - 1 - T.emin_exp) * T.M + j;
    if (k > T.n - 2) {
      k = T.n - 2;
      fr += 1.0f;
    }
  } else if (e >= T.f_last) {
    k = T.n - 2;
    fr = (e - T.f_pen) * T.f_inv_last_span;
  } else {
    k = 0;
    fr = (e - T.f_first) * T.f_inv_first_span;
  }
  const float4 v = ld_cell(T.pcf, (unsigned)k);
  p = v.x + fr * v.y;
  cs2 = v.z + fr * v.w;
}

// eos_p_cs2 / eos_p in the arithmetic type F (float only for the conformal
// and tabulated EoS).
template <int EOS, typename F>
__device__ __forceinline__ void eos_p_cs2_t(const EosTable &T, F e, F &p, F &cs2) {
  if constexpr (std::is_same<F, double>::value) {
    eos_p_cs2<EOS>(T, e, p, cs2);
  } else {
    static_assert(EOS != EOS_WB, "the WB fits stay in FP64");
    if constexpr (EOS == EOS_CONFORMAL) {
      p = e / 3.0f;
      cs2 = 0.0f;
    } else {
      tab_p_cs2_f(T, e, p, cs2);
    }
  }
}
template <int EOS, typename F>
__device__ __forceinline__ F eos_p_t(const EosTable &T, F e) {
  if constexpr (std::is_same<F, double>::value) {
    return eos_p<EOS>(T, e);
  } else {
    F p, c;
    eos_p_cs2_t<EOS, F>(T, e, p, c);
    return p;
  }
}

// effectiveTemperature, EquationOfState.cu:119-148
template <int EOS>
__device__ __forceinline__ double eos_T(const EosTable &T, double e) {
  if constexpr (EOS == EOS_CONFORMAL) {
    return pow(e / EOS_FACTOR, 0.25);
  } else if constexpr (EOS == EOS_TABLE) {
    TabPos q = tab_locate(T, e);
    const double2 v = ld_cell(T.tt, (unsigned)q.k);
    return v.x + q.fr * v.y;
  } else {
    double e2 = e * e, e3 = e2 * e, e4 = e3 * e, e5 = e4 * e, e6 = e5 * e;
    double e7 = e6 * e, e8 = e7 * e, e9 = e8 * e, e10 = e9 * e, e11 = e10 * e;
    double n = 1.510073201405604e-29 + 8.014062800678687e-18 * e +
               2.4954778310451065e-10 * e2 + 0.000063810382643387 * e3 +
               0.4873490574161924 * e4 + 207.48582344326206 * e5 +
               6686.07424325115 * e6 + 14109.766109389702 * e7 +
               1471.6180520527757 * e8 + 14.055788949565482 * e9 +
               0.015421252394182246 * e10 + 1.5780479034557783e-6 * e11;
    double d = 7.558667139355393e-28 + 1.3686372302041508e-16 * e +
               2.998130743142826e-9 * e2 + 0.0005036835870305458 * e3 +
               2.316902328874072 * e4 + 578.0778724946719 * e5 +
               11179.193315394154 * e6 + 17965.67607192861 * e7 +
               1051.0730543534657 * e8 + 5.916312075925817 * e9 +
               0.003778342768228011 * e10 + 1.8472801679382593e-7 * e11;
    return n / d;
  }
}

// bulkViscosityToEntropyDensity, TransportCoefficients.cu:27-36
__device__ __forceinline__ double zeta_over_s(double T, double tpeak) {
  double x = T / tpeak;
  if (x > 1.05) return 0.9 * exp(-(x - 1) / 0.025) + 0.25 * exp(-(x - 1) / 0.13) + 0.001;
  if (x < 0.995) return 0.9 * exp((x - 1) / 0.0025) + 0.22 * exp((x - 1) / 0.022) + 0.03;
  return -13.77 * x * x + 27.55 * x - 13.45;
}
// The same in FP32 for the mixed form's source terms: divisions by the
// constant widths as products with their reciprocals, the hardware exp.
__device__ __forceinline__ float zeta_over_s(float T, float tpeak) {
  const float x = T * __builtin_amdgcn_rcpf(tpeak), d = x - 1.0f;
  if (x > 1.05f) return 0.9f * __expf(d * -40.0f) + 0.25f * __expf(d * (-1.0f / 0.13f)) + 0.001f;
  if (x < 0.995f) return 0.9f * __expf(d * 400.0f) + 0.22f * __expf(d * (1.0f / 0.022f)) + 0.03f;
  return -13.77f * x * x + 27.55f * x - 13.45f;
}

// ------------------------------------------------------------ root-find
// energyDensityFromConservedVariables, EnergyMomentumTensor.cu:19-48:
// Newton on e with the caller's guess, stopping at |de| <= 1e-3 |e|; the
// step is damped by 0.9 after MAX_ITERS/2 iterations.  The update
// e0 - f/fp with fp = 1 - num/den is evaluated as e0 - f den/(den - num),
// and 1/e0 is shared: three divisions per iteration instead of five, equal
// to the reference's expression up to rounding.
// One Newton update of e0 given p(e0) and c_s^2(e0).
template <bool DAMPED, typename F>
__device__ __forceinline__ bool newton_update(F &e0, F p, F cs2, F M0, F M, F Pi) {
  const F ie0 = rcp_t(e0);
  const F cst2 = p * ie0;
  const F A = fma(M0, F(1) - cst2, Pi);
  const F B = fma(M0, M0 + Pi, -M);
  const F H = sqrt_t(fabs(A * A + F(4) * cst2 * B));
  const F D = div_t(A - H, F(2) * cst2);
  const F f = e0 + D;
  const F dc = cs2 - cst2;
  const F num = dc * (B + D * H - dc * cst2 * D * M0 * ie0);
  const F den = cst2 * e0 * H;
  const F step = div_t((DAMPED ? F(0.9) * f : f) * den, den - num);
  const F e = e0 - step;
  const bool done = fabs(e - e0) <= F(0.001) * fabs(e);
  e0 = e;
  return done;
}
template <int EOS, bool DAMPED, typename F = double>
__device__ __forceinline__ bool newton_step(const EosTable &T, F &e0, F M0, F M, F Pi) {
  F p, cs2;
  eos_p_cs2_t<EOS, F>(T, e0, p, cs2);
  return newton_update<DAMPED, F>(e0, p, cs2, M0, M, Pi);
}

// max_iters <= MAX_ITERS bounds the iterations (the reference: MAX_ITERS);
// the damping still starts at MAX_ITERS/2.
template <int EOS>
__device__ __forceinline__ double solve_e(const EosTable &T, double ePrev, double M0,
                                          double M, double Pi,
                                          unsigned long long *diag, int max_iters) {
  if constexpr (EOS == EOS_CONFORMAL) {
    return fabs(sqrt(fabs(4.0 * M0 * M0 - 3.0 * M)) - M0);
  } else {
    // A NaN iterate never meets the stopping rule and every later iterate is
    // NaN, so the reference's 10^6 iterations end on NaN: leaving at the
    // first NaN returns the same value.
    double e = ePrev;
    bool nan = false;
    const int lim1 = max_iters < MAX_ITERS / 2 ? max_iters : MAX_ITERS / 2;
    for (int it = 0; it < lim1 && !nan; ++it) {
      if (newton_step<EOS, false>(T, e, M0, M, Pi)) return e;
      nan = isnan(e);
    }
    for (int it = MAX_ITERS / 2; it < max_iters && !nan; ++it) {
      if (newton_step<EOS, true>(T, e, M0, M, Pi)) return e;
      nan = isnan(e);
    }
    atomicAdd(&diag[DIAG_NOT_CONVERGED], 1ull);
    return e;  // the last iterate, as the reference returns e0 after e0 = e
  }
}

// N independent root-finds sharing the guess ePrev, run in lockstep: every
// iteration evaluates all N Newton steps without branches and a chain keeps
// its iterate once it has converged, so each result equals solve_e's for
// that chain.  A wave already iterates until its slowest lane converges, so
// the instruction count matches N sequential solves while the N table
// gathers and dependency chains overlap.
template <int EOS, int N, typename F = double>
__device__ __forceinline__ void solve_e_lockstep(const EosTable &T, F ePrev, const F *M0,
                                                 const F *M, const F *Pi, F *e,
                                                 unsigned long long *diag, int max_iters) {
  if constexpr (EOS == EOS_CONFORMAL) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      const F r = fabs(F(4) * M0[k] * M0[k] - F(3) * M[k]);
      e[k] = fabs((std::is_same<F, double>::value ? sqrt(r) : sqrt_t(r)) - M0[k]);
    }
  } else {
    unsigned act = (1u << N) - 1u;
#pragma unroll
    for (int k = 0; k < N; ++k) e[k] = ePrev;
    unsigned failed = 0;  // chains that reached a NaN iterate (see solve_e)
    const int lim1 = max_iters < MAX_ITERS / 2 ? max_iters : MAX_ITERS / 2;
    for (int it = 0; it < lim1 && act; ++it) {
      // the N table gathers first, all in flight together
      F pk[N], ck[N];
#pragma unroll
      for (int k = 0; k < N; ++k) eos_p_cs2_t<EOS, F>(T, e[k], pk[k], ck[k]);
#pragma unroll
      for (int k = 0; k < N; ++k) {
        F ek = e[k];
        const bool done = newton_update<false, F>(ek, pk[k], ck[k], M0[k], M[k], Pi[k]);
        const bool live = (act >> k) & 1u;
        e[k] = live ? ek : e[k];
        if (live && done) act &= ~(1u << k);
        if (live && !done && isnan(ek)) {
          act &= ~(1u << k);
          failed |= 1u << k;
        }
      }
    }
    for (int it = MAX_ITERS / 2; it < max_iters && act; ++it) {
      // the N table gathers first, all in flight together
      F pk[N], ck[N];
#pragma unroll
      for (int k = 0; k < N; ++k) eos_p_cs2_t<EOS, F>(T, e[k], pk[k], ck[k]);
#pragma unroll
      for (int k = 0; k < N; ++k) {
        F ek = e[k];
        const bool done = newton_update<true, F>(ek, pk[k], ck[k], M0[k], M[k], Pi[k]);
        const bool live = (act >> k) & 1u;
        e[k] = live ? ek : e[k];
        if (live && done) act &= ~(1u << k);
        if (live && !done && isnan(ek)) {
          act &= ~(1u << k);
          failed |= 1u << k;
        }
      }
    }
    act |= failed;
    if (act) atomicAdd(&diag[DIAG_NOT_CONVERGED], (unsigned long long)__popc(act));
  }
}

struct Prim {
  double e, p, ut, ux, uy, un;
};

// getInferredVariables, EnergyMomentumTensor.cu:51-108 (the '#ifdef Pi'
// consistency fix at :82-85 is never compiled in the reference; omitted)
template <int EOS>
__device__ __forceinline__ Prim inferred(const EosTable &T, double t, double ttt,
                                         double ttx, double tty, double ttn,
                                         double pitt, double pitx, double pity,
                                         double pitn, double Pi, double ePrev,
                                         unsigned long long *diag, int max_iters) {
  double M0 = ttt - pitt, M1 = ttx - pitx, M2 = tty - pity, M3 = ttn - pitn;
  double M = M1 * M1 + M2 * M2 + t * t * M3 * M3;
  Prim r;
  r.e = (ePrev <= 0.1) ? M0 - M / M0 : solve_e<EOS>(T, ePrev, M0, M, Pi, diag, max_iters);
  if (isnan(r.e)) atomicAdd(&diag[DIAG_E_NAN], 1ull);
  r.p = eos_p<EOS>(T, r.e);
  if (r.e < 1.e-7) {
    r.e = 1.e-7;
    r.p = 1.e-7;
    atomicAdd(&diag[DIAG_E_CLAMPED], 1ull);
  }
  double P = r.p + Pi;
  double E = 1.0 / (r.e + P);
  r.ut = sqrt(fabs((M0 + P) * E));
  double E2 = E / r.ut;
  r.ux = M1 * E2;
  r.uy = M2 * E2;
  r.un = M3 * E2;
  return r;
}

// Flux velocity v = u_dir/u^tau of a half-site state, as getInferredVariables
// (EnergyMomentumTensor.cu:51-108) followed by u_dir/u^tau
// (SpectralRadius.cu, FluxFunctions.cu).  With E = 1/(e+P),
// u^tau = sqrt|(M0+P)E| and u_dir = M_dir E/u^tau, so
// v = M_dir E/|(M0+P)E| = sign(e+P) M_dir/|M0+P|: one division and no square
// root, equal up to rounding wherever e+P is finite and nonzero.
template <int EOS, int DIR>
__device__ __forceinline__ double flux_velocity(const EosTable &T, double t, double ttt,
                                                double ttx, double tty, double ttn, double pitt,
                                                double pitx, double pity, double pitn,
                                                double Pi, double ePrev,
                                                unsigned long long *diag, int max_iters) {
  const double M0 = ttt - pitt, M1 = ttx - pitx, M2 = tty - pity, M3 = ttn - pitn;
  const double M = M1 * M1 + M2 * M2 + t * t * M3 * M3;
  double e = (ePrev <= 0.1) ? M0 - M / M0 : solve_e<EOS>(T, ePrev, M0, M, Pi, diag, max_iters);
  if (isnan(e)) atomicAdd(&diag[DIAG_E_NAN], 1ull);
  double p = eos_p<EOS>(T, e);
  if (e < 1.e-7) {
    e = 1.e-7;
    p = 1.e-7;
    atomicAdd(&diag[DIAG_E_CLAMPED], 1ull);
  }
  const double P = p + Pi;
  const double Md = DIR == 0 ? M1 : (DIR == 1 ? M2 : M3);
  return ((e + P > 0.0) ? Md : -Md) / fabs(M0 + P);
}

// flux_velocity of N half-site states of one direction that share the
// guess ePrev, with their root-finds in lockstep (solve_e_lockstep).
// W[k] = (T^tt, T^tx, T^ty, T^tn, pi^tt, pi^tx, pi^ty, pi^tn, Pi).
// In the mixed form (R = float) the root-finds of the flux states run in FP32
// for the conformal and tabulated EoS: their results only set the flux
// velocities of FP32 fluxes.  The WB fits (powers up to e^13) overflow FP32
// and stay FP64, as does every cell's own inversion.
// ROOT64 keeps them in FP64 in every form (the mixed form's split as the
// GVH_MIXED_F64_ROOTS precision states it).
template <int EOS, typename R, bool ROOT64 = false>
using flux_root_t =
    typename std::conditional<(sizeof(R) == 4 && EOS != EOS_WB && !ROOT64), float, double>::type;

template <int EOS, int DIR, int N, typename R, bool ROOT64 = false>
__device__ __forceinline__ void flux_velocity_lockstep(const EosTable &T, double t,
                                                       const R (*W)[9], double ePrev,
                                                       unsigned long long *diag, int max_iters,
                                                       double *v) {
  using F = flux_root_t<EOS, R, ROOT64>;
  F M0[N], M[N], Md[N], Pi[N], e[N];
  const F tf = F(t);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    M0[k] = (F)W[k][0] - (F)W[k][4];
    const F M1 = (F)W[k][1] - (F)W[k][5], M2 = (F)W[k][2] - (F)W[k][6],
            M3 = (F)W[k][3] - (F)W[k][7];
    M[k] = M1 * M1 + M2 * M2 + tf * tf * M3 * M3;
    Md[k] = DIR == 0 ? M1 : (DIR == 1 ? M2 : M3);
    Pi[k] = (F)W[k][8];
  }
  if (ePrev <= 0.1) {
#pragma unroll
    for (int k = 0; k < N; ++k) e[k] = M0[k] - div_t(M[k], M0[k]);
  } else {
    solve_e_lockstep<EOS, N, F>(T, F(ePrev), M0, M, Pi, e, diag, max_iters);
  }
  // the N pressure gathers first, so that they are in flight together (an
  // atomic between them would order them)
  F pk[N];
#pragma unroll
  for (int k = 0; k < N; ++k) pk[k] = eos_p_t<EOS, F>(T, e[k]);
#pragma unroll
  for (int k = 0; k < N; ++k) {
    F ek = e[k];
    if (isnan(ek)) atomicAdd(&diag[DIAG_E_NAN], 1ull);
    F p = pk[k];
    if (ek < F(1.e-7)) {
      ek = F(1.e-7);
      p = F(1.e-7);
      atomicAdd(&diag[DIAG_E_CLAMPED], 1ull);
    }
    const F P = p + Pi[k];
    v[k] = (double)div_t((ek + P > F(0)) ? Md[k] : -Md[k], fabs(M0[k] + P));
  }
}

}  // namespace gvh
