// gvh_stage.hip -- one fused kernel per Runge-Kutta stage of the KT step.
//
// The reference runs, per stage, a Source kernel, X/Y/Z flux kernels that
// read-modify-write the output, an inferred-variables kernel, a regulation
// kernel, three ghost-cell kernels on transient streams and (stage 2) a
// convex-average kernel (FullyDiscreteKurganovTadmorScheme.cu:109-139).
// Here one thread owns one interior cell and does all of it in registers:
//
//   acc = q + dt*S                    (eulerStepKernelSource_1D, EulerStep.cu:947)
//   acc += dt*(flux diff + S_x)       (eulerStepKernelX_1D, EulerStep.cu:1012)
//   acc += dt*(flux diff + S_y)       (..Y_1D, :1101)
//   acc += dt*(flux diff + S_z)       (..Z_1D, :1190)
//   acc = (acc + q^n)/2               (stage 2, convexCombination..., :60-104)
//   (e,p,u) = root-find(acc; e_in)    (setInferredVariablesKernel)
//   pi *= regulation factor           (regulateDissipativeCurrents)
//   write acc, e, p, u; boundary cells also write their two ghost layers
//
// The summation order of the reference is kept.  Inputs and outputs are
// distinct buffers (q, e, p and u are ping-ponged by the host), so no thread
// reads what another writes within a launch.
#include "gvh_device.h"

namespace gvh {

// ------------------------------------------------------ source terms
struct Grad {
  double ut[3], ux[3], uy[3], un[3], p[3];  // [x, y, eta] derivatives
};

// loadSourceTerms2 (SourceTerms.cu:570-755) with setPimunuSourceTerms
// (:30-256): T^{tau mu} geometric and pressure-gradient sources and the
// relaxation right-hand sides of pi^{mu nu} and Pi.
template <int EOS, bool BULK>
__device__ __forceinline__ void source_terms(const StageArgs &A, const double *Q,
                                             double e, double p, double ut,
                                             double ux, double uy, double un,
                                             double utp, double uxp, double uyp,
                                             double unp, const Grad &g, double *S) {
  const double t = A.t;
  const double inv_t = 1.0 / t;
  const double t2 = t * t, t3 = t * t2;
  const double inv_t2 = 1.0 / t2;
  const double Pi = BULK ? Q[14] : 0.0;
  const double pitt = Q[4], pitx = Q[5], pity = Q[6], pitn = Q[7], pixx = Q[8],
               pixy = Q[9], pixn = Q[10], piyy = Q[11], piyn = Q[12], pinn = Q[13];
  const double inv_ut = 1.0 / ut;

  // T^{tau mu} sources, SourceTerms.cu:726-737
  const double PP = p + Pi;
  const double tnn = (e + PP) * un * un + PP * inv_t2 + pinn;
  const double vx = ux * inv_ut, vy = uy * inv_ut, vn = un * inv_ut;
  const double dkvk = (g.ux[0] - vx * g.ut[0]) * inv_ut + (g.uy[1] - vy * g.ut[1]) * inv_ut +
                      (g.un[2] - vn * g.ut[2]) * inv_ut;
  S[0] = -(Q[0] * inv_t + t * tnn) + dkvk * (pitt - p - Pi) - vx * g.p[0] - vy * g.p[1] -
         vn * g.p[2];
  S[1] = -Q[1] * inv_t - g.p[0] + dkvk * pitx;
  S[2] = -Q[2] * inv_t - g.p[1] + dkvk * pity;
  S[3] = -3.0 * Q[3] * inv_t - g.p[2] * inv_t2 + dkvk * pitn;

  // transport coefficients, SourceTerms.cu:46-60
  const double T = eos_T<EOS>(A.tab, e);
  const double taupiInv = (double)0.2f * (T / A.etabar);
  const double beta_pi = (double)0.2f * (e + p);

  // backward-difference time derivative and covariant acceleration, :67-86
  const double dtut = (ut - utp) * A.inv_dt, dtux = (ux - uxp) * A.inv_dt;
  const double dtuy = (uy - uyp) * A.inv_dt, dtun = (un - unp) * A.inv_dt;
  const double un2 = un * un, ut2 = ut * ut;
  const double dut = ut * dtut + ux * g.ut[0] + uy * g.ut[1] + un * g.ut[2];
  const double Dut = dut + t * un2;
  const double dux = ut * dtux + ux * g.ux[0] + uy * g.ux[1] + un * g.ux[2];
  const double duy = ut * dtuy + ux * g.uy[0] + uy * g.uy[1] + un * g.uy[2];
  const double dun = ut * dtun + ux * g.un[0] + uy * g.un[1] + un * g.un[2];
  const double Dux = -dux, Duy = -duy;
  const double Dun = -t2 * dun - 2.0 * t * ut * un;

  const double theta = ut * inv_t + dtut + g.ux[0] + g.uy[1] + g.un[2];
  const double th3 = theta * (1.0 / 3.0);

  // sigma^{mu nu}, :96-106
  const double stt = -t * ut * un2 + (dtut - ut * dut) + (ut2 - 1.0) * th3;
  const double stx = -(t * un2 * ux) * 0.5 + (dtux - g.ut[0]) * 0.5 -
                     (ux * dut + ut * dux) * 0.5 + ut * ux * th3;
  const double sty = -(t * un2 * uy) * 0.5 + (dtuy - g.ut[1]) * 0.5 -
                     (uy * dut + ut * duy) * 0.5 + ut * uy * th3;
  const double stn = -un * (2.0 * ut2 + t2 * un2) * (0.5 * inv_t) +
                     (dtun - g.ut[2] * inv_t2) * 0.5 - (un * dut + ut * dun) * 0.5 +
                     ut * un * th3;
  const double sxx = -(g.ux[0] + ux * dux) + (1.0 + ux * ux) * th3;
  const double sxy = -(g.uy[0] + g.ux[1]) * 0.5 - (uy * dux + ux * duy) * 0.5 + ux * uy * th3;
  const double sxn = -ut * ux * un * inv_t - (g.un[0] + g.ux[2] * inv_t2) * 0.5 -
                     (un * dux + ux * dun) * 0.5 + ux * un * th3;
  const double syy = -(g.uy[1] + uy * duy) + (1.0 + uy * uy) * th3;
  const double syn = -ut * uy * un * inv_t - (g.un[1] + g.uy[2] * inv_t2) * 0.5 -
                     (un * duy + uy * dun) * 0.5 + uy * un * th3;
  const double snn = -ut * (1.0 + 2.0 * t2 * un2) / t3 - g.un[2] * inv_t2 - un * dun +
                     (inv_t2 + un2) * th3;

  // omega^mu_nu, :123-136
  const double wtx = (-dtux - g.ut[0] + ut * dux - ux * dut - t * un2 * ux) * 0.5;
  const double wty = (-dtuy - g.ut[1] + ut * duy - uy * dut - t * un2 * uy) * 0.5;
  const double wtn = (-t2 * dtun - 2.0 * t * un - g.ut[2] - ut * Dun - t2 * un * dut -
                      t3 * un * un2) * 0.5;
  const double wxy = (g.uy[0] - g.ux[1] + ux * duy - uy * dux) * 0.5;
  const double wxn = (t2 * g.un[0] - g.ux[2] - ux * Dun - t2 * un * dux) * 0.5;
  const double wyn = (t2 * g.un[1] - g.uy[2] - uy * Dun - t2 * un * duy) * 0.5;
  const double wxt = wtx, wyt = wty, wnt = wtn * inv_t2, wyx = -wxy;
  const double wnx = -wxn * inv_t2, wny = -wyn * inv_t2;

  // pi^{mu lambda} D u_lambda for each row mu, then I1 = u^mu a^nu + u^nu a^mu
  const double at = pitt * Dut + pitx * Dux + pity * Duy + pitn * Dun;
  const double ax = pitx * Dut + pixx * Dux + pixy * Duy + pixn * Dun;
  const double ay = pity * Dut + pixy * Dux + piyy * Duy + piyn * Dun;
  const double an = pitn * Dut + pixn * Dux + piyn * Duy + pinn * Dun;

  // pi:sigma, :185-187
  const double ps = pitt * stt - 2.0 * pitx * stx - 2.0 * pity * sty + pixx * sxx +
                    2.0 * pixy * sxy + piyy * syy - 2.0 * pitn * stn * t2 +
                    2.0 * pixn * sxn * t2 + 2.0 * piyn * syn * t2 + pinn * snn * t2 * t2;
  const double ps3 = ps * (1.0 / 3.0);

  constexpr double d_pp = 1.33333, t_pp = 1.42857, l_pP = 1.2;  // TransportCoefficients.cuh:14-16
  const double lPi = l_pP * Pi;

  // I = I1 + delta I2 - I3 + tau I4 - lambda Pi sigma, relaxation RHS with
  // Milne Christoffel terms, :141-233; RHS/u^tau + pi dkvk, :243-252
  const double b2 = 2.0 * beta_pi;
  {
    const double I1 = 2.0 * ut * at;
    const double I3 = 2.0 * (pitx * wtx + pity * wty + pitn * wtn);
    const double I4 = (pitt * stt - pitx * stx - pity * sty - t2 * pitn * stn) - (1.0 - ut2) * ps3;
    const double I = I1 + d_pp * theta * pitt - I3 + t_pp * I4 - lPi * stt;
    const double r = b2 * stt - pitt * taupiInv - I - 2.0 * un * t * pitn;
    S[4] = r * inv_ut + pitt * dkvk;
  }
  {
    const double I1 = ux * at + ut * ax;
    const double I3 = pitt * wxt + pity * wxy + pitn * wxn + pixx * wtx + pixy * wty + pixn * wtn;
    const double I4 = (pitt * stx + pitx * stt) * 0.5 - (pitx * sxx + pixx * stx) * 0.5 -
                      (pity * sxy + pixy * sty) * 0.5 - t2 * (pitn * sxn + pixn * stn) * 0.5 +
                      (ut * ux) * ps3;
    const double I = I1 + d_pp * theta * pitx - I3 + t_pp * I4 - lPi * stx;
    const double r = b2 * stx - pitx * taupiInv - I - un * t * pixn;
    S[5] = r * inv_ut + pitx * dkvk;
  }
  {
    const double I1 = uy * at + ut * ay;
    const double I3 = pitt * wyt + pitx * wyx + pitn * wyn + pixy * wtx + piyy * wty + piyn * wtn;
    const double I4 = (pitt * sty + pity * stt) * 0.5 - (pitx * sxy + pixy * stx) * 0.5 -
                      (pity * syy + piyy * sty) * 0.5 - t2 * (pitn * syn + piyn * stn) * 0.5 +
                      (ut * uy) * ps3;
    const double I = I1 + d_pp * theta * pity - I3 + t_pp * I4 - lPi * sty;
    const double r = b2 * sty - pity * taupiInv - I - un * t * piyn;
    S[6] = r * inv_ut + pity * dkvk;
  }
  {
    const double I1 = un * at + ut * an;
    const double I3 = pitt * wnt + pitx * wnx + pity * wny + pixn * wtx + piyn * wty + pinn * wtn;
    const double I4 = (pitt * stn + pitn * stt) * 0.5 - (pitx * sxn + pixn * stx) * 0.5 -
                      (pity * syn + piyn * sty) * 0.5 - t2 * (pitn * snn + pinn * stn) * 0.5 +
                      (ut * un) * ps3;
    const double I = I1 + d_pp * theta * pitn - I3 + t_pp * I4 - lPi * stn;
    const double r = b2 * stn - pitn * taupiInv - I - un * t * pinn -
                     (ut * pitn + un * pitt) * inv_t;
    S[7] = r * inv_ut + pitn * dkvk;
  }
  {
    const double I1 = 2.0 * ux * ax;
    const double I3 = 2.0 * (pitx * wxt + pixy * wxy + pixn * wxn);
    const double I4 = (pitx * stx - pixx * sxx - pixy * sxy - t2 * pixn * sxn) + (1.0 + ux * ux) * ps3;
    const double I = I1 + d_pp * theta * pixx - I3 + t_pp * I4 - lPi * sxx;
    const double r = b2 * sxx - pixx * taupiInv - I;
    S[8] = r * inv_ut + pixx * dkvk;
  }
  {
    const double I1 = uy * ax + ux * ay;
    const double I3 = pitx * wyt + pity * wxt + pixx * wyx + piyy * wxy + pixn * wyn + piyn * wxn;
    const double I4 = (pitx * sty + pity * stx) * 0.5 - (pixx * sxy + pixy * sxx) * 0.5 -
                      (pixy * syy + piyy * sxy) * 0.5 - t2 * (pixn * syn + piyn * sxn) * 0.5 +
                      (ux * uy) * ps3;
    const double I = I1 + d_pp * theta * pixy - I3 + t_pp * I4 - lPi * sxy;
    const double r = b2 * sxy - pixy * taupiInv - I;
    S[9] = r * inv_ut + pixy * dkvk;
  }
  {
    const double I1 = un * ax + ux * an;
    const double I3 = pitx * wnt + pitn * wxt + pixx * wnx + pixy * wny + piyn * wxy + pinn * wxn;
    const double I4 = (pitx * stn + pitn * stx) * 0.5 - (pixx * sxn + pixn * sxx) * 0.5 -
                      (pixy * syn + piyn * sxy) * 0.5 - t2 * (pixn * snn + pinn * sxn) * 0.5 +
                      (ux * un) * ps3;
    const double I = I1 + d_pp * theta * pixn - I3 + t_pp * I4 - lPi * sxn;
    const double r = b2 * sxn - pixn * taupiInv - I - (ut * pixn + un * pitx) * inv_t;
    S[10] = r * inv_ut + pixn * dkvk;
  }
  {
    const double I1 = 2.0 * uy * ay;
    const double I3 = 2.0 * (pity * wyt + pixy * wyx + piyn * wyn);
    const double I4 = (pity * sty - pixy * sxy - piyy * syy - t2 * piyn * syn) + (1.0 + uy * uy) * ps3;
    const double I = I1 + d_pp * theta * piyy - I3 + t_pp * I4 - lPi * syy;
    const double r = b2 * syy - piyy * taupiInv - I;
    S[11] = r * inv_ut + piyy * dkvk;
  }
  {
    const double I1 = un * ay + uy * an;
    const double I3 = pity * wnt + pitn * wyt + pixy * wnx + pixn * wyx + piyy * wny + pinn * wyn;
    const double I4 = (pity * stn + pitn * sty) * 0.5 - (pixy * sxn + pixn * sxy) * 0.5 -
                      (piyy * syn + piyn * syy) * 0.5 - t2 * (piyn * snn + pinn * syn) * 0.5 +
                      (uy * un) * ps3;
    const double I = I1 + d_pp * theta * piyn - I3 + t_pp * I4 - lPi * syn;
    const double r = b2 * syn - piyn * taupiInv - I - (ut * piyn + un * pity) * inv_t;
    S[12] = r * inv_ut + piyn * dkvk;
  }
  {
    const double I1 = 2.0 * un * an;
    const double I3 = 2.0 * (pitn * wnt + pixn * wnx + piyn * wny);
    const double I4 = (pitn * stn - pixn * sxn - piyn * syn - t2 * pinn * snn) + (inv_t2 + un2) * ps3;
    const double I = I1 + d_pp * theta * pinn - I3 + t_pp * I4 - lPi * snn;
    const double r = b2 * snn - pinn * taupiInv - I - 2.0 * (ut * pinn + un * pitn) * inv_t;
    S[13] = r * inv_ut + pinn * dkvk;
  }
  if constexpr (BULK) {
    // bulk relaxation, :53-60 and :238, :254
    double p_, cs2;
    eos_p_cs2<EOS>(A.tab, e, p_, cs2);
    const double a = (double)0.333333f - cs2;
    const double a2 = a * a;
    const double beta_Pi = 15.0 * a2 * (e + p);
    const double lambda_Pipi = (double)1.6f * a;
    const double tauPiInv = 15.0 * a2 * (T / zeta_over_s(T, A.zeta_tpeak));
    constexpr double d_PP = 0.666667;  // TransportCoefficients.cuh:18
    const double dPi = -beta_Pi * theta - Pi * tauPiInv - d_PP * Pi * theta + lambda_Pipi * ps;
    S[14] = dPi * inv_ut + Pi * dkvk;
  }
}

// ------------------------------------------------- directional update
// Threads per block of the stage kernel; the per-thread accumulator of the 15
// conserved variables lives in LDS as acc[n * BLOCK + threadIdx.x] so that the
// variable loops can stay rolled (dynamic index) and registers stay few.
constexpr int BLOCK = 256;

// One direction of eulerStepKernel{X,Y,Z}_1D: the two KT interface fluxes of
// flux() (SemiDiscreteKurganovTadmorScheme.cu:16-57), each with both
// half-site states inverted using this cell's e as the guess, plus the
// central-difference pi / Pi gradient sources of loadSourceTerms{X,Y,Z}
// (SourceTerms.cu:444-567).  acc[n] += ((H_b - H_f)/h + S_dir) dt.
//
// Half-site states (HalfSiteExtrapolation.cu:16-29): right/left at i+1/2
// are q_{i+1} - D_{i+1}/2 and q_i + D_i/2, right/left at i-1/2 are
// q_i - D_i/2 and q_{i-1} + D_{i-1}/2.  The nine variables the inversion
// needs (T^{tau mu}, pi^{tau mu}, Pi) are reconstructed state by state for
// the four solves, then every variable once more for its fluxes.
// Every stencil load is a uniform (scalar) base plus the thread's 32-bit
// cell offset.
template <int EOS, bool BULK, int DIR, bool KEEP_STATES>
__device__ __forceinline__ void direction_update(const StageArgs &A, unsigned s, double ePrev,
                                                 double v_s, double *acc) {
  constexpr int NV = BULK ? 15 : 14;
  const long long st = DIR == 0 ? 1 : (DIR == 1 ? (long long)A.ncx : (long long)A.ncx * A.ncy);
  const double *__restrict__ q = A.qin;
  const long long G = A.G;
  const double t = A.t;
  const double hc = A.half_inv_h[DIR];

  double vRf, vLf, vRb, vLb;
  double Rf[9], Lf[9], Rb[9], Lb[9];  // used when KEEP_STATES
  double cdv[5];
  if constexpr (KEEP_STATES) {
    // all four states of the nine inversion variables stay in registers
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      if (a == 8 && !BULK) {
        Rf[8] = Lf[8] = Rb[8] = Lb[8] = cdv[4] = 0.0;
        continue;
      }
      const int n = a < 8 ? a : 14;
      const double *b = q + n * G;
      const double vm2 = (b - 2 * st)[s], vm1 = (b - st)[s], v0 = b[s], vp1 = (b + st)[s],
                   vp2 = (b + 2 * st)[s];
      const double D0 = slope(vm1, v0, vp1);
      Rf[a] = vp1 - slope(v0, vp1, vp2) * 0.5;
      Lf[a] = v0 + D0 * 0.5;
      Rb[a] = v0 - D0 * 0.5;
      Lb[a] = vm1 + slope(vm2, vm1, v0) * 0.5;
      if (a >= 4) cdv[a - 4] = (vp1 - vm1) * hc;
    }
    vRf = flux_velocity<EOS, DIR>(A.tab, t, Rf[0], Rf[1], Rf[2], Rf[3], Rf[4], Rf[5], Rf[6],
                                  Rf[7], Rf[8], ePrev, A.diag);
    vLf = flux_velocity<EOS, DIR>(A.tab, t, Lf[0], Lf[1], Lf[2], Lf[3], Lf[4], Lf[5], Lf[6],
                                  Lf[7], Lf[8], ePrev, A.diag);
    vRb = flux_velocity<EOS, DIR>(A.tab, t, Rb[0], Rb[1], Rb[2], Rb[3], Rb[4], Rb[5], Rb[6],
                                  Rb[7], Rb[8], ePrev, A.diag);
    vLb = flux_velocity<EOS, DIR>(A.tab, t, Lb[0], Lb[1], Lb[2], Lb[3], Lb[4], Lb[5], Lb[6],
                                  Lb[7], Lb[8], ePrev, A.diag);
  } else {
    // the four states one after another (rolled: few live registers)
    vRf = vLf = vRb = vLb = 0.0;
#pragma unroll 1
    for (int state = 0; state < 4; ++state) {
      const long long o = state == 0 ? st : (state == 3 ? -st : 0);
      const double sg = (state & 1) ? 0.5 : -0.5;
      double w[9];
#pragma unroll
      for (int a = 0; a < 9; ++a) {
        if (a == 8 && !BULK) {
          w[8] = 0.0;
          continue;
        }
        const int n = a < 8 ? a : 14;
        const double *b = q + n * G + o;
        const double xm = (b - st)[s], x0 = b[s], xp = (b + st)[s];
        w[a] = x0 + slope(xm, x0, xp) * sg;
      }
      const double v = flux_velocity<EOS, DIR>(A.tab, t, w[0], w[1], w[2], w[3], w[4], w[5], w[6],
                                               w[7], w[8], ePrev, A.diag);
      vRf = state == 0 ? v : vRf;
      vLf = state == 1 ? v : vLf;
      vRb = state == 2 ? v : vRb;
      vLb = state == 3 ? v : vLb;
    }
#pragma unroll
    for (int a = 4; a < 9; ++a) {
      if (a == 8 && !BULK) {
        cdv[4] = 0.0;
        continue;
      }
      const int n = a < 8 ? a : 14;
      cdv[a - 4] = ((q + n * G + st)[s] - (q + n * G - st)[s]) * hc;
    }
  }
  // local propagation speeds, max(|v_L|, |v_R|) (LocalPropagationSpeed.cu:16-24)
  const double aF = fmax(fabs(vLf), fabs(vRf));
  const double aB = fmax(fabs(vLb), fabs(vRb));

  // directional sources: central differences of pi^{tau mu} and Pi and of
  // the three spatial pi components of this direction
  auto cd = [&](int n) { return ((q + n * G + st)[s] - (q + n * G - st)[s]) * hc; };
  const double cPi = cdv[4];
  double Sd0, Sd1, Sd2, Sd3;
  if constexpr (DIR == 0) {
    Sd0 = cdv[0] * v_s - cdv[1] - (BULK ? v_s * cPi : 0.0);
    Sd1 = cdv[1] * v_s - cd(8) - cPi;
    Sd2 = cdv[2] * v_s - cd(9);
    Sd3 = cdv[3] * v_s - cd(10);
  } else if constexpr (DIR == 1) {
    Sd0 = cdv[0] * v_s - cdv[2] - (BULK ? v_s * cPi : 0.0);
    Sd2 = cdv[2] * v_s - cd(11) - cPi;
    Sd1 = cdv[1] * v_s - cd(9);
    Sd3 = cdv[3] * v_s - cd(12);
  } else {
    Sd0 = cdv[0] * v_s - cdv[3] - (BULK ? v_s * cPi : 0.0);
    Sd3 = cdv[3] * v_s - cd(13) - (BULK ? cPi / (t * t) : 0.0);
    Sd1 = cdv[1] * v_s - cd(10);
    Sd2 = cdv[2] * v_s - cd(12);
  }

  const double ih = A.inv_h[DIR], dt = A.dt;
  // H = (F(q_R) + F(q_L) - a (q_R - q_L))/2 with F = v q
#define GVH_FLUXDIFF(rf, lf, rb, lb)                                         \
  (((vRb * (rb) + vLb * (lb) - aB * ((rb) - (lb))) * 0.5 -                   \
    (vRf * (rf) + vLf * (lf) - aF * ((rf) - (lf))) * 0.5) * ih)
  int n0 = 0;
  if constexpr (KEEP_STATES) {
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      if (a == 8 && !BULK) continue;
      const int n = a < 8 ? a : 14;
      double r = GVH_FLUXDIFF(Rf[a], Lf[a], Rb[a], Lb[a]);
      if (a == 0) r += Sd0;
      if (a == 1) r += Sd1;
      if (a == 2) r += Sd2;
      if (a == 3) r += Sd3;
      acc[n * BLOCK] += r * dt;
    }
    n0 = 8;
  }
  const int n1 = KEEP_STATES ? 14 : NV;
#pragma unroll 1
  for (int n = n0; n < n1; ++n) {
    const double *b = q + n * G;
    const double vm2 = (b - 2 * st)[s], vm1 = (b - st)[s], v0 = b[s], vp1 = (b + st)[s],
                 vp2 = (b + 2 * st)[s];
    const double D0 = slope(vm1, v0, vp1);
    const double rf = vp1 - slope(v0, vp1, vp2) * 0.5;
    const double lf = v0 + D0 * 0.5;
    const double rb = v0 - D0 * 0.5;
    const double lb = vm1 + slope(vm2, vm1, v0) * 0.5;
    double r = GVH_FLUXDIFF(rf, lf, rb, lb);
    if (n < 4) r += n == 0 ? Sd0 : (n == 1 ? Sd1 : (n == 2 ? Sd2 : Sd3));
    acc[n * BLOCK] += r * dt;
  }
#undef GVH_FLUXDIFF
}

// Cell of a thread.  Traversal order: x fastest, then the rows of a strip
// of A.strip rows in y, then eta, then the strips -- so the eta neighbours
// of a cell (+-2 planes) are a few strips-of-rows away in launch order rather
// than a whole x-y plane, and stay in L2.  Blocks are renumbered so that each
// of the 8 XCDs (which receive consecutive workgroups round-robin) gets one
// contiguous run of that order.  The grid is a multiple of 8 blocks; threads
// past the lattice (or in the padding rows of the last strip) exit.
struct CellPos {
  int i, j, k;
  unsigned s;
  bool valid;
};

__device__ __forceinline__ CellPos cell_of_thread(const StageArgs &A) {
  const unsigned nb = gridDim.x, b = blockIdx.x;
  const unsigned lb = (b & 7u) * (nb >> 3) + (b >> 3);
  const int by = A.strip;
  const long long tid = (long long)lb * BLOCK + threadIdx.x;
  const int i = (int)(tid % A.nx);
  const long long r = tid / A.nx;
  const int jj = (int)(r % by);
  const long long r2 = r / by;
  const int k = (int)(r2 % A.nz);
  const int yb = (int)(r2 / A.nz);
  const int j = yb * by + jj;
  CellPos c;
  c.valid = j < A.ny && yb < A.nstrips;
  c.i = i + NG;
  c.j = (c.valid ? j : 0) + NG;
  c.k = k + NG;
  c.s = (unsigned)(c.i + (long long)A.ncx * (c.j + (long long)A.ncy * c.k));
  return c;
}

// Source part of the stage (eulerStepKernelSource_1D, EulerStep.cu:947-1009):
// acc = q + dt S written to the stage's output buffer at the cell.
template <int EOS, bool BULK, bool HAS_Z>
__global__ __launch_bounds__(BLOCK) void kt_source_kernel(StageArgs A) {
  constexpr int NV = BULK ? 15 : 14;
  const CellPos c = cell_of_thread(A);
  if (!c.valid) return;
  const unsigned s = c.s;
  const long long sy = A.ncx, sz = (long long)A.ncx * A.ncy;
  const long long G = A.G;
  const double *u = A.uin;
  const double ut = u[s], ux = (u + G)[s], uy = (u + 2 * G)[s], un = (u + 3 * G)[s];
  const double e_s = A.ein[s], p_s = A.pin[s];
  // minmod gradients of u and p (SourceTerms.cu:662-719)
  Grad g;
  const long long st[3] = {1, sy, sz};
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (d == 2 && !HAS_Z) {
      g.ut[2] = g.ux[2] = g.uy[2] = g.un[2] = g.p[2] = 0.0;
      continue;
    }
    const double ih = A.inv_h[d];
    const double *um = u - st[d], *upl = u + st[d];
    g.ut[d] = slope(um[s], ut, upl[s]) * ih;
    g.ux[d] = slope((um + G)[s], ux, (upl + G)[s]) * ih;
    g.uy[d] = slope((um + 2 * G)[s], uy, (upl + 2 * G)[s]) * ih;
    g.un[d] = slope((um + 3 * G)[s], un, (upl + 3 * G)[s]) * ih;
    g.p[d] = slope((A.pin - st[d])[s], p_s, (A.pin + st[d])[s]) * ih;
  }
  double Q[15], S[15];
#pragma unroll
  for (int n = 0; n < 15; ++n) Q[n] = (n < NV) ? (A.qin + n * G)[s] : 0.0;
  S[14] = 0.0;
  const double *up = A.uprev;
  source_terms<EOS, BULK>(A, Q, e_s, p_s, ut, ux, uy, un, up[s], (up + G)[s], (up + 2 * G)[s],
                          (up + 3 * G)[s], g, S);
#pragma unroll
  for (int n = 0; n < NV; ++n) (A.qout + n * G)[s] = Q[n] + A.dt * S[n];
}

// Flux part of the stage plus the cell's finishing work: the three
// directional updates in the reference's order, the convex average (stage
// 2), the root-find, regulation, and the writes of the cell and its ghost
// copies.  acc starts from what kt_source_kernel stored at the cell.
// FUSED: the source terms are computed here first (one kernel per stage);
// otherwise kt_source_kernel has stored acc = q + dt S at the cell.
template <int EOS, bool BULK, bool STAGE2, bool HAS_Z, bool FUSED, bool KEEP>
__global__ __launch_bounds__(BLOCK) void kt_flux_kernel(StageArgs A) {
  constexpr int NV = BULK ? 15 : 14;
  __shared__ double acc_lds[15 * BLOCK];
  double *acc = acc_lds + threadIdx.x;
  const CellPos c = cell_of_thread(A);
  if (!c.valid) return;
  const int i = c.i, j = c.j, k = c.k;
  const unsigned s = c.s;
  const long long sy = A.ncx, sz = (long long)A.ncx * A.ncy;
  const long long G = A.G;
  const double *u = A.uin;
  const double ut = u[s], ux = (u + G)[s], uy = (u + 2 * G)[s], un = (u + 3 * G)[s];
  const double e_s = A.ein[s];
  if constexpr (FUSED) {
    const double p_s = A.pin[s];
    Grad g;
    const long long st[3] = {1, sy, sz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (d == 2 && !HAS_Z) {
        g.ut[2] = g.ux[2] = g.uy[2] = g.un[2] = g.p[2] = 0.0;
        continue;
      }
      const double ih = A.inv_h[d];
      const double *um = u - st[d], *upl = u + st[d];
      g.ut[d] = slope(um[s], ut, upl[s]) * ih;
      g.ux[d] = slope((um + G)[s], ux, (upl + G)[s]) * ih;
      g.uy[d] = slope((um + 2 * G)[s], uy, (upl + 2 * G)[s]) * ih;
      g.un[d] = slope((um + 3 * G)[s], un, (upl + 3 * G)[s]) * ih;
      g.p[d] = slope((A.pin - st[d])[s], p_s, (A.pin + st[d])[s]) * ih;
    }
    double Q[15], S[15];
#pragma unroll
    for (int n = 0; n < 15; ++n) Q[n] = (n < NV) ? (A.qin + n * G)[s] : 0.0;
    S[14] = 0.0;
    const double *up = A.uprev;
    source_terms<EOS, BULK>(A, Q, e_s, p_s, ut, ux, uy, un, up[s], (up + G)[s],
                            (up + 2 * G)[s], (up + 3 * G)[s], g, S);
#pragma unroll
    for (int n = 0; n < 15; ++n) acc[n * BLOCK] = Q[n] + A.dt * S[n];
  } else {
#pragma unroll
    for (int n = 0; n < NV; ++n) acc[n * BLOCK] = (A.qout + n * G)[s];
  }

  // flux velocity of the cell itself for the directional sources (v = u/u^tau)
  direction_update<EOS, BULK, 0, KEEP>(A, s, e_s, ux / ut, acc);
  direction_update<EOS, BULK, 1, KEEP>(A, s, e_s, uy / ut, acc);
  if constexpr (HAS_Z) direction_update<EOS, BULK, 2, KEEP>(A, s, e_s, un / ut, acc);

  if constexpr (STAGE2) {
#pragma unroll 1
    for (int n = 0; n < NV; ++n) acc[n * BLOCK] = (acc[n * BLOCK] + (A.qn + n * G)[s]) * 0.5;
  }

  // inferred variables with this cell's e as the guess
  const Prim pr = inferred<EOS>(A.tab, A.tinf, acc[0], acc[BLOCK], acc[2 * BLOCK], acc[3 * BLOCK],
                                acc[4 * BLOCK], acc[5 * BLOCK], acc[6 * BLOCK], acc[7 * BLOCK],
                                BULK ? acc[14 * BLOCK] : 0.0, e_s, A.diag);

  // regulateDissipativeCurrents (RegulateDissipativeCurrents.cu:16-134): only
  // cells at or below e_reg are rescaled; the bulk factor is 1 unless p is 0
  if (!(pr.e > A.e_reg)) {
    const double t = A.tinf, t2 = t * t;
    const double pitt = acc[4 * BLOCK], pitx = acc[5 * BLOCK], pity = acc[6 * BLOCK],
                 pitn = acc[7 * BLOCK], pixx = acc[8 * BLOCK], pixy = acc[9 * BLOCK],
                 pixn = acc[10 * BLOCK], piyy = acc[11 * BLOCK], piyn = acc[12 * BLOCK],
                 pinn = acc[13 * BLOCK];
    const double pipi = pitt * pitt - 2 * pitx * pitx - 2 * pity * pity + pixx * pixx +
                        2 * pixy * pixy + piyy * piyy - 2 * pitn * pitn * t2 +
                        2 * pixn * pixn * t2 + 2 * piyn * piyn * t2 + pinn * pinn * t2 * t2;
    const double spipi = sqrt(fabs(pipi));
    const double pimumu = pitt - pixx - piyy - pinn * t2;
    const double piu0 = -pitn * t2 * pr.un + pitt * pr.ut - pitx * pr.ux - pity * pr.uy;
    const double piu1 = -pixn * t2 * pr.un + pitx * pr.ut - pixx * pr.ux - pixy * pr.uy;
    const double piu2 = -piyn * t2 * pr.un + pity * pr.ut - pixy * pr.ux - piyy * pr.uy;
    const double piu3 = -pinn * t2 * pr.un + pitn * pr.ut - pixn * pr.ux - piyn * pr.uy;
    const double a1 = spipi * (1.0 / sqrt(pr.e * pr.e + 3 * pr.p * pr.p));  // rho_max = 1
    const double den = 0.1 * spipi;                                          // xi0 rho_max
    const double rho = fmax(a1, fmax(pimumu / den, fmax(piu0 / den, fmax(piu1 / den,
                       fmax(piu2 / den, piu3 / den)))));
    double fs = tanh(rho) / rho;
    if (fabs(rho) < 1.e-7) fs = 1.0;
#pragma unroll 1
    for (int n = 4; n < 14; ++n) acc[n * BLOCK] *= fs;
    if constexpr (BULK) {
      const double rb = 0.0 / pr.p;
      double fb = tanh(rb) / rb;
      if (fabs(rb) < 1.e-7) fb = 1.0;
      acc[14 * BLOCK] *= fb;
    }
  }

  // write the cell and, on the boundary, its two ghost layers
  // (GhostCells.cu:15-144: zero-gradient copy, edges and corners untouched)
  // offsets of this cell's ghost copies: two per lattice face it touches (a
  // cell of a one-point axis touches both faces of that axis)
  long long gh[12];
  int ng = 0;
  if (i == NG) { gh[ng++] = -1; gh[ng++] = -2; }
  if (i == A.nx + 1) { gh[ng++] = 1; gh[ng++] = 2; }
  if (j == NG) { gh[ng++] = -sy; gh[ng++] = -2 * sy; }
  if (j == A.ny + 1) { gh[ng++] = sy; gh[ng++] = 2 * sy; }
  if (k == NG) { gh[ng++] = -sz; gh[ng++] = -2 * sz; }
  if (k == A.nz + 1) { gh[ng++] = sz; gh[ng++] = 2 * sz; }

  auto put = [&](double *base, double val) {
    base[s] = val;
    for (int m = 0; m < ng; ++m) base[(long long)s + gh[m]] = val;
  };
#pragma unroll 1
  for (int n = 0; n < NV; ++n) put(A.qout + n * G, acc[n * BLOCK]);
  put(A.eout, pr.e);
  put(A.pout, pr.p);
  put(A.uout, pr.ut);
  put(A.uout + G, pr.ux);
  put(A.uout + 2 * G, pr.uy);
  put(A.uout + 3 * G, pr.un);
}

// ------------------------------------------------------------ launch
// Kernel arrangement of a stage (StageArgs::variant): 0 fused kernel with
// the inversion states kept in registers, 1 fused with states rebuilt one at
// a time, 2 separate source kernel + flux kernel (states rebuilt).
template <int EOS, bool BULK, bool STAGE2, bool HAS_Z>
static hipError_t launch_z(const StageArgs &A, hipStream_t st) {
  const long long ncell = (long long)A.nx * A.strip * A.nstrips * A.nz;
  const int block = BLOCK;
  const int grid = (int)(((ncell + block - 1) / block + 7) / 8 * 8);  // multiple of 8 (XCD order)
  switch (A.variant) {
    case 0:
      hipLaunchKernelGGL((kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, true, true>), dim3(grid),
                         dim3(block), 0, st, A);
      break;
    case 1:
      hipLaunchKernelGGL((kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, true, false>), dim3(grid),
                         dim3(block), 0, st, A);
      break;
    default:
      hipLaunchKernelGGL((kt_source_kernel<EOS, BULK, HAS_Z>), dim3(grid), dim3(block), 0, st, A);
      hipLaunchKernelGGL((kt_flux_kernel<EOS, BULK, STAGE2, HAS_Z, false, false>), dim3(grid),
                         dim3(block), 0, st, A);
  }
  return hipGetLastError();
}

template <int EOS, bool BULK, bool STAGE2>
static hipError_t launch_t(const StageArgs &A, bool has_z, hipStream_t st) {
  return has_z ? launch_z<EOS, BULK, STAGE2, true>(A, st) : launch_z<EOS, BULK, STAGE2, false>(A, st);
}

template <int EOS>
static hipError_t launch_e(const StageArgs &A, bool bulk, bool stage2, bool has_z, hipStream_t st) {
  if (bulk) return stage2 ? launch_t<EOS, true, true>(A, has_z, st) : launch_t<EOS, true, false>(A, has_z, st);
  return stage2 ? launch_t<EOS, false, true>(A, has_z, st) : launch_t<EOS, false, false>(A, has_z, st);
}

hipError_t launch_stage_fp64(const StageArgs &A, int eos, bool bulk, bool stage2, hipStream_t st) {
  const bool has_z = A.nz > 1;
  switch (eos) {
    case EOS_CONFORMAL: return launch_e<EOS_CONFORMAL>(A, bulk, stage2, has_z, st);
    case EOS_WB: return launch_e<EOS_WB>(A, bulk, stage2, has_z, st);
    default: return launch_e<EOS_TABLE>(A, bulk, stage2, has_z, st);
  }
}

// Termination scan of the run loop (HydroPlugin.cpp:396-407): count interior
// cells with e > threshold; one atomic per block.
__global__ __launch_bounds__(256) void count_above_kernel(const double *__restrict__ e, int nx,
                                                          int ny, int nz, int ncx, int ncy,
                                                          double thr,
                                                          unsigned long long *__restrict__ out) {
  __shared__ unsigned int wsum[4];
  const int nint = nx * ny * nz;
  unsigned int c = 0;
  for (int tid = blockIdx.x * blockDim.x + threadIdx.x; tid < nint; tid += gridDim.x * blockDim.x) {
    const int i = tid % nx + NG, r = tid / nx, j = r % ny + NG, k = r / ny + NG;
    c += e[i + (long long)ncx * (j + (long long)ncy * k)] > thr ? 1u : 0u;
  }
  // wavefront reduction (64 lanes) then across the 4 waves of the block
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) wsum[w] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tot = 0;
    for (int m = 0; m < (int)(blockDim.x >> 6); ++m) tot += wsum[m];
    if (tot) atomicAdd(out, tot);
  }
}

hipError_t launch_count_above(const double *e, int nx, int ny, int nz, int ncx, int ncy, double thr,
                              unsigned long long *out, hipStream_t st) {
  const int nint = nx * ny * nz;
  int grid = (nint + 255) / 256;
  if (grid > 2048) grid = 2048;
  hipLaunchKernelGGL(count_above_kernel, dim3(grid), dim3(256), 0, st, e, nx, ny, nz, ncx, ncy, thr,
                     out);
  return hipGetLastError();
}

// Fill kernel for clearing small device buffers inside a stream without
// hipMemsetAsync.
__global__ void fill_u64_kernel(unsigned long long *p, int n, unsigned long long v) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
}

hipError_t launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t st) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3((n + 63) / 64), dim3(64), 0, st, p, n, v);
  return hipGetLastError();
}

}  // namespace gvh
