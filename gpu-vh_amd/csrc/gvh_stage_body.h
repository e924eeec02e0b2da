// gvh_stage_body.h -- one fused kernel per Runge-Kutta stage of the KT step
// (launched by gvh_stage.hip for one event and by gvh_batch_kernels.hip for
// batches of events).
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
#pragma once
#include "../../include/gpuvh.h"
#include "gvh_device.h"

// Waves per SIMD asked of the one-event stage kernel (the second launch
// bound is amdgpu_waves_per_eu; 1 leaves the register allocation free).
// Measured on the 201x201x101 table-EoS event: the FP32 (mixed) form at three
// waves (166-168 VGPRs) runs 13% faster than at its free two; the FP64 form
// runs 5% slower when held to two than when free (same register count,
// different schedule), and at three its spills cost more than the wave gains.
#ifndef GVH_WAVES_FP32
#define GVH_WAVES_FP32 3
#endif
#ifndef GVH_WAVES_FP64
#define GVH_WAVES_FP64 1
#endif
// The mixed form with FP64 half-site root-finds (arrangement LOCK 3): its
// FP64 lockstep chains spill 212 B per lane at three waves.
#ifndef GVH_WAVES_FP32_ROOT64
#define GVH_WAVES_FP32_ROOT64 2
#endif
template <typename R, int LOCK = 0>
constexpr int stage_waves_per_simd() {
  return sizeof(R) == 4 ? (LOCK == 3 ? GVH_WAVES_FP32_ROOT64 : GVH_WAVES_FP32) : GVH_WAVES_FP64;
}

namespace gvh {

// ------------------------------------------------------ source terms
template <typename R>
struct GradT {
  R ut[3], ux[3], uy[3], un[3], p[3];  // [x, y, eta] derivatives
};
using Grad = GradT<double>;

// loadSourceTerms2 (SourceTerms.cu:570-755) with setPimunuSourceTerms
// (:30-256): T^{tau mu} geometric and pressure-gradient sources and the
// relaxation right-hand sides of pi^{mu nu} and Pi.
template <int EOS, bool BULK, typename R>
__device__ __forceinline__ void source_terms(const StageArgs &A, const R *Q, double e_eos,
                                             R e, R p, R ut, R ux, R uy, R un, R utp, R uxp,
                                             R uyp, R unp, const GradT<R> &g, R *S) {
  const R t = R(A.t);
  const R inv_t = div_st(R(1.0), t);
  const R t2 = t * t, t3 = t * t2;
  const R inv_t2 = div_st(R(1.0), t2);
  const R Pi = BULK ? Q[14] : R(0.0);
  const R pitt = Q[4], pitx = Q[5], pity = Q[6], pitn = Q[7], pixx = Q[8],
               pixy = Q[9], pixn = Q[10], piyy = Q[11], piyn = Q[12], pinn = Q[13];
  const R inv_ut = div_st(R(1.0), ut);

  // T^{tau mu} sources, SourceTerms.cu:726-737
  const R PP = p + Pi;
  const R tnn = (e + PP) * un * un + PP * inv_t2 + pinn;
  const R vx = ux * inv_ut, vy = uy * inv_ut, vn = un * inv_ut;
  const R dkvk = (g.ux[0] - vx * g.ut[0]) * inv_ut + (g.uy[1] - vy * g.ut[1]) * inv_ut +
                      (g.un[2] - vn * g.ut[2]) * inv_ut;
  S[0] = -(Q[0] * inv_t + t * tnn) + dkvk * (pitt - p - Pi) - vx * g.p[0] - vy * g.p[1] -
         vn * g.p[2];
  S[1] = -Q[1] * inv_t - g.p[0] + dkvk * pitx;
  S[2] = -Q[2] * inv_t - g.p[1] + dkvk * pity;
  S[3] = -R(3.0) * Q[3] * inv_t - g.p[2] * inv_t2 + dkvk * pitn;

  // transport coefficients, SourceTerms.cu:46-60
  const double T_eos = eos_T<EOS>(A.tab, e_eos);
  const R T = R(T_eos);
  const R taupiInv = R(0.2f) * div_st(T, R(A.etabar));
  const R beta_pi = R(0.2f) * (e + p);

  // backward-difference time derivative and covariant acceleration, :67-86
  const R dtut = (ut - utp) * R(A.inv_dt), dtux = (ux - uxp) * R(A.inv_dt);
  const R dtuy = (uy - uyp) * R(A.inv_dt), dtun = (un - unp) * R(A.inv_dt);
  const R un2 = un * un, ut2 = ut * ut;
  const R dut = ut * dtut + ux * g.ut[0] + uy * g.ut[1] + un * g.ut[2];
  const R Dut = dut + t * un2;
  const R dux = ut * dtux + ux * g.ux[0] + uy * g.ux[1] + un * g.ux[2];
  const R duy = ut * dtuy + ux * g.uy[0] + uy * g.uy[1] + un * g.uy[2];
  const R dun = ut * dtun + ux * g.un[0] + uy * g.un[1] + un * g.un[2];
  const R Dux = -dux, Duy = -duy;
  const R Dun = -t2 * dun - R(2.0) * t * ut * un;

  const R theta = ut * inv_t + dtut + g.ux[0] + g.uy[1] + g.un[2];
  const R th3 = theta * (R(1.0) / R(3.0));

  // sigma^{mu nu}, :96-106
  const R stt = -t * ut * un2 + (dtut - ut * dut) + (ut2 - R(1.0)) * th3;
  const R stx = -(t * un2 * ux) * R(0.5) + (dtux - g.ut[0]) * R(0.5) -
                     (ux * dut + ut * dux) * R(0.5) + ut * ux * th3;
  const R sty = -(t * un2 * uy) * R(0.5) + (dtuy - g.ut[1]) * R(0.5) -
                     (uy * dut + ut * duy) * R(0.5) + ut * uy * th3;
  const R stn = -un * (R(2.0) * ut2 + t2 * un2) * (R(0.5) * inv_t) +
                     (dtun - g.ut[2] * inv_t2) * R(0.5) - (un * dut + ut * dun) * R(0.5) +
                     ut * un * th3;
  const R sxx = -(g.ux[0] + ux * dux) + (R(1.0) + ux * ux) * th3;
  const R sxy = -(g.uy[0] + g.ux[1]) * R(0.5) - (uy * dux + ux * duy) * R(0.5) + ux * uy * th3;
  const R sxn = -ut * ux * un * inv_t - (g.un[0] + g.ux[2] * inv_t2) * R(0.5) -
                     (un * dux + ux * dun) * R(0.5) + ux * un * th3;
  const R syy = -(g.uy[1] + uy * duy) + (R(1.0) + uy * uy) * th3;
  const R syn = -ut * uy * un * inv_t - (g.un[1] + g.uy[2] * inv_t2) * R(0.5) -
                     (un * duy + uy * dun) * R(0.5) + uy * un * th3;
  const R snn = div_st(-ut * (R(1.0) + R(2.0) * t2 * un2), t3) - g.un[2] * inv_t2 - un * dun +
                     (inv_t2 + un2) * th3;

  // omega^mu_nu, :123-136
  const R wtx = (-dtux - g.ut[0] + ut * dux - ux * dut - t * un2 * ux) * R(0.5);
  const R wty = (-dtuy - g.ut[1] + ut * duy - uy * dut - t * un2 * uy) * R(0.5);
  const R wtn = (-t2 * dtun - R(2.0) * t * un - g.ut[2] - ut * Dun - t2 * un * dut -
                      t3 * un * un2) * R(0.5);
  const R wxy = (g.uy[0] - g.ux[1] + ux * duy - uy * dux) * R(0.5);
  const R wxn = (t2 * g.un[0] - g.ux[2] - ux * Dun - t2 * un * dux) * R(0.5);
  const R wyn = (t2 * g.un[1] - g.uy[2] - uy * Dun - t2 * un * duy) * R(0.5);
  const R wxt = wtx, wyt = wty, wnt = wtn * inv_t2, wyx = -wxy;
  const R wnx = -wxn * inv_t2, wny = -wyn * inv_t2;

  // pi^{mu lambda} D u_lambda for each row mu, then I1 = u^mu a^nu + u^nu a^mu
  const R at = pitt * Dut + pitx * Dux + pity * Duy + pitn * Dun;
  const R ax = pitx * Dut + pixx * Dux + pixy * Duy + pixn * Dun;
  const R ay = pity * Dut + pixy * Dux + piyy * Duy + piyn * Dun;
  const R an = pitn * Dut + pixn * Dux + piyn * Duy + pinn * Dun;

  // pi:sigma, :185-187
  const R ps = pitt * stt - R(2.0) * pitx * stx - R(2.0) * pity * sty + pixx * sxx +
                    R(2.0) * pixy * sxy + piyy * syy - R(2.0) * pitn * stn * t2 +
                    R(2.0) * pixn * sxn * t2 + R(2.0) * piyn * syn * t2 + pinn * snn * t2 * t2;
  const R ps3 = ps * (R(1.0) / R(3.0));

  constexpr R d_pp = R(1.33333), t_pp = R(1.42857), l_pP = R(1.2);  // TransportCoefficients.cuh:14-16
  const R lPi = l_pP * Pi;

  // I = I1 + delta I2 - I3 + tau I4 - lambda Pi sigma, relaxation RHS with
  // Milne Christoffel terms, :141-233; RHS/u^tau + pi dkvk, :243-252
  const R b2 = R(2.0) * beta_pi;
  {
    const R I1 = R(2.0) * ut * at;
    const R I3 = R(2.0) * (pitx * wtx + pity * wty + pitn * wtn);
    const R I4 = (pitt * stt - pitx * stx - pity * sty - t2 * pitn * stn) - (R(1.0) - ut2) * ps3;
    const R I = I1 + d_pp * theta * pitt - I3 + t_pp * I4 - lPi * stt;
    const R r = b2 * stt - pitt * taupiInv - I - R(2.0) * un * t * pitn;
    S[4] = r * inv_ut + pitt * dkvk;
  }
  {
    const R I1 = ux * at + ut * ax;
    const R I3 = pitt * wxt + pity * wxy + pitn * wxn + pixx * wtx + pixy * wty + pixn * wtn;
    const R I4 = (pitt * stx + pitx * stt) * R(0.5) - (pitx * sxx + pixx * stx) * R(0.5) -
                      (pity * sxy + pixy * sty) * R(0.5) - t2 * (pitn * sxn + pixn * stn) * R(0.5) +
                      (ut * ux) * ps3;
    const R I = I1 + d_pp * theta * pitx - I3 + t_pp * I4 - lPi * stx;
    const R r = b2 * stx - pitx * taupiInv - I - un * t * pixn;
    S[5] = r * inv_ut + pitx * dkvk;
  }
  {
    const R I1 = uy * at + ut * ay;
    const R I3 = pitt * wyt + pitx * wyx + pitn * wyn + pixy * wtx + piyy * wty + piyn * wtn;
    const R I4 = (pitt * sty + pity * stt) * R(0.5) - (pitx * sxy + pixy * stx) * R(0.5) -
                      (pity * syy + piyy * sty) * R(0.5) - t2 * (pitn * syn + piyn * stn) * R(0.5) +
                      (ut * uy) * ps3;
    const R I = I1 + d_pp * theta * pity - I3 + t_pp * I4 - lPi * sty;
    const R r = b2 * sty - pity * taupiInv - I - un * t * piyn;
    S[6] = r * inv_ut + pity * dkvk;
  }
  {
    const R I1 = un * at + ut * an;
    const R I3 = pitt * wnt + pitx * wnx + pity * wny + pixn * wtx + piyn * wty + pinn * wtn;
    const R I4 = (pitt * stn + pitn * stt) * R(0.5) - (pitx * sxn + pixn * stx) * R(0.5) -
                      (pity * syn + piyn * sty) * R(0.5) - t2 * (pitn * snn + pinn * stn) * R(0.5) +
                      (ut * un) * ps3;
    const R I = I1 + d_pp * theta * pitn - I3 + t_pp * I4 - lPi * stn;
    const R r = b2 * stn - pitn * taupiInv - I - un * t * pinn -
                     (ut * pitn + un * pitt) * inv_t;
    S[7] = r * inv_ut + pitn * dkvk;
  }
  {
    const R I1 = R(2.0) * ux * ax;
    const R I3 = R(2.0) * (pitx * wxt + pixy * wxy + pixn * wxn);
    const R I4 = (pitx * stx - pixx * sxx - pixy * sxy - t2 * pixn * sxn) + (R(1.0) + ux * ux) * ps3;
    const R I = I1 + d_pp * theta * pixx - I3 + t_pp * I4 - lPi * sxx;
    const R r = b2 * sxx - pixx * taupiInv - I;
    S[8] = r * inv_ut + pixx * dkvk;
  }
  {
    const R I1 = uy * ax + ux * ay;
    const R I3 = pitx * wyt + pity * wxt + pixx * wyx + piyy * wxy + pixn * wyn + piyn * wxn;
    const R I4 = (pitx * sty + pity * stx) * R(0.5) - (pixx * sxy + pixy * sxx) * R(0.5) -
                      (pixy * syy + piyy * sxy) * R(0.5) - t2 * (pixn * syn + piyn * sxn) * R(0.5) +
                      (ux * uy) * ps3;
    const R I = I1 + d_pp * theta * pixy - I3 + t_pp * I4 - lPi * sxy;
    const R r = b2 * sxy - pixy * taupiInv - I;
    S[9] = r * inv_ut + pixy * dkvk;
  }
  {
    const R I1 = un * ax + ux * an;
    const R I3 = pitx * wnt + pitn * wxt + pixx * wnx + pixy * wny + piyn * wxy + pinn * wxn;
    const R I4 = (pitx * stn + pitn * stx) * R(0.5) - (pixx * sxn + pixn * sxx) * R(0.5) -
                      (pixy * syn + piyn * sxy) * R(0.5) - t2 * (pixn * snn + pinn * sxn) * R(0.5) +
                      (ux * un) * ps3;
    const R I = I1 + d_pp * theta * pixn - I3 + t_pp * I4 - lPi * sxn;
    const R r = b2 * sxn - pixn * taupiInv - I - (ut * pixn + un * pitx) * inv_t;
    S[10] = r * inv_ut + pixn * dkvk;
  }
  {
    const R I1 = R(2.0) * uy * ay;
    const R I3 = R(2.0) * (pity * wyt + pixy * wyx + piyn * wyn);
    const R I4 = (pity * sty - pixy * sxy - piyy * syy - t2 * piyn * syn) + (R(1.0) + uy * uy) * ps3;
    const R I = I1 + d_pp * theta * piyy - I3 + t_pp * I4 - lPi * syy;
    const R r = b2 * syy - piyy * taupiInv - I;
    S[11] = r * inv_ut + piyy * dkvk;
  }
  {
    const R I1 = un * ay + uy * an;
    const R I3 = pity * wnt + pitn * wyt + pixy * wnx + pixn * wyx + piyy * wny + pinn * wyn;
    const R I4 = (pity * stn + pitn * sty) * R(0.5) - (pixy * sxn + pixn * sxy) * R(0.5) -
                      (piyy * syn + piyn * syy) * R(0.5) - t2 * (piyn * snn + pinn * syn) * R(0.5) +
                      (uy * un) * ps3;
    const R I = I1 + d_pp * theta * piyn - I3 + t_pp * I4 - lPi * syn;
    const R r = b2 * syn - piyn * taupiInv - I - (ut * piyn + un * pity) * inv_t;
    S[12] = r * inv_ut + piyn * dkvk;
  }
  {
    const R I1 = R(2.0) * un * an;
    const R I3 = R(2.0) * (pitn * wnt + pixn * wnx + piyn * wny);
    const R I4 = (pitn * stn - pixn * sxn - piyn * syn - t2 * pinn * snn) + (inv_t2 + un2) * ps3;
    const R I = I1 + d_pp * theta * pinn - I3 + t_pp * I4 - lPi * snn;
    const R r = b2 * snn - pinn * taupiInv - I - R(2.0) * (ut * pinn + un * pitn) * inv_t;
    S[13] = r * inv_ut + pinn * dkvk;
  }
  if constexpr (BULK) {
    // bulk relaxation, :53-60 and :238, :254; c_s^2 from the FP64 EoS (1/3 -
    // c_s^2 is small where the fluid is nearly conformal), zeta/s in the
    // arithmetic of the source terms
    using FE = R;
    double p_, cs2;
    eos_p_cs2<EOS>(A.tab, e_eos, p_, cs2);
    const R a = R(0.333333f) - R(cs2);
    const R a2 = a * a;
    const R beta_Pi = R(15.0) * a2 * (e + p);
    const R lambda_Pipi = R(1.6f) * a;
    const R tauPiInv = R(15.0) * a2 * div_st(T, R(zeta_over_s(FE(T_eos), FE(A.zeta_tpeak))));
    constexpr R d_PP = R(0.666667);  // TransportCoefficients.cuh:18
    const R dPi = -beta_Pi * theta - Pi * tauPiInv - d_PP * Pi * theta + lambda_Pipi * ps;
    S[14] = dPi * inv_ut + Pi * dkvk;
  }
}

// ------------------------------------------------- directional update
// Threads per block of the stage kernel; the per-thread accumulator of the 15
// conserved variables lives in LDS as acc[n * BLOCK + threadIdx.x] so that the
// variable loops can stay rolled (dynamic index) and registers stay few.
constexpr int BLOCK = GVH_BLOCK;

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
// YSTAGE (y direction of the four-row blocks, kt_flux_kernel_ystaged): the
// 45 stencil values of the nine inversion variables come from the block's
// LDS copy of rows j-2 .. j+2 (ystage[variable][row slot][lane]) instead of
// from memory.
template <int EOS, bool BULK, int DIR, bool KEEP_STATES, int LOCK, typename R, typename PS,
          bool YSTAGE = false>
__device__ __forceinline__ void direction_update(const StageArgs &A, unsigned s, double ePrev,
                                                 R v_s, double *acc,
                                                 const R *ystage = nullptr) {
  constexpr int NV = BULK ? 15 : 14;
  const long long st = DIR == 0 ? 1 : (DIR == 1 ? (long long)A.ncx : (long long)A.ncx * A.ncy);
  const long long G = A.G;
  const double *__restrict__ qT = A.qin;                                    // T^{tau mu}
  const PS *__restrict__ qP = reinterpret_cast<const PS *>(A.qin + 4 * G);  // pi, Pi
  const double t = A.t;
  const R hc = R(A.half_inv_h[DIR]);
  // conserved variable n at cell s + o (n < 4: T^{tau mu}, else pi / Pi)
  auto ld = [&](int n, long long o) -> R {
    return n < 4 ? R(ld_cell(qT + n * G + o, s)) : R(ld_cell(qP + (n - 4) * G + o, s));
  };
  auto ldp = [&](int n, long long o) -> R { return R(ld_cell(qP + (n - 4) * G + o, s)); };

  // the six spatial pi components, whose flux comes after that of the kept
  // states: their 30 stencil values in one batch.  In FP32 they are loaded
  // before the root-finds and stay in flight during them; in FP64 the 60
  // registers this takes would cost a wave per SIMD, so they are loaded
  // after the root-finds.
  constexpr bool FLUX_FROM_KEPT = KEEP_STATES && LOCK != 2;
  constexpr bool EARLY_RAW2 = FLUX_FROM_KEPT && sizeof(R) == 4;
  R raw2[6][5];
  auto load_raw2 = [&]() {
#pragma unroll
    for (int n = 8; n < 14; ++n)
#pragma unroll
      for (int m = 0; m < 5; ++m) raw2[n - 8][m] = ldp(n, (m - 2) * st);
  };
  double vRf, vLf, vRb, vLb;
  // the four states when KEEP_STATES: [0] right of i+1/2, [1] left of i+1/2,
  // [2] right of i-1/2, [3] left of i-1/2
  R Wst[4][9];
  R cdv[5];
  if constexpr (KEEP_STATES) {
    // all four states of the nine inversion variables stay in registers; the
    // 45 stencil values are loaded in one batch (one memory round trip for
    // the direction instead of one per variable)
    R raw[9][5];
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      if (a == 8 && !BULK) continue;
      const int n = a < 8 ? a : 14;
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        if constexpr (YSTAGE && DIR == 1)
          raw[a][m] = ystage[(a * 8 + (threadIdx.x >> 6) + m) * 64 + (threadIdx.x & 63)];
        else
          raw[a][m] = ld(n, (m - 2) * st);
      }
    }
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      if (a == 8 && !BULK) {
        Wst[0][8] = Wst[1][8] = Wst[2][8] = Wst[3][8] = cdv[4] = R(0);
        continue;
      }
      const R vm2 = raw[a][0], vm1 = raw[a][1], v0 = raw[a][2], vp1 = raw[a][3],
              vp2 = raw[a][4];
      const R h0 = half_slope_t<R>(vm1, v0, vp1);
      Wst[0][a] = vp1 - half_slope_t<R>(v0, vp1, vp2);
      Wst[1][a] = v0 + h0;
      Wst[2][a] = v0 - h0;
      Wst[3][a] = vm1 + half_slope_t<R>(vm2, vm1, v0);
      if (a >= 4) cdv[a - 4] = (vp1 - vm1) * hc;
    }
    if constexpr (EARLY_RAW2) load_raw2();
    double v4[4];
    if constexpr (LOCK) {
      flux_velocity_lockstep<EOS, DIR, 4, R, LOCK == 3>(A.tab, t, Wst, ePrev, A.diag,
                                                        A.max_iters, v4);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        v4[k] = flux_velocity<EOS, DIR>(A.tab, t, (double)Wst[k][0], (double)Wst[k][1],
                                        (double)Wst[k][2], (double)Wst[k][3], (double)Wst[k][4],
                                        (double)Wst[k][5], (double)Wst[k][6], (double)Wst[k][7],
                                        (double)Wst[k][8], ePrev, A.diag, A.max_iters);
    }
    vRf = v4[0];
    vLf = v4[1];
    vRb = v4[2];
    vLb = v4[3];
  } else {
    // the four states one after another (rolled: few live registers)
    vRf = vLf = vRb = vLb = 0.0;
#pragma unroll 1
    for (int state = 0; state < 4; ++state) {
      const long long o = state == 0 ? st : (state == 3 ? -st : 0);
      const R sg = (state & 1) ? R(1) : R(-1);
      R w[9];
#pragma unroll
      for (int a = 0; a < 9; ++a) {
        if (a == 8 && !BULK) {
          w[8] = R(0);
          continue;
        }
        const int n = a < 8 ? a : 14;
        const R xm = ld(n, o - st), x0 = ld(n, o), xp = ld(n, o + st);
        w[a] = x0 + half_slope_t<R>(xm, x0, xp) * sg;
      }
      const double v = flux_velocity<EOS, DIR>(A.tab, t, (double)w[0], (double)w[1], (double)w[2],
                                               (double)w[3], (double)w[4], (double)w[5],
                                               (double)w[6], (double)w[7], (double)w[8], ePrev,
                                               A.diag, A.max_iters);
      vRf = state == 0 ? v : vRf;
      vLf = state == 1 ? v : vLf;
      vRb = state == 2 ? v : vRb;
      vLb = state == 3 ? v : vLb;
    }
#pragma unroll
    for (int a = 4; a < 9; ++a) {
      if (a == 8 && !BULK) {
        cdv[4] = R(0);
        continue;
      }
      const int n = a < 8 ? a : 14;
      cdv[a - 4] = (ld(n, st) - ld(n, -st)) * hc;
    }
  }
  // local propagation speeds, max(|v_L|, |v_R|) (LocalPropagationSpeed.cu:16-24)
  const R rRf = R(vRf), rLf = R(vLf), rRb = R(vRb), rLb = R(vLb);
  const R aF = fmax(fabs(rLf), fabs(rRf));
  const R aB = fmax(fabs(rLb), fabs(rRb));


  if constexpr (FLUX_FROM_KEPT && !EARLY_RAW2) load_raw2();

  // directional sources: central differences of pi^{tau mu} and Pi and of
  // the three spatial pi components of this direction
  auto cd = [&](int n) {
    if constexpr (FLUX_FROM_KEPT) return (raw2[n - 8][3] - raw2[n - 8][1]) * hc;
    return (ldp(n, st) - ldp(n, -st)) * hc;
  };
  const R cPi = cdv[4];
  R Sd0, Sd1, Sd2, Sd3;
  if constexpr (DIR == 0) {
    Sd0 = cdv[0] * v_s - cdv[1] - (BULK ? v_s * cPi : R(0));
    Sd1 = cdv[1] * v_s - cd(8) - cPi;
    Sd2 = cdv[2] * v_s - cd(9);
    Sd3 = cdv[3] * v_s - cd(10);
  } else if constexpr (DIR == 1) {
    Sd0 = cdv[0] * v_s - cdv[2] - (BULK ? v_s * cPi : R(0));
    Sd2 = cdv[2] * v_s - cd(11) - cPi;
    Sd1 = cdv[1] * v_s - cd(9);
    Sd3 = cdv[3] * v_s - cd(12);
  } else {
    Sd0 = cdv[0] * v_s - cdv[3] - (BULK ? v_s * cPi : R(0));
    Sd3 = cdv[3] * v_s - cd(13) - (BULK ? div_st(cPi, R(t * t)) : R(0));
    Sd1 = cdv[1] * v_s - cd(10);
    Sd2 = cdv[2] * v_s - cd(12);
  }

  const R hih = R(A.half_inv_h[DIR]);
  const double dt = A.dt;
  // (H_b - H_f)/h with H = (F(q_R) + F(q_L) - a (q_R - q_L))/2, F = v q,
  // written per interface as 2H = (v_R - a) q_R + (v_L + a) q_L: the two
  // weights are formed once per direction, so each variable's flux is one
  // multiply and one FMA (equal to the reference's expression up to
  // rounding).  The halving is applied once to the difference,
  // (2H_b - 2H_f)/(2h).
  const R wRb = rRb - aB, wLb = rLb + aB, wRf = rRf - aF, wLf = rLf + aF;
#define GVH_FLUXDIFF(rf, lf, rb, lb)                                         \
  (((wRb * (rb) + wLb * (lb)) - (wRf * (rf) + wLf * (lf))) * hih)
  // the kept states give the flux of the nine inversion variables, unless
  // LOCK == 2, which rebuilds every state from the stencil after the
  // root-finds so that none stays live through them
  int n0 = 0;
  if constexpr (FLUX_FROM_KEPT) {
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      if (a == 8 && !BULK) continue;
      const int n = a < 8 ? a : 14;
      R r = GVH_FLUXDIFF(Wst[0][a], Wst[1][a], Wst[2][a], Wst[3][a]);
      if (a == 0) r += Sd0;
      if (a == 1) r += Sd1;
      if (a == 2) r += Sd2;
      if (a == 3) r += Sd3;
      acc[n * BLOCK] += (double)r * dt;
    }
    n0 = 8;
  }
  auto flux_of = [&](R vm2, R vm1, R v0, R vp1, R vp2) {
    const R h0 = half_slope_t<R>(vm1, v0, vp1);
    const R rf = vp1 - half_slope_t<R>(v0, vp1, vp2);
    const R lf = v0 + h0;
    const R rb = v0 - h0;
    const R lb = vm1 + half_slope_t<R>(vm2, vm1, v0);
    return GVH_FLUXDIFF(rf, lf, rb, lb);
  };
  if constexpr (FLUX_FROM_KEPT) {
#pragma unroll
    for (int n = 8; n < 14; ++n) {
      const R *w = raw2[n - 8];
      acc[n * BLOCK] += (double)flux_of(w[0], w[1], w[2], w[3], w[4]) * dt;
    }
  } else {
#pragma unroll 1
    for (int n = n0; n < NV; ++n) {
      R r = flux_of(ld(n, -2 * st), ld(n, -st), ld(n, 0), ld(n, st), ld(n, 2 * st));
      if (n < 4) r += n == 0 ? Sd0 : (n == 1 ? Sd1 : (n == 2 ? Sd2 : Sd3));
      acc[n * BLOCK] += (double)r * dt;
    }
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

// Cell of this thread in tile b of nb tiles (a multiple of 8): consecutive
// tiles go to the 8 XCDs in turn, so tile b is renumbered to
// lb = (b mod 8) nb/8 + b/8 and each XCD sweeps a contiguous range.
// The thread index and its quotients are 32-bit: a lattice's ghosted cell
// count is below 2^29 (every array of 8-byte values is below 4 GiB), and the
// grid exceeds the cell count by less than one strip plus 8 blocks.
__device__ __forceinline__ CellPos cell_of_tile(const StageArgs &A, unsigned b, unsigned nb) {
  const unsigned lb = (b & 7u) * (nb >> 3) + (b >> 3);
  const unsigned by = (unsigned)A.strip, nx = (unsigned)A.nx, nz = (unsigned)A.nz;
  const unsigned tid = lb * BLOCK + threadIdx.x;
  const unsigned r = tid / nx;
  const int i = (int)(tid - r * nx);
  const unsigned r2 = r / by;
  const int jj = (int)(r - r2 * by);
  const unsigned yb_ = r2 / nz;
  const int k = (int)(r2 - yb_ * nz);
  const int yb = (int)yb_;
  const int j = yb * (int)by + jj;
  CellPos c;
  c.valid = j < A.ny && yb < A.nstrips;
  c.i = i + NG;
  c.j = (c.valid ? j : 0) + NG;
  c.k = k + NG;
  c.s = (unsigned)c.i + (unsigned)A.ncx * ((unsigned)c.j + (unsigned)A.ncy * (unsigned)c.k);
  return c;
}

__device__ __forceinline__ CellPos cell_of_thread(const StageArgs &A) {
  return cell_of_tile(A, blockIdx.x, gridDim.x);
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
  source_terms<EOS, BULK, double>(A, Q, e_s, e_s, p_s, ut, ux, uy, un, up[s], (up + G)[s],
                                  (up + 2 * G)[s], (up + 3 * G)[s], g, S);
#pragma unroll
  for (int n = 0; n < NV; ++n) (A.qout + n * G)[s] = Q[n] + A.dt * S[n];
}

// Flux part of the stage plus the cell's finishing work: the three
// directional updates in the reference's order, the convex average (stage
// 2), the root-find, regulation, and the writes of the cell and its ghost
// copies.  acc starts from what kt_source_kernel stored at the cell.
// FUSED: the source terms are computed here first (one kernel per stage);
// otherwise kt_source_kernel has stored acc = q + dt S at the cell.
// YSTAGE: the thread belongs to a four-row block that staged the y stencil
// in LDS; every thread of the block reaches the barrier before the y
// direction, and only threads of valid cells compute (the others only
// staged rows for their neighbours).
template <int EOS, bool BULK, bool STAGE2, bool HAS_Z, bool FUSED, bool KEEP, int LOCK,
          typename R, typename PS, bool YSTAGE = false>
__device__ __forceinline__ void stage_cell(const StageArgs &A, const CellPos &c, double *acc,
                                           const R *ystage = nullptr) {
  constexpr int NV = BULK ? 15 : 14;
  const int i = c.i, j = c.j, k = c.k;
  const unsigned s = c.s;
  const long long sy = A.ncx, sz = (long long)A.ncx * A.ncy;
  const long long G = A.G;
  // conserved variable n of a buffer at this cell: T^{tau mu} in FP64, pi
  // and Pi in the storage type PS
  auto q_at = [&](const double *base, int n) -> double {
    return n < 4 ? ld_cell(base + n * G, s)
                 : (double)ld_cell(reinterpret_cast<const PS *>(base + 4 * G) + (n - 4) * G, s);
  };
  const double *u = A.uin;
  double ut = 1.0, ux = 0.0, uy = 0.0, un = 0.0, e_s = 0.0;
  if (!YSTAGE || c.valid) {
    ut = ld_cell(u, s);
    ux = ld_cell(u + G, s);
    uy = ld_cell(u + 2 * G, s);
    un = ld_cell(u + 3 * G, s);
    e_s = ld_cell(A.ein, s);
    if constexpr (FUSED) {
      const double p_s = ld_cell(A.pin, s);
      GradT<R> g;
      const long long st[3] = {1, sy, sz};
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (d == 2 && !HAS_Z) {
          g.ut[2] = g.ux[2] = g.uy[2] = g.un[2] = g.p[2] = R(0);
          continue;
        }
        const R ih = R(A.inv_h[d]);
        const double *um = u - st[d], *upl = u + st[d];
        g.ut[d] = slope_t<R>(R(ld_cell(um, s)), R(ut), R(ld_cell(upl, s))) * ih;
        g.ux[d] = slope_t<R>(R(ld_cell(um + G, s)), R(ux), R(ld_cell(upl + G, s))) * ih;
        g.uy[d] = slope_t<R>(R(ld_cell(um + 2 * G, s)), R(uy), R(ld_cell(upl + 2 * G, s))) * ih;
        g.un[d] = slope_t<R>(R(ld_cell(um + 3 * G, s)), R(un), R(ld_cell(upl + 3 * G, s))) * ih;
        g.p[d] = slope_t<R>(R(ld_cell(A.pin - st[d], s)), R(p_s), R(ld_cell(A.pin + st[d], s))) * ih;
      }
      double Qd[15];
      R Q[15], S[15];
#pragma unroll
      for (int n = 0; n < 15; ++n) {
        Qd[n] = (n < NV) ? q_at(A.qin, n) : 0.0;
        Q[n] = R(Qd[n]);
      }
      S[14] = R(0);
      const double *up = A.uprev;
      source_terms<EOS, BULK, R>(A, Q, e_s, R(e_s), R(p_s), R(ut), R(ux), R(uy), R(un),
                                 R(ld_cell(up, s)), R(ld_cell(up + G, s)),
                                 R(ld_cell(up + 2 * G, s)), R(ld_cell(up + 3 * G, s)), g, S);
#pragma unroll
      for (int n = 0; n < 15; ++n) acc[n * BLOCK] = Qd[n] + A.dt * (double)S[n];
    } else {
#pragma unroll
      for (int n = 0; n < NV; ++n) acc[n * BLOCK] = ld_cell(A.qout + n * G, s);
    }

    // flux velocity of the cell itself for the directional sources (v = u/u^tau)
    direction_update<EOS, BULK, 0, KEEP, LOCK, R, PS>(A, s, e_s, div_st(R(ux), R(ut)), acc);
  }
  if constexpr (YSTAGE) {
    __syncthreads();
    if (!c.valid) return;
  }
  direction_update<EOS, BULK, 1, KEEP, LOCK, R, PS, YSTAGE>(A, s, e_s, div_st(R(uy), R(ut)), acc,
                                                            ystage);
  if constexpr (HAS_Z) direction_update<EOS, BULK, 2, KEEP, LOCK, R, PS>(A, s, e_s, div_st(R(un), R(ut)), acc);

  if constexpr (STAGE2) {
    // q^n of the convex average: all NV loads in flight together
    double qn[NV];
#pragma unroll
    for (int n = 0; n < NV; ++n) qn[n] = q_at(A.qn, n);
#pragma unroll
    for (int n = 0; n < NV; ++n) acc[n * BLOCK] = (acc[n * BLOCK] + qn[n]) * 0.5;
  }

  // inferred variables with this cell's e as the guess
  const Prim pr = inferred<EOS>(A.tab, A.tinf, acc[0], acc[BLOCK], acc[2 * BLOCK], acc[3 * BLOCK],
                                acc[4 * BLOCK], acc[5 * BLOCK], acc[6 * BLOCK], acc[7 * BLOCK],
                                BULK ? acc[14 * BLOCK] : 0.0, e_s, A.diag, A.max_iters);

  // regulateDissipativeCurrents (RegulateDissipativeCurrents.cu:16-134): only
  // cells at or below e_reg are rescaled; the bulk factor is 1 unless p is 0.
  // Stage 2 records the shear factor when asked (:132, VALIDITY_DOMAIN).
  double fs_rec = 1.0;
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
    // max_i(x_i/den) as (max_i x_i)/den: a correctly rounded division by a
    // positive den is monotone in x, so both give the same value; with
    // den = 0 both leave rho = a1 unless some x_i > 0 (then +inf)
    const double rho = fmax(a1, fmax(pimumu, fmax(piu0, fmax(piu1, fmax(piu2, piu3)))) / den);
    double fs = tanh(rho) / rho;
    if (fabs(rho) < 1.e-7) fs = 1.0;
    fs_rec = fs;
#pragma unroll 1
    for (int n = 4; n < 14; ++n) acc[n * BLOCK] *= fs;
    if constexpr (BULK) {
      const double rb = 0.0 / pr.p;
      double fb = tanh(rb) / rb;
      if (fabs(rb) < 1.e-7) fb = 1.0;
      acc[14 * BLOCK] *= fb;
    }
  }

  if constexpr (STAGE2) {
    if (A.reg) st_cell(A.reg, s, fs_rec);
  }

  // write the cell and, on the boundary, its two ghost layers
  // (GhostCells.cu:15-144: zero-gradient copy, edges and corners untouched);
  // a cell of a one-point axis touches both faces of that axis
  const bool gx0 = i == NG, gx1 = i == A.nx + 1, gy0 = j == NG, gy1 = j == A.ny + 1;
  const bool gz0 = k == NG, gz1 = k == A.nz + 1;
  const long long sy2 = 2 * sy, sz2 = 2 * sz;
  auto put = [&](auto *base, auto val) {
    st_cell(base, s, val);
    if (gx0) { st_cell(base - 1, s, val); st_cell(base - 2, s, val); }
    if (gx1) { st_cell(base + 1, s, val); st_cell(base + 2, s, val); }
    if (gy0) { st_cell(base - sy, s, val); st_cell(base - sy2, s, val); }
    if (gy1) { st_cell(base + sy, s, val); st_cell(base + sy2, s, val); }
    if (gz0) { st_cell(base - sz, s, val); st_cell(base - sz2, s, val); }
    if (gz1) { st_cell(base + sz, s, val); st_cell(base + sz2, s, val); }
  };
  PS *qoutP = reinterpret_cast<PS *>(A.qout + 4 * G);
#pragma unroll 1
  for (int n = 0; n < 4; ++n) put(A.qout + n * G, acc[n * BLOCK]);
#pragma unroll 1
  for (int n = 4; n < NV; ++n) put(qoutP + (n - 4) * G, (PS)acc[n * BLOCK]);
  put(A.eout, pr.e);
  put(A.pout, pr.p);
  put(A.uout, pr.ut);
  put(A.uout + G, pr.ux);
  put(A.uout + 2 * G, pr.uy);
  put(A.uout + 3 * G, pr.un);
}

template <int EOS, bool BULK, bool STAGE2, bool HAS_Z, bool FUSED, bool KEEP, int LOCK,
          typename R, typename PS>
__global__ __launch_bounds__(BLOCK, (stage_waves_per_simd<R, LOCK>())) void kt_flux_kernel(StageArgs A) {
  __shared__ double acc_lds[15 * BLOCK];
  const CellPos c = cell_of_thread(A);
  if (!c.valid) return;
  stage_cell<EOS, BULK, STAGE2, HAS_Z, FUSED, KEEP, LOCK, R, PS>(A, c, acc_lds + threadIdx.x);
}

// ------------------------------------- y stencil staged in LDS (arrangement 5)
// Blocks of four waves on four neighbouring rows of a strip: lane L of every
// wave has the same x and row group g, and wave w takes row 4 g + w (the
// strip height is a multiple of 4).  For the nine inversion variables the
// block loads rows j0 - 2 .. j0 + 5 of its lanes once into LDS -- each wave
// its own row and one halo row -- so a cell's y stencil (rows j - 2 .. j + 2)
// is read from LDS instead of five times from memory.  Rows past the
// ghosted lattice are not loaded; only cells past the lattice would read them.
constexpr int YROWS = 4;

struct CellRows {
  CellPos c;
  int jg;  // ghosted row of the thread's wave, not clamped to the lattice
};

__device__ __forceinline__ CellRows cell_of_block_rows(const StageArgs &A, unsigned b,
                                                       unsigned nb) {
  const unsigned lb = (b & 7u) * (nb >> 3) + (b >> 3);
  const unsigned by = (unsigned)A.strip, nx = (unsigned)A.nx, nz = (unsigned)A.nz;
  const unsigned lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
  const unsigned T = lb * 64u + lane;
  const unsigned r = T / nx;
  const int i = (int)(T - r * nx);
  const unsigned gpr = by / YROWS;  // row groups per strip
  const unsigned r2 = r / gpr;
  const int jj = (int)(r - r2 * gpr) * YROWS + (int)w;
  const unsigned yb_ = r2 / nz;
  const int k = (int)(r2 - yb_ * nz);
  const int yb = (int)yb_;
  const int j = yb * (int)by + jj;
  CellRows o;
  o.c.valid = j < A.ny && yb < A.nstrips;
  o.c.i = i + NG;
  o.c.j = (o.c.valid ? j : 0) + NG;
  o.c.k = k + NG;
  o.c.s = (unsigned)o.c.i + (unsigned)A.ncx * ((unsigned)o.c.j + (unsigned)A.ncy * (unsigned)o.c.k);
  o.jg = j + NG;
  return o;
}

template <int EOS, bool BULK, bool STAGE2, bool HAS_Z, int LOCK, typename R, typename PS>
__global__ __launch_bounds__(YROWS * BLOCK, (stage_waves_per_simd<R, LOCK>())) void
kt_flux_kernel_ystaged(StageArgs A) {
  __shared__ double acc_lds[YROWS][15 * BLOCK];
  __shared__ R ystage[9 * 8 * BLOCK];
  const CellRows cr = cell_of_block_rows(A, blockIdx.x, gridDim.x);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  // own row -> slot w + 2; halo: waves 0, 1 rows j - 2 (slots 0, 1), waves
  // 2, 3 rows j + 2 (slots 6, 7)
  const int hrow = w < 2 ? cr.jg - 2 : cr.jg + 2;
  const int hslot = w < 2 ? w : w + 4;
  const long long G = A.G;
  const double *qT = A.qin;
  const PS *qP = reinterpret_cast<const PS *>(A.qin + 4 * G);
  const unsigned ncx = (unsigned)A.ncx, ncy = (unsigned)A.ncy;
  const unsigned colbase = (unsigned)cr.c.i + ncx * ncy * (unsigned)cr.c.k;
#pragma unroll
  for (int pass = 0; pass < 2; ++pass) {
    const int row = pass == 0 ? cr.jg : hrow;
    const int slot = pass == 0 ? w + 2 : hslot;
    const bool in = row >= 0 && row < A.ncy;
    const unsigned sr = in ? colbase + ncx * (unsigned)row : cr.c.s;
#pragma unroll
    for (int a = 0; a < 9; ++a) {
      if (a == 8 && !BULK) continue;
      const int n = a < 8 ? a : 14;
      const R v = n < 4 ? R(ld_cell(qT + n * G, sr)) : R(ld_cell(qP + (n - 4) * G, sr));
      ystage[(a * 8 + slot) * BLOCK + lane] = in ? v : R(0);
    }
  }
  stage_cell<EOS, BULK, STAGE2, HAS_Z, true, true, LOCK, R, PS, true>(A, cr.c, acc_lds[w] + lane,
                                                                     ystage);
}

}  // namespace gvh
