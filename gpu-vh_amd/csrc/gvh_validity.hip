// gvh_validity.hip -- Knudsen and inverse Reynolds numbers per cell
// (checkValidityKernel, HydrodynamicValidity.cu:19-229) and the C ABI of the
// validity arrays.
//
// One thread per interior cell, cells in x-fastest order with the blocks of
// each XCD on a contiguous range, so the +-1 row and +-1 plane neighbours of
// the central differences are mostly in that XCD's L2.  Per cell the kernel
// reads e, p, u, u_prev, 10 pi and Pi (21 values; the 24 u neighbours come
// from cache) and writes 10 values: 248 bytes in FP64, an HBM-bound pass.
#include <cstring>

#include "gvh_internal.h"

namespace gvh {

hipError_t launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t st);

namespace {

constexpr int VBLOCK = 256;

struct ValidityArgs {
  const double *__restrict__ u;   // u^mu, 4 blocks of G
  const double *__restrict__ up;  // u^mu of the previous step
  const double *__restrict__ e;
  const double *__restrict__ p;
  const double *__restrict__ q;   // conserved buffer: T^{tau mu} doubles, then pi, Pi
  double *__restrict__ out;       // GVH_V_REGULATIONS blocks of G
  long long G;
  int nx, ny, ncx, ncy;
  unsigned ncell, nblocks;
  double t, dt, fac[3];           // fac = 1 / h / 2 (HydrodynamicValidity.cu:50-52)
  double etabar, zeta_tpeak;
  EosTable tab;
};

template <int EOS, bool BULK, typename PS>
__global__ __launch_bounds__(VBLOCK) void validity_kernel(ValidityArgs A) {
  // XCD-aware order: block b runs on XCD b % 8; give each XCD a contiguous
  // eighth of the cells (nblocks is a multiple of 8)
  const unsigned b = blockIdx.x;
  const unsigned lb = (b & 7u) * (A.nblocks >> 3) + (b >> 3);
  const unsigned idx = lb * VBLOCK + threadIdx.x;
  if (idx >= A.ncell) return;
  const unsigned i = idx % (unsigned)A.nx, r = idx / (unsigned)A.nx;
  const unsigned j = r % (unsigned)A.ny, k = r / (unsigned)A.ny;
  const long long G = A.G;
  const long long s = (i + 2) + (long long)A.ncx * ((j + 2) + (long long)A.ncy * (k + 2));
  const long long st[3] = {1, A.ncx, (long long)A.ncx * A.ncy};

  double uu[4], upv[4], du[3][4];
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    const double *um = A.u + m * G;
    uu[m] = um[s];
    upv[m] = (A.up + m * G)[s];
#pragma unroll
    for (int d = 0; d < 3; ++d) du[d][m] = (um[s + st[d]] - um[s - st[d]]) * A.fac[d];
  }
  const double ut = uu[0], ux = uu[1], uy = uu[2], un = uu[3];
  const double dxut = du[0][0], dxux = du[0][1], dxuy = du[0][2], dxun = du[0][3];
  const double dyut = du[1][0], dyux = du[1][1], dyuy = du[1][2], dyun = du[1][3];
  const double dnut = du[2][0], dnux = du[2][1], dnuy = du[2][2], dnun = du[2][3];
  const double e_s = A.e[s], p_s = A.p[s];
  const double t = A.t, t2 = t * t, t3 = t * t2, ut2 = ut * ut, un2 = un * un;
  const double etabar = A.etabar;

  // transport coefficients (:76-83); the float literals keep their float value
  double pdummy, cs2;
  eos_p_cs2<EOS>(A.tab, e_s, pdummy, cs2);
  const double T = eos_T<EOS>(A.tab, e_s);
  const double a = (double)0.333333f - cs2;
  const double lambda_Pipi = (double)1.6f * a;
  const double zetabar = zeta_over_s(T, A.zeta_tpeak);
  const double tauPiInv = 15 * (a * a) * (T / zetabar);

  // time derivatives, comoving derivatives and expansion rate (:86-100)
  const double dtut = (ut - upv[0]) / A.dt, dtux = (ux - upv[1]) / A.dt;
  const double dtuy = (uy - upv[2]) / A.dt, dtun = (un - upv[3]) / A.dt;
  const double Dut = ut * dtut + ux * dxut + uy * dyut + un * dnut + t * un * un;
  const double dut = Dut - t * un * un;
  const double dux = ut * dtux + ux * dxux + uy * dyux + un * dnux;
  const double duy = ut * dtuy + ux * dxuy + uy * dyuy + un * dnuy;
  const double dun = ut * dtun + ux * dxun + uy * dyun + un * dnun;
  const double Dun = -t2 * dun - 2 * t * ut * un;
  const double theta = ut / t + dtut + dxux + dyuy + dnun;
  const double th3 = theta / 3;

  // sigma^{mu nu} (:103-113)
  const double stt = -t * ut * un2 + (dtut - ut * dut) + (ut2 - 1) * th3;
  const double stx = -(t * un2 * ux) / 2 + (dtux - dxut) / 2 - (ux * dut + ut * dux) / 2 + ut * ux * th3;
  const double sty = -(t * un2 * uy) / 2 + (dtuy - dyut) / 2 - (uy * dut + ut * duy) / 2 + ut * uy * th3;
  const double stn = -un * (2 * ut2 + t2 * un2) / (2 * t) + (dtun - dnut / t2) / 2 -
                     (un * dut + ut * dun) / 2 + ut * un * th3;
  const double sxx = -(dxux + ux * dux) + (1 + ux * ux) * th3;
  const double sxy = -(dxuy + dyux) / 2 - (uy * dux + ux * duy) / 2 + ux * uy * th3;
  const double sxn = -ut * ux * un / t - (dxun + dnux / t2) / 2 - (un * dux + ux * dun) / 2 + ux * un * th3;
  const double syy = -(dyuy + uy * duy) + (1 + uy * uy) * th3;
  const double syn = -ut * uy * un / t - (dyun + dnuy / t2) / 2 - (un * duy + uy * dun) / 2 + uy * un * th3;
  const double snn = -ut * (1 + 2 * t2 * un2) / t3 - dnun / t2 - un * dun + (1 / t2 + un2) * th3;

  // vorticity in this kernel's own convention (:116-128), which differs
  // from the source terms' (SourceTerms.cu:123-128)
  const double wtx = (dtux + dxut) / 2 + (ux * dut - ut * dux) / 2 + t * un2 * ux / 2;
  const double wty = (dtuy + dyut) / 2 + (uy * dut - ut * duy) / 2 + t * un2 * uy / 2;
  const double wtn = (t2 * dtun + 2 * t * un + dnut) / 2 + (t2 * un * dut - ut * Dun) + t3 * un * un2 / 2;
  const double wxy = (dyux - dxuy) / 2 + (uy * dux - ux * duy) / 2;
  const double wxn = (dnux - t2 * dxun) / 2 + (t2 * un * dux - ux * Dun) / 2;
  const double wyn = (dnuy - t2 * dyun) / 2 + (t2 * un * duy - uy * Dun) / 2;
  const double wnt = wtn / t2, wyx = -wxy, wnx = -wxn / t2, wny = -wyn / t2;

  const PS *pq = reinterpret_cast<const PS *>(A.q + 4 * G);
  const double pitt = pq[s], pitx = pq[G + s], pity = pq[2 * G + s], pitn = pq[3 * G + s];
  const double pixx = pq[4 * G + s], pixy = pq[5 * G + s], pixn = pq[6 * G + s];
  const double piyy = pq[7 * G + s], piyn = pq[8 * G + s], pinn = pq[9 * G + s];
  const double Pi = BULK ? (double)pq[10 * G + s] : 0.0;

  // pi:sigma, pi:pi, sigma:sigma (:174-197)
  const double ps = pitt * stt - 2 * pitx * stx - 2 * pity * sty + pixx * sxx + 2 * pixy * sxy +
                    piyy * syy - 2 * pitn * stn * t2 + 2 * pixn * sxn * t2 + 2 * piyn * syn * t2 +
                    pinn * snn * t2 * t2;
  const double ps3 = ps / 3;
  const double pipi = pitt * pitt - 2 * pitx * pitx - 2 * pity * pity + pixx * pixx +
                      2 * pixy * pixy + piyy * piyy - 2 * pitn * pitn * t2 + 2 * pixn * pixn * t2 +
                      2 * piyn * piyn * t2 + pinn * pinn * t2 * t2;
  const double ss = stt * stt - 2 * stx * stx - 2 * sty * sty + sxx * sxx + 2 * sxy * sxy +
                    syy * syy - 2 * stn * stn * t2 + 2 * sxn * sxn * t2 + 2 * syn * syn * t2 +
                    snn * snn * t2 * t2;

  // J^{mu nu} = delta I2 - I3 + tau I4 - lambda Pi sigma (:148-208), one
  // component at a time, folded into J:J in the reference's term order
  constexpr double delta_pipi = 1.33333, tau_pipi = 1.42857, lambda_piPi = 1.2;
  constexpr double delta_PiPi = 0.666667;  // TransportCoefficients.cuh:14-18
  auto Jc = [&](double pimn, double I3, double I4, double smn) {
    return delta_pipi * (theta * pimn) - I3 + tau_pipi * I4 - lambda_piPi * Pi * smn;
  };
  const double Jtt = Jc(pitt, 2 * (pitx * wtx + pity * wty + pitn * wtn),
                        (pitt * stt - pitx * stx - pity * sty - t2 * pitn * stn) - (1 - ut2) * ps3, stt);
  const double Jtx = Jc(pitx, pitt * wtx + pity * wxy + pitn * wxn + pixx * wtx + pixy * wty + pixn * wtn,
                        (pitt * stx + pitx * stt) / 2 - (pitx * sxx + pixx * stx) / 2 -
                            (pity * sxy + pixy * sty) / 2 - t2 * (pitn * sxn + pixn * stn) / 2 +
                            (ut * ux) * ps3, stx);
  const double Jty = Jc(pity, pitt * wty + pitx * wyx + pitn * wyn + pixy * wtx + piyy * wty + piyn * wtn,
                        (pitt * sty + pity * stt) / 2 - (pitx * sxy + pixy * stx) / 2 -
                            (pity * syy + piyy * sty) / 2 - t2 * (pitn * syn + piyn * stn) / 2 +
                            (ut * uy) * ps3, sty);
  const double Jtn = Jc(pitn, pitt * wnt + pitx * wnx + pity * wny + pixn * wtx + piyn * wty + pinn * wtn,
                        (pitt * stn + pitn * stt) / 2 - (pitx * sxn + pixn * stx) / 2 -
                            (pity * syn + piyn * sty) / 2 - t2 * (pitn * snn + pinn * stn) / 2 +
                            (ut * un) * ps3, stn);
  const double Jxx = Jc(pixx, 2 * (pitx * wtx + pixy * wxy + pixn * wxn),
                        (pitx * stx - pixx * sxx - pixy * sxy - t2 * pixn * sxn) + (1 + ux * ux) * ps3, sxx);
  const double Jxy = Jc(pixy, pitx * wty + pity * wtx + pixx * wyx + piyy * wxy + pixn * wyn + piyn * wxn,
                        (pitx * sty + pity * stx) / 2 - (pixx * sxy + pixy * sxx) / 2 -
                            (pixy * syy + piyy * sxy) / 2 - t2 * (pixn * syn + piyn * sxn) / 2 +
                            (ux * uy) * ps3, sxy);
  const double Jxn = Jc(pixn, pitx * wnt + pitn * wtx + pixx * wnx + pixy * wny + piyn * wxy + pinn * wxn,
                        (pitx * stn + pitn * stx) / 2 - (pixx * sxn + pixn * sxx) / 2 -
                            (pixy * syn + piyn * sxy) / 2 - t2 * (pixn * snn + pinn * sxn) / 2 +
                            (ux * un) * ps3, sxn);
  const double Jyy = Jc(piyy, 2 * (pity * wty + pixy * wyx + piyn * wyn),
                        (pity * sty - pixy * sxy - piyy * syy - t2 * piyn * syn) + (1 + uy * uy) * ps3, syy);
  const double Jyn = Jc(piyn, pity * wnt + pitn * wty + pixy * wnx + pixn * wyx + piyy * wny + pinn * wyn,
                        (pity * stn + pitn * sty) / 2 - (pixy * sxn + pixn * sxy) / 2 -
                            (piyy * syn + piyn * syy) / 2 - t2 * (piyn * snn + pinn * syn) / 2 +
                            (uy * un) * ps3, syn);
  const double Jnn = Jc(pinn, 2 * (pitn * wnt + pixn * wnx + piyn * wny),
                        (pitn * stn - pixn * sxn - piyn * syn - t2 * pinn * snn) + (1 / t2 + un2) * ps3, snn);
  const double J = -delta_PiPi * Pi * theta + lambda_Pipi * ps;
  const double JJ = Jtt * Jtt - 2 * Jtx * Jtx - 2 * Jty * Jty + Jxx * Jxx + 2 * Jxy * Jxy +
                    Jyy * Jyy - 2 * Jtn * Jtn * t2 + 2 * Jxn * Jxn * t2 + 2 * Jyn * Jyn * t2 +
                    Jnn * Jnn * t2 * t2;

  // (:214-227)
  double *o = A.out + s;
  o[GVH_V_RE_PI * G] = sqrt(fabs(pipi)) / p_s;
  o[GVH_V_RE2_PI * G] = T / 2 / etabar / (e_s + p_s) * sqrt(fabs(JJ / ss));
  o[GVH_V_RE_BULK * G] = fabs(Pi) / p_s;
  o[GVH_V_RE2_BULK * G] = fabs(J / zetabar / theta) * T / (e_s + p_s);
  o[GVH_V_KN_TAUPI * G] = 5 * etabar * theta / T;
  o[GVH_V_KN_TAUBULK * G] = theta / tauPiInv;
  o[GVH_V_TAUPI * G] = 5 * etabar / T;
  o[GVH_V_DXUX * G] = dxux;
  o[GVH_V_DYUY * G] = dyuy;
  o[GVH_V_THETA * G] = theta;
}

template <int EOS, bool BULK>
hipError_t launch_validity_b(const ValidityArgs &A, bool mixed, hipStream_t st) {
  if (mixed)
    hipLaunchKernelGGL((validity_kernel<EOS, BULK, float>), dim3(A.nblocks), dim3(VBLOCK), 0, st, A);
  else
    hipLaunchKernelGGL((validity_kernel<EOS, BULK, double>), dim3(A.nblocks), dim3(VBLOCK), 0, st, A);
  return hipGetLastError();
}

template <int EOS>
hipError_t launch_validity_e(const ValidityArgs &A, bool bulk, bool mixed, hipStream_t st) {
  return bulk ? launch_validity_b<EOS, true>(A, mixed, st)
              : launch_validity_b<EOS, false>(A, mixed, st);
}

}  // namespace
}  // namespace gvh

using namespace gvh;

extern "C" int gvh_validity_enable(gvh_context *c) {
  if (!c) return gvh_fail(GVH_EINVAL, "gvh_validity_enable: NULL context");
  if (c->validity) return GVH_OK;
  const size_t n = (size_t)GVH_VALIDITY_FIELDS * c->G;
  if (n > 0x7fffffffull) return gvh_fail(GVH_EINVAL, "validity arrays too large");
  double *d = nullptr;
  if (hipMalloc(&d, n * sizeof(double)) != hipSuccess) {
    (void)hipGetLastError();
    return gvh_fail(GVH_ENOMEM, "validity arrays: %zu bytes", n * sizeof(double));
  }
  // zero, and the regulation factors 1 (DynamicalVariables.cu:81)
  unsigned long long one;
  const double v1 = 1.0;
  memcpy(&one, &v1, sizeof one);
  hipError_t e = launch_fill_u64(reinterpret_cast<unsigned long long *>(d),
                                 (int)(GVH_V_REGULATIONS * c->G), 0ull, c->stream);
  if (e == hipSuccess)
    e = launch_fill_u64(reinterpret_cast<unsigned long long *>(d + GVH_V_REGULATIONS * c->G),
                        (int)c->G, one, c->stream);
  if (e != hipSuccess) {
    (void)hipFree(d);
    return gvh_fail(GVH_EHIP, "validity fill: %s", hipGetErrorString(e));
  }
  c->validity = d;
  return GVH_OK;
}

extern "C" int gvh_check_validity(gvh_context *c, double t, int flags) {
  if (!c) return gvh_fail(GVH_EINVAL, "gvh_check_validity: NULL context");
  if (!c->initialized) return gvh_fail(GVH_EINVAL, "state not initialised (gvh_set_initial_state)");
  if (flags & ~GVH_VALIDITY_CURRENT_PI) return gvh_fail(GVH_EINVAL, "unknown validity flags %d", flags);
  if (!(t > 0.0)) return gvh_fail(GVH_EINVAL, "validity needs a proper time > 0");
  int rc = gvh_validity_enable(c);
  if (rc) return rc;
  ValidityArgs A{};
  const bool prev = c->steps_done > 0 && !(flags & GVH_VALIDITY_CURRENT_PI);
  A.u = c->ubuf[c->iu];
  A.up = c->ubuf[c->iup];
  A.e = c->ebuf[c->ie];
  A.p = c->pbuf[c->ie];
  A.q = c->qbuf[prev ? c->iQ() : c->iq];
  A.out = c->validity;
  A.G = c->G;
  A.nx = c->nx;
  A.ny = c->ny;
  A.ncx = c->ncx;
  A.ncy = c->ncy;
  A.ncell = (unsigned)c->nx * c->ny * c->nz;
  A.nblocks = ((A.ncell + VBLOCK - 1) / VBLOCK + 7) / 8 * 8;
  A.t = t;
  A.dt = c->L.latticeSpacingProperTime;
  const double h[3] = {c->L.latticeSpacingX, c->L.latticeSpacingY, c->L.latticeSpacingRapidity};
  for (int d = 0; d < 3; ++d) A.fac[d] = 1 / h[d] / 2;
  A.etabar = c->Hp.shearViscosityToEntropyDensity;
  A.zeta_tpeak = c->zeta_tpeak;
  A.tab = c->tab;
  const bool mixed = c->precision != GVH_FP64;
  hipError_t e;
  switch (c->eos.kind) {
    case EOS_CONFORMAL: e = launch_validity_e<EOS_CONFORMAL>(A, c->bulk, mixed, c->stream); break;
    case EOS_TABLE: e = launch_validity_e<EOS_TABLE>(A, c->bulk, mixed, c->stream); break;
    default: e = launch_validity_e<EOS_WB>(A, c->bulk, mixed, c->stream); break;
  }
  if (e != hipSuccess) return gvh_fail(GVH_EHIP, "validity kernel: %s", hipGetErrorString(e));
  return GVH_OK;
}

extern "C" int gvh_get_validity(gvh_context *c, int field, int ghosted, double *host_out) {
  if (!c || !host_out) return gvh_fail(GVH_EINVAL, "gvh_get_validity: NULL argument");
  if (field < 0 || field >= GVH_VALIDITY_FIELDS)
    return gvh_fail(GVH_EINVAL, "unknown validity field %d", field);
  if (!c->validity)
    return gvh_fail(GVH_EINVAL, "validity arrays not enabled (gvh_check_validity)");
  return gvh_copy_ghosted_field(c, c->validity + (size_t)field * c->G, ghosted, host_out);
}
