// gvh_api.hip -- C ABI of the KT step: context, device state, stepping.
//
// Device state of one lattice (the module globals of the reference's
// DynamicalVariables.cu, made a context): three conserved buffers (q^n, the
// predictor qS, the corrector Q), three velocity buffers (u^n, u^{n-1}, u*)
// and two (e, p) buffers.  Buffers are var-major blocks of G doubles, G the
// ghosted lattice size, in the reference's column-major cell order
// s = i + ncx (j + ncy k) (DynamicalVariables.cu:33-35), so host transfers
// are plain copies.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gpuvh.h"
#include "gvh_device.h"
#include "gvh_internal.h"

namespace gvh {
hipError_t launch_stage(const StageArgs &A, int eos, bool bulk, bool stage2, int precision,
                        hipStream_t st);
hipError_t launch_count_above(const double *e, int nx, int ny, int nz, int ncx, int ncy, double thr,
                              unsigned long long *out, hipStream_t st);
hipError_t launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t st);
}  // namespace gvh

using namespace gvh;

static thread_local std::string g_err;

int gvh_fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

extern "C" const char *gvh_last_error(void) { return g_err.c_str(); }

namespace gvh {
// ------------------------------------------------------------- host EoS
// Host evaluation of equilibriumEnergyDensity (EquationOfState.cu:151-204)
// for the freeze-out and regulation thresholds; the tabulated form inverts
// the monotone T(e) table.
static double node_e(const HostEos &H, int k) {
  return std::ldexp(1.0 + (double)(k % H.M) / H.M, H.emin_exp + k / H.M);
}

double host_e_of_T(const HostEos &H, double T) {
  if (H.kind == EOS_CONFORMAL) return EOS_FACTOR * std::pow(T, 4.0);
  if (H.kind == EOS_TABLE) {
    const int N = H.n;
    const double *Tt = H.T.data();
    if (T <= Tt[0]) return T <= 0 ? 0.0 : node_e(H, 0) * std::pow(T / Tt[0], 4.0);
    if (T >= Tt[N - 1]) return node_e(H, N - 1) * std::pow(T / Tt[N - 1], 4.0);
    int lo = 0, hi = N - 1;
    while (hi - lo > 1) {
      const int mid = (lo + hi) / 2;
      if (Tt[mid] <= T) lo = mid; else hi = mid;
    }
    const double f = (T - Tt[lo]) / (Tt[hi] - Tt[lo]);
    return node_e(H, lo) + f * (node_e(H, hi) - node_e(H, lo));
  }
  static const double num[24] = {
      -0.011958188410851651, 119.89423098138208, -3156.9475699248055, 32732.86844374939,
      -187899.8994764422, 712537.3610845465, -1.557049803609345e6, 1.4852519861308339e6,
      532132.6079941876, -1.963099445042592e6, -4484.44579242679, 1.7984228830058286e6,
      119345.25619517374, -1.3499773937058165e6, -207838.4995663606, 654970.2138652403,
      -78643.00334616247, 40274.00078068926, 422619.58977657766, -409688.07836393174,
      -62005.75915066359, 46788.14270090656, 40784.330477857235, -12589.47744840392};
  static const double den[24] = {
      31630.074365558292, -127100.88940643385, 173528.1225422275, -39403.297956865215,
      -85582.57873541754, 9320.560804233442, 50882.74198960172, 20335.926473421183,
      -14897.725710713818, -23836.484117457, -13726.013896090335, 4517.908673107615,
      18056.19917986404, 14954.82860467155, 2569.623976952738, -9304.046211514986,
      -15606.429173842751, 8383.710735812094, 1591.3177623932843, -678.748230997762,
      -33.58687934953277, 3.2520554133126285, -0.19647288043440464, 0.005443394551264717};
  double a = num[0], b = den[0], Tk = 1.0;
  for (int k = 1; k < 24; ++k) {
    Tk *= T;
    a += num[k] * Tk;
    b += den[k] * Tk;
  }
  return a / b;
}

// Host evaluation of equilibriumPressure (EquationOfState.cu:31-80) with the
// same table lookup as the device (initial pressures, diagnostics).
double host_eos_p(const HostEos &H, double e) {
  if (H.kind == EOS_CONFORMAL) return e / 3.0;
  if (H.kind == EOS_TABLE) {
    const int N = H.n, M = H.M;
    const double e_first = node_e(H, 0), e_last = node_e(H, N - 1);
    int k;
    double fr;
    if (e >= e_first && e < e_last) {
      int E;
      const double m = std::frexp(e, &E);
      const double f = (2.0 * m - 1.0) * (double)M;
      const int j = (int)f;
      fr = f - (double)j;
      k = (E - 1 - H.emin_exp) * M + j;
      if (k > N - 2) {
        k = N - 2;
        fr += 1.0;
      }
    } else if (e >= e_last) {
      k = N - 2;
      fr = (e - node_e(H, N - 2)) / (e_last - node_e(H, N - 2));
    } else {
      k = 0;
      fr = (e - e_first) / (node_e(H, 1) - e_first);
    }
    return H.p[k] + fr * (H.p[k + 1] - H.p[k]);
  }
  static const double a[13] = {-0.25181736420168666, 9737.845799644809, 1.077580993288114e6,
                               3.1729694865420084e6, 1.6357487344679043e6, 334334.4309240126,
                               41913.439282708554, 6340.448389300905, 141.5073484468774,
                               0.7158279081255019, 0.0009417586777847889, 3.1188455176941583e-7,
                               1.9531729608963267e-11};
  static const double b[13] = {45829.44617893836, 4.0574329080826794e6, 2.0931169138134286e7,
                               1.3512402226067686e7, 1.7851642641834426e6, 278581.2989342773,
                               26452.34905933697, 499.04919730607065, 2.3405487982094204,
                               0.002962497695527404, 9.601103399348206e-7, 5.928138360995685e-11,
                               3.2581066229887368e-18};
  double ek[13];
  ek[0] = 1.0;
  ek[1] = e;
  for (int i = 2; i < 13; ++i) ek[i] = ek[i - 1] * e;
  double na = std::fma(a[1], ek[1], a[0]), nb = std::fma(b[1], ek[1], b[0]);
  for (int i = 2; i < 13; ++i) {
    na = std::fma(a[i], ek[i], na);
    nb = std::fma(b[i], ek[i], nb);
  }
  return na / nb;
}

// ---------------------------------------------------------- context

static int check_dims(const gvh_lattice_parameters *L) {
  if (!L) return gvh_fail(GVH_EINVAL, "lattice parameters are NULL");
  if (L->numLatticePointsX < 1 || L->numLatticePointsY < 1 || L->numLatticePointsRapidity < 1)
    return gvh_fail(GVH_EINVAL, "lattice dimensions must be positive (%d x %d x %d)",
                    L->numLatticePointsX, L->numLatticePointsY, L->numLatticePointsRapidity);
  const double G = (double)(L->numLatticePointsX + 4) * (L->numLatticePointsY + 4) *
                   (L->numLatticePointsRapidity + 4);
  // the kernels address a cell of one array by a 32-bit byte offset
  if (G * 8.0 >= 4294967296.0)
    return gvh_fail(GVH_EINVAL, "lattice too large: %.0f ghosted cells, at most 2^29 (an array "
                    "of doubles must stay below 4 GiB)", G);
  if (!(L->latticeSpacingX > 0 && L->latticeSpacingY > 0 && L->latticeSpacingRapidity > 0 &&
        L->latticeSpacingProperTime > 0))
    return gvh_fail(GVH_EINVAL, "lattice spacings must be positive");
  return GVH_OK;
}

int check_event_params(const gvh_lattice_parameters *L, const gvh_hydro_parameters *H) {
  int rc = check_dims(L);
  if (rc) return rc;
  if (!H) return gvh_fail(GVH_EINVAL, "hydro parameters are NULL");
  if (!(H->shearViscosityToEntropyDensity > 0))
    return gvh_fail(GVH_EINVAL, "eta/s must be positive");
  if (!(H->initialProperTimePoint > 0))
    return gvh_fail(GVH_EINVAL, "the initial proper time must be positive");
  return GVH_OK;
}

int eos_from_options(const gvh_options *o, HostEos &H) {
  H.kind = o ? o->eos : EOS_WB;
  if (H.kind == EOS_TABLE) {
    if (!o->tab_p || !o->tab_cs2 || !o->tab_T || o->tab_n < 2 || o->tab_per_octave < 1)
      return gvh_fail(GVH_EINVAL, "tabulated EoS needs tab_p, tab_cs2, tab_T, tab_n>=2");
    H.n = o->tab_n;
    H.M = o->tab_per_octave;
    H.emin_exp = o->tab_emin_exp;
    H.p.assign(o->tab_p, o->tab_p + H.n);
    H.cs2.assign(o->tab_cs2, o->tab_cs2 + H.n);
    H.T.assign(o->tab_T, o->tab_T + H.n);
    for (int k = 1; k < H.n; ++k)
      if (!(H.T[k] > H.T[k - 1])) return gvh_fail(GVH_EINVAL, "EoS table T(e) not increasing at %d", k);
  } else if (H.kind != EOS_CONFORMAL && H.kind != EOS_WB) {
    return gvh_fail(GVH_EINVAL, "unknown EoS kind %d", H.kind);
  }
  return GVH_OK;
}

int upload_eos_table(const HostEos &H, void **dev_pc, void **dev_tt, EosTable &T, hipStream_t st) {
  T = EosTable{};
  if (H.kind != EOS_TABLE) return GVH_OK;
  const int N = H.n;
  // one allocation: 4N doubles, then the same 4N values as floats
  std::vector<double> pc(4 * (size_t)N + 2 * (size_t)N), tt(2 * (size_t)N);
  float *pcf = reinterpret_cast<float *>(pc.data() + 4 * (size_t)N);
  for (int k = 0; k < N; ++k) {
    const int k1 = k + 1 < N ? k + 1 : k;
    pc[4 * k + 0] = H.p[k];
    pc[4 * k + 1] = H.p[k1] - H.p[k];
    pc[4 * k + 2] = H.cs2[k];
    pc[4 * k + 3] = H.cs2[k1] - H.cs2[k];
    tt[2 * k + 0] = H.T[k];
    tt[2 * k + 1] = H.T[k1] - H.T[k];
    for (int m = 0; m < 4; ++m) pcf[4 * k + m] = (float)pc[4 * k + m];
  }
  HIP_TRY(hipMalloc(dev_pc, pc.size() * sizeof(double)));
  HIP_TRY(hipMalloc(dev_tt, tt.size() * sizeof(double)));
  HIP_TRY(hipMemcpyAsync(*dev_pc, pc.data(), pc.size() * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(*dev_tt, tt.data(), tt.size() * sizeof(double), hipMemcpyHostToDevice, st));
  HIP_TRY(hipStreamSynchronize(st));
  T.pc = (const double4 *)*dev_pc;
  T.pcf = (const float4 *)((const double *)*dev_pc + 4 * (size_t)N);
  T.tt = (const double2 *)*dev_tt;
  T.n = N;
  T.M = H.M;
  T.emin_exp = H.emin_exp;
  T.e_first = node_e(H, 0);
  T.e_second = node_e(H, 1);
  T.e_pen = node_e(H, N - 2);
  T.e_last = node_e(H, N - 1);
  T.inv_first_span = 1.0 / (T.e_second - T.e_first);
  T.inv_last_span = 1.0 / (T.e_last - T.e_pen);
  T.f_first = (float)T.e_first;
  T.f_pen = (float)T.e_pen;
  T.f_last = (float)T.e_last;
  T.f_inv_first_span = (float)T.inv_first_span;
  T.f_inv_last_span = (float)T.inv_last_span;
  return GVH_OK;
}

}  // namespace gvh

extern "C" int gvh_create(const gvh_lattice_parameters *lattice, const gvh_hydro_parameters *hydro,
                          const gvh_options *opts, gvh_context **out) {
  if (!out) return gvh_fail(GVH_EINVAL, "out is NULL");
  *out = nullptr;
  int rc = check_event_params(lattice, hydro);
  if (rc) return rc;
  if (opts && opts->precision != GVH_FP64 && opts->precision != GVH_MIXED &&
      opts->precision != GVH_MIXED_F64_ROOTS)
    return gvh_fail(GVH_EINVAL, "unknown precision %d", opts->precision);

  gvh_context *c = new gvh_context();
  c->L = *lattice;
  c->Hp = *hydro;
  c->nx = lattice->numLatticePointsX;
  c->ny = lattice->numLatticePointsY;
  c->nz = lattice->numLatticePointsRapidity;
  c->ncx = c->nx + 4;
  c->ncy = c->ny + 4;
  c->ncz = c->nz + 4;
  c->G = (long long)c->ncx * c->ncy * c->ncz;
  c->bulk = opts ? (opts->bulk != 0) : true;
  c->precision = opts ? opts->precision : GVH_FP64;
  c->zeta_tpeak = (opts && opts->zeta_tpeak > 0) ? opts->zeta_tpeak : 1.01355;
  rc = eos_from_options(opts, c->eos);
  if (rc) {
    delete c;
    return rc;
  }
  if (opts && opts->stream) {
    c->stream = (hipStream_t)opts->stream;
    c->own_stream = false;
  } else {
    hipError_t e = hipStreamCreate(&c->stream);
    if (e != hipSuccess) {
      delete c;
      return gvh_fail(GVH_EHIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    c->own_stream = true;
  }
  // one allocation for all state: 3 conserved buffers, 3 x 4 velocity,
  // 2 x 2 (e, p); a mixed conserved buffer holds 4 G doubles + 11 G floats
  const size_t Gs = (size_t)c->G;
  c->qstride = c->precision != GVH_FP64 ? 4 * Gs + ((11 * Gs + 1) / 2 + 63) / 64 * 64 : 15 * Gs;
  const size_t bytes = (3 * c->qstride + 16 * Gs) * sizeof(double);
  hipError_t e = hipMalloc(&c->arena, bytes);
  if (e != hipSuccess) {
    gvh_destroy(c);
    return gvh_fail(GVH_ENOMEM, "hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
  }
  e = hipMemsetAsync(c->arena, 0, bytes, c->stream);
  if (e == hipSuccess) e = hipMalloc((void **)&c->diag, DIAG_N * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc((void **)&c->count, sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMemsetAsync(c->diag, 0, DIAG_N * sizeof(unsigned long long), c->stream);
  if (e == hipSuccess) e = hipEventCreate(&c->ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&c->ev[1]);
  if (e == hipSuccess) e = hipEventCreate(&c->ev[2]);
  if (e != hipSuccess) {
    gvh_destroy(c);
    return gvh_fail(GVH_EHIP, "context set-up: %s", hipGetErrorString(e));
  }
  double *a = (double *)c->arena;
  for (int b = 0; b < 3; ++b) c->qbuf[b] = a + (size_t)b * c->qstride;
  a += 3 * c->qstride;
  for (int b = 0; b < 3; ++b) c->ubuf[b] = a + (size_t)b * 4 * c->G;
  a += 12 * (size_t)c->G;
  for (int b = 0; b < 2; ++b) {
    c->ebuf[b] = a + (size_t)(2 * b) * c->G;
    c->pbuf[b] = a + (size_t)(2 * b + 1) * c->G;
  }
  rc = upload_eos_table(c->eos, &c->tab_pc, &c->tab_tt, c->tab, c->stream);
  if (rc) {
    gvh_destroy(c);
    return rc;
  }
  const double hbarc = 0.197326938;  // HydroPlugin.cpp:132
  c->e_reg = host_e_of_T(c->eos, hydro->regulationTemperatureGeV / hbarc);
  c->iq = 0;
  c->iu = 0;
  c->iup = 1;
  // kernel arrangement and traversal strip height; GVH_VARIANT / GVH_STRIP
  // override them for tuning sweeps
  if (const char *v = getenv("GVH_VARIANT")) c->variant = atoi(v);
  if (const char *v = getenv("GVH_STRIP")) c->strip = atoi(v);
  if (c->strip < 1) c->strip = 1;
  if (c->strip > c->ny) c->strip = c->ny;
  *out = c;
  return GVH_OK;
}

extern "C" int gvh_destroy(gvh_context *c) {
  if (!c) return GVH_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int k = 0; k < 3; ++k)
    if (c->ev[k]) (void)hipEventDestroy(c->ev[k]);
  if (c->arena) (void)hipFree(c->arena);
  if (c->diag) (void)hipFree(c->diag);
  if (c->count) (void)hipFree(c->count);
  if (c->tab_pc) (void)hipFree(c->tab_pc);
  if (c->tab_tt) (void)hipFree(c->tab_tt);
  if (c->validity) (void)hipFree(c->validity);
  fo_free(c->fo);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return GVH_OK;
}

// ghost fill on the host copy of the initial state (GhostCells.cu:66-144)
static void host_ghosts(std::vector<double *> &arrs, int nx, int ny, int nz, int ncx, int ncy) {
  auto id = [&](int i, int j, int k) { return (size_t)i + (size_t)ncx * (j + (size_t)ncy * k); };
  for (double *a : arrs) {
    for (int k = 2; k < nz + 2; ++k)
      for (int j = 2; j < ny + 2; ++j) {
        a[id(0, j, k)] = a[id(1, j, k)] = a[id(2, j, k)];
        a[id(nx + 2, j, k)] = a[id(nx + 3, j, k)] = a[id(nx + 1, j, k)];
      }
    for (int k = 2; k < nz + 2; ++k)
      for (int i = 2; i < nx + 2; ++i) {
        a[id(i, 0, k)] = a[id(i, 1, k)] = a[id(i, 2, k)];
        a[id(i, ny + 2, k)] = a[id(i, ny + 3, k)] = a[id(i, ny + 1, k)];
      }
    for (int j = 2; j < ny + 2; ++j)
      for (int i = 2; i < nx + 2; ++i) {
        a[id(i, j, 0)] = a[id(i, j, 1)] = a[id(i, j, 2)];
        a[id(i, j, nz + 2)] = a[id(i, j, nz + 3)] = a[id(i, j, nz + 1)];
      }
  }
}

extern "C" int gvh_set_initial_state(gvh_context *c, double t0, const double *e, const double *p,
                                     const double *ut, const double *ux, const double *uy,
                                     const double *un, const double *pi10, const double *Pi) {
  (void)t0;
  if (!c || !e || !p || !ut || !ux || !uy || !un)
    return gvh_fail(GVH_EINVAL, "gvh_set_initial_state: NULL argument");
  const int nx = c->nx, ny = c->ny, nz = c->nz, ncx = c->ncx, ncy = c->ncy;
  const size_t G = (size_t)c->G, nint = (size_t)nx * ny * nz;
  // host image of one conserved buffer, e, p and u (zero outside the interior)
  std::vector<double> hq(15 * G, 0.0), he(G, 0.0), hp(G, 0.0), hu(4 * G, 0.0), hup(4 * G, 0.0);
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const size_t a = i + (size_t)nx * (j + (size_t)ny * k);
        const size_t s = (i + 2) + (size_t)ncx * ((j + 2) + (size_t)ncy * (k + 2));
        const double Pis = (Pi && c->bulk) ? Pi[a] : 0.0;
        he[s] = e[a];
        hp[s] = p[a];
        hu[s] = ut[a];
        hu[G + s] = ux[a];
        hu[2 * G + s] = uy[a];
        hu[3 * G + s] = un[a];
        for (int m = 0; m < 10; ++m) hq[(4 + m) * G + s] = pi10 ? pi10[m * nint + a] : 0.0;
        if (c->bulk) hq[14 * G + s] = Pis;
        // setConservedVariables with Ttt..Ttn (EnergyMomentumTensor.cu:159-177)
        const double P = p[a] + Pis;
        hq[0 * G + s] = (e[a] + P) * ut[a] * ut[a] - P + hq[4 * G + s];
        hq[1 * G + s] = (e[a] + P) * ut[a] * ux[a] + hq[5 * G + s];
        hq[2 * G + s] = (e[a] + P) * ut[a] * uy[a] + hq[6 * G + s];
        hq[3 * G + s] = (e[a] + P) * ut[a] * un[a] + hq[7 * G + s];
      }
  hup = hu;  // the previous-step velocity has zero ghosts (DynamicalVariables.cu:280-284)
  std::vector<double *> arrs;
  for (int n = 0; n < 15; ++n) arrs.push_back(hq.data() + n * G);
  arrs.push_back(he.data());
  arrs.push_back(hp.data());
  for (int m = 0; m < 4; ++m) arrs.push_back(hu.data() + m * G);
  host_ghosts(arrs, nx, ny, nz, ncx, ncy);

  c->iq = 0;
  c->iu = 0;
  c->iup = 1;
  c->ie = 0;
  const size_t B = sizeof(double);
  std::vector<float> hqf;
  if (c->precision != GVH_FP64) {
    // T^{tau mu} in FP64, pi and Pi stored in FP32
    hqf.assign(hq.begin() + 4 * G, hq.end());
    HIP_TRY(hipMemcpyAsync(c->qbuf[c->iq], hq.data(), 4 * G * B, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipMemcpyAsync(c->qbuf[c->iq] + 4 * G, hqf.data(), 11 * G * sizeof(float),
                           hipMemcpyHostToDevice, c->stream));
  } else {
    HIP_TRY(hipMemcpyAsync(c->qbuf[c->iq], hq.data(), 15 * G * B, hipMemcpyHostToDevice, c->stream));
  }
  HIP_TRY(hipMemcpyAsync(c->ebuf[c->ie], he.data(), G * B, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->pbuf[c->ie], hp.data(), G * B, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->ubuf[c->iu], hu.data(), 4 * G * B, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(c->ubuf[c->iup], hup.data(), 4 * G * B, hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->initialized = true;
  c->steps_done = 0;
  return GVH_OK;
}

// Device address of field f of the current state; *is_float tells a pi / Pi
// field stored in FP32 (mixed form).
static void *field_dev(gvh_context *c, int f, bool *is_float) {
  *is_float = false;
  if (f >= GVH_TTT && f <= GVH_PI) {
    double *q = c->qbuf[c->iq];
    if (c->precision != GVH_FP64 && f >= GVH_PITT) {
      *is_float = true;
      return reinterpret_cast<float *>(q + 4 * c->G) + (size_t)(f - GVH_PITT) * c->G;
    }
    return q + (size_t)f * c->G;
  }
  if (f == GVH_E) return c->ebuf[c->ie];
  if (f == GVH_P) return c->pbuf[c->ie];
  if (f >= GVH_UT && f <= GVH_UN) return c->ubuf[c->iu] + (size_t)(f - GVH_UT) * c->G;
  if (f >= GVH_UTP && f <= GVH_UNP) return c->ubuf[c->iup] + (size_t)(f - GVH_UTP) * c->G;
  return nullptr;
}

int gvh_copy_q_field(gvh_context *c, const double *qbase, int f, double *host) {
  const size_t G = (size_t)c->G;
  if (c->precision != GVH_FP64 && f >= GVH_PITT) {
    std::vector<float> tmp(G);
    HIP_TRY(hipMemcpyAsync(tmp.data(), reinterpret_cast<const float *>(qbase + 4 * G) + (f - GVH_PITT) * G,
                           G * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < G; ++i) host[i] = tmp[i];
    return GVH_OK;
  }
  HIP_TRY(hipMemcpyAsync(host, qbase + f * G, G * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GVH_OK;
}

extern "C" int gvh_set_field(gvh_context *c, int field, const double *host_ghosted) {
  if (!c || !host_ghosted) return gvh_fail(GVH_EINVAL, "gvh_set_field: NULL argument");
  bool fl;
  void *d = field_dev(c, field, &fl);
  if (!d) return gvh_fail(GVH_EINVAL, "unknown field %d", field);
  if (fl) {
    std::vector<float> tmp(host_ghosted, host_ghosted + c->G);
    HIP_TRY(hipMemcpyAsync(d, tmp.data(), (size_t)c->G * sizeof(float), hipMemcpyHostToDevice,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GVH_OK;
  }
  HIP_TRY(hipMemcpyAsync(d, host_ghosted, (size_t)c->G * sizeof(double), hipMemcpyHostToDevice,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GVH_OK;
}

extern "C" int gvh_get_field(gvh_context *c, int field, int ghosted, double *host_out) {
  if (!c || !host_out) return gvh_fail(GVH_EINVAL, "gvh_get_field: NULL argument");
  bool fl;
  void *dv = field_dev(c, field, &fl);
  if (!dv) return gvh_fail(GVH_EINVAL, "unknown field %d", field);
  if (fl) {
    std::vector<double> full(c->G);
    int rc = gvh_copy_q_field(c, c->qbuf[c->iq], field, full.data());
    if (rc) return rc;
    if (ghosted) {
      std::copy(full.begin(), full.end(), host_out);
    } else {
      for (int k = 0; k < c->nz; ++k)
        for (int j = 0; j < c->ny; ++j)
          for (int i = 0; i < c->nx; ++i)
            host_out[i + (size_t)c->nx * (j + (size_t)c->ny * k)] =
                full[(i + 2) + (size_t)c->ncx * ((j + 2) + (size_t)c->ncy * (k + 2))];
    }
    return GVH_OK;
  }
  return gvh_copy_ghosted_field(c, (const double *)dv, ghosted, host_out);
}

// One ghosted block of G doubles to the host: the whole block with one copy,
// then (ghosted == 0) the interior cells picked on the host.  Strided 2-D
// copies into pageable memory returned stale rows when a profiler was
// attached, so none are used.
int gvh_copy_ghosted_field(gvh_context *c, const double *d, int ghosted, double *host_out) {
  if (ghosted) {
    HIP_TRY(hipMemcpyAsync(host_out, d, (size_t)c->G * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return GVH_OK;
  }
  std::vector<double> full(c->G);
  HIP_TRY(hipMemcpyAsync(full.data(), d, (size_t)c->G * sizeof(double), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int k = 0; k < c->nz; ++k)
    for (int j = 0; j < c->ny; ++j)
      std::copy_n(full.data() + 2 + (size_t)c->ncx * ((j + 2) + (size_t)c->ncy * (k + 2)), c->nx,
                  host_out + (size_t)c->nx * (j + (size_t)c->ny * k));
  return GVH_OK;
}

// Arguments of one stage launch for the current buffer roles.
namespace gvh {
Buffers context_buffers(const gvh_context *c) {
  Buffers B;
  for (int b = 0; b < 3; ++b) {
    B.q[b] = c->qbuf[b];
    B.u[b] = c->ubuf[b];
  }
  for (int b = 0; b < 2; ++b) {
    B.e[b] = c->ebuf[b];
    B.p[b] = c->pbuf[b];
  }
  return B;
}

Roles context_roles(const gvh_context *c) { return Roles{c->iq, c->iu, c->iup, c->ie}; }

StageArgs make_stage_args(gvh_context *c, int stage, double t) {
  StageArgs A{};
  stage_pointers(A, context_buffers(c), context_roles(c), stage);
  A.diag = c->diag;
  A.reg = (stage == 2 && c->validity) ? c->validity + GVH_V_REGULATIONS * c->G : nullptr;
  A.G = c->G;
  A.nx = c->nx;
  A.ny = c->ny;
  A.nz = c->nz;
  A.ncx = c->ncx;
  A.ncy = c->ncy;
  const double dt = c->L.latticeSpacingProperTime;
  A.dt = dt;
  A.inv_dt = 1.0 / dt;
  stage_times(A, t, stage);
  const double h[3] = {c->L.latticeSpacingX, c->L.latticeSpacingY, c->L.latticeSpacingRapidity};
  for (int d = 0; d < 3; ++d) {
    A.h[d] = h[d];
    A.inv_h[d] = 1.0 / h[d];
    A.half_inv_h[d] = 1.0 / h[d] / 2;  // 1 / d_dx / 2, SourceTerms.cu:449
  }
  A.etabar = c->Hp.shearViscosityToEntropyDensity;
  A.zeta_tpeak = c->zeta_tpeak;
  A.e_reg = c->e_reg;
  A.tab = c->tab;
  A.strip = c->strip;
  A.nstrips = (c->ny + c->strip - 1) / c->strip;
  A.variant = c->variant;
  A.max_iters = c->max_iters;
  return A;
}
}  // namespace gvh

int launch_stage(gvh_context *c, const StageArgs &A, bool stage2);

// Buffer rotation of one step: stage 1 writes qS, e/p[1-ie], u*; stage 2
// writes Q, e/p[ie], and u^{n+1} into the u^{n-1} buffer; then the roles
// swap like setCurrentConservedVariables and swapFluidVelocity.
static void rotate_after_step(gvh_context *c) {
  Roles R = context_roles(c);
  rotate_roles(R);
  c->iq = R.iq;
  c->iu = R.iu;
  c->iup = R.iup;
  ++c->steps_done;
}

static int step_impl(gvh_context *c, double t, float *ms1, float *ms2) {
  if (!c) return gvh_fail(GVH_EINVAL, "context is NULL");
  if (!c->initialized) return gvh_fail(GVH_EINVAL, "state not initialised (gvh_set_initial_state)");
  const bool timed = ms1 || ms2;
  if (timed) HIP_TRY(hipEventRecord(c->ev[0], c->stream));
  StageArgs A1 = make_stage_args(c, 1, t);
  int rc = launch_stage(c, A1, false);
  if (rc) return rc;
  if (timed) HIP_TRY(hipEventRecord(c->ev[1], c->stream));
  StageArgs A2 = make_stage_args(c, 2, t);
  rc = launch_stage(c, A2, true);
  if (rc) return rc;
  if (timed) {
    HIP_TRY(hipEventRecord(c->ev[2], c->stream));
    HIP_TRY(hipEventSynchronize(c->ev[2]));
    float a = 0, b = 0;
    HIP_TRY(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
    HIP_TRY(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
    if (ms1) *ms1 = a;
    if (ms2) *ms2 = b;
  }
  rotate_after_step(c);
  return GVH_OK;
}

int launch_stage(gvh_context *c, const StageArgs &A, bool stage2) {
  hipError_t e = launch_stage(A, c->eos.kind, c->bulk, stage2, c->precision, c->stream);
  if (e != hipSuccess) return gvh_fail(GVH_EHIP, "stage kernel launch: %s", hipGetErrorString(e));
  return GVH_OK;
}

extern "C" int gvh_step(gvh_context *c, double t) { return step_impl(c, t, nullptr, nullptr); }

extern "C" int gvh_steps(gvh_context *c, double t0, int n_begin, int nsteps) {
  if (!c) return gvh_fail(GVH_EINVAL, "context is NULL");
  if (nsteps < 0) return gvh_fail(GVH_EINVAL, "nsteps must be >= 0");
  const double dt = c->L.latticeSpacingProperTime;
  for (int n = n_begin; n < n_begin + nsteps; ++n) {
    int rc = step_impl(c, t0 + n * dt, nullptr, nullptr);
    if (rc) return rc;
  }
  return GVH_OK;
}

// nsteps steps with an event before the first stage and after every stage,
// one wait at the end: the stage kernel times of exactly the steps that the
// caller's clock brackets.
extern "C" int gvh_steps_timed(gvh_context *c, double t0, int n_begin, int nsteps,
                               float *ms_stage1, float *ms_stage2, float *ms_total) {
  if (!c) return gvh_fail(GVH_EINVAL, "context is NULL");
  if (!c->initialized) return gvh_fail(GVH_EINVAL, "state not initialised (gvh_set_initial_state)");
  if (nsteps < 1) return gvh_fail(GVH_EINVAL, "nsteps must be >= 1");
  const double dt = c->L.latticeSpacingProperTime;
  std::vector<hipEvent_t> ev(2 * (size_t)nsteps + 1, nullptr);
  int rc = GVH_OK;
  hipError_t e = hipSuccess;
  for (auto &x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  if (e == hipSuccess) e = hipEventRecord(ev[0], c->stream);
  for (int k = 0; k < nsteps && e == hipSuccess && !rc; ++k) {
    const double t = t0 + (n_begin + k) * dt;
    StageArgs A1 = make_stage_args(c, 1, t);
    if ((rc = launch_stage(c, A1, false))) break;
    if ((e = hipEventRecord(ev[2 * k + 1], c->stream)) != hipSuccess) break;
    StageArgs A2 = make_stage_args(c, 2, t);
    if ((rc = launch_stage(c, A2, true))) break;
    if ((e = hipEventRecord(ev[2 * k + 2], c->stream)) != hipSuccess) break;
    rotate_after_step(c);
  }
  float s1 = 0, s2 = 0, tot = 0;
  if (!rc && e == hipSuccess) e = hipEventSynchronize(ev[2 * nsteps]);
  for (int k = 0; k < nsteps && !rc && e == hipSuccess; ++k) {
    float a = 0, b = 0;
    e = hipEventElapsedTime(&a, ev[2 * k], ev[2 * k + 1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[2 * k + 1], ev[2 * k + 2]);
    s1 += a;
    s2 += b;
  }
  if (!rc && e == hipSuccess) e = hipEventElapsedTime(&tot, ev[0], ev[2 * nsteps]);
  for (auto &x : ev)
    if (x) (void)hipEventDestroy(x);
  if (rc) return rc;
  if (e != hipSuccess) return gvh_fail(GVH_EHIP, "gvh_steps_timed: %s", hipGetErrorString(e));
  if (ms_stage1) *ms_stage1 = s1;
  if (ms_stage2) *ms_stage2 = s2;
  if (ms_total) *ms_total = tot;
  return GVH_OK;
}

extern "C" int gvh_step_timed(gvh_context *c, double t, float *ms1, float *ms2) {
  float a = 0, b = 0;
  int rc = step_impl(c, t, &a, &b);
  if (ms1) *ms1 = a;
  if (ms2) *ms2 = b;
  return rc;
}

extern "C" int gvh_synchronize(gvh_context *c) {
  if (!c) return gvh_fail(GVH_EINVAL, "context is NULL");
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GVH_OK;
}

extern "C" int gvh_count_above(gvh_context *c, double thr, long long *count) {
  if (!c || !count) return gvh_fail(GVH_EINVAL, "gvh_count_above: NULL argument");
  HIP_TRY(launch_fill_u64(c->count, 1, 0ull, c->stream));
  HIP_TRY(launch_count_above(c->ebuf[c->ie], c->nx, c->ny, c->nz, c->ncx, c->ncy, thr, c->count,
                             c->stream));
  unsigned long long h = 0;
  HIP_TRY(hipMemcpyAsync(&h, c->count, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  *count = (long long)h;
  return GVH_OK;
}

extern "C" int gvh_get_diagnostics(gvh_context *c, gvh_diagnostics *out) {
  if (!c || !out) return gvh_fail(GVH_EINVAL, "gvh_get_diagnostics: NULL argument");
  unsigned long long h[DIAG_N] = {0};
  HIP_TRY(hipMemcpyAsync(h, c->diag, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  out->newton_not_converged = h[DIAG_NOT_CONVERGED];
  out->e_nan = h[DIAG_E_NAN];
  out->e_clamped = h[DIAG_E_CLAMPED];
  return GVH_OK;
}

extern "C" int gvh_energy_density_of_T(gvh_context *c, double T, double *e) {
  if (!c || !e) return gvh_fail(GVH_EINVAL, "gvh_energy_density_of_T: NULL argument");
  *e = host_e_of_T(c->eos, T);
  return GVH_OK;
}

extern "C" int gvh_equilibrium_pressure(gvh_context *c, long long n, const double *e, double *p) {
  if (!c || (n > 0 && (!e || !p))) return gvh_fail(GVH_EINVAL, "gvh_equilibrium_pressure: NULL argument");
  for (long long i = 0; i < n; ++i) p[i] = host_eos_p(c->eos, e[i]);
  return GVH_OK;
}

extern "C" int gvh_dims(gvh_context *c, int *nx, int *ny, int *nz) {
  if (!c) return gvh_fail(GVH_EINVAL, "context is NULL");
  if (nx) *nx = c->nx;
  if (ny) *ny = c->ny;
  if (nz) *nz = c->nz;
  return GVH_OK;
}
