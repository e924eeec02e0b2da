// gvh_run.cpp -- one hydrodynamic event from files to freeze-out surface:
// the run loop of HydroPlugin.cpp:107-434 and HYDRO::run_hydro
// (HydroWrapper.cpp:17-87) over the device context.
//
// Differences from the reference, all at its undefined or crashing edges:
// a missing or malformed input file is GVH_EIO (the reference prints and then
// calls fclose(NULL)); initial-condition types other than the file readers 10
// and 11 are GVH_EINVAL; the surface file is <root_dir>/output/surface.dat
// (the reference opens "output/surface.dat" relative to the working
// directory, which its driver also passes as the root directory).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include <unistd.h>

#include "../../include/gpuvh.h"
#include "gvh_internal.h"

namespace {

constexpr double HBARC = 0.197326938;  // HydroPlugin.cpp:132

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point a) {
  return std::chrono::duration<double, std::milli>(Clock::now() - a).count();
}

struct Ctx {
  gvh_context *c = nullptr;
  ~Ctx() {
    if (c) gvh_destroy(c);
  }
};

struct File {
  FILE *f = nullptr;
  ~File() {
    if (f) fclose(f);
  }
};

// Interior arrays of the initial state, x fastest.
struct InitialState {
  std::vector<double> e, p, ut, ux, uy, un, pi10, Pi;
  explicit InitialState(size_t n)
      : e(n), p(n), ut(n), ux(n), uy(n), un(n), pi10(10 * n), Pi(n) {}
};

// One "x y z value" file per interior cell, rows in the reference's loop
// order (x outermost, eta innermost), values parsed as float
// (InitialConditions.cpp:98-489).
int read_column_file(const std::string &path, int nx, int ny, int nz, double *out, double add) {
  File F;
  F.f = fopen(path.c_str(), "r");
  if (!F.f) return gvh_fail(GVH_EIO, "cannot open %s", path.c_str());
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        float x, y, z, v;
        if (fscanf(F.f, "%f %f %f %f\n", &x, &y, &z, &v) != 4)
          return gvh_fail(GVH_EIO, "%s: malformed or short at cell (%d, %d, %d)", path.c_str(), i,
                          j, k);
        out[i + (size_t)nx * (j + (size_t)ny * k)] = (double)v + add;
      }
  return GVH_OK;
}

// setInitialTmunuFromFiles (InitialConditions.cpp:98-489): e.dat (+1e-3,
// p from the EoS), ut/ux/uy/un.dat, the ten pi*.dat and, with bulk, bulk.dat.
int read_ic_files(gvh_context *c, const std::string &root, InitialState &S) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const size_t n = (size_t)nx * ny * nz;
  const std::string in = root + "/input/";
  int rc = read_column_file(in + "e.dat", nx, ny, nz, S.e.data(), 1.0e-3);
  if (rc) return rc;
  for (size_t s = 0; s < n; ++s) S.p[s] = gvh::host_eos_p(c->eos, S.e[s]);
  const char *u_names[4] = {"ut", "ux", "uy", "un"};
  double *u_arrs[4] = {S.ut.data(), S.ux.data(), S.uy.data(), S.un.data()};
  for (int k = 0; k < 4; ++k)
    if ((rc = read_column_file(in + u_names[k] + ".dat", nx, ny, nz, u_arrs[k], 0.0))) return rc;
  const char *pi_names[10] = {"pitt", "pitx", "pity", "pitn", "pixx",
                              "pixy", "pixn", "piyy", "piyn", "pinn"};
  for (int k = 0; k < 10; ++k)
    if ((rc = read_column_file(in + pi_names[k] + ".dat", nx, ny, nz, S.pi10.data() + k * n, 0.0)))
      return rc;
  if (c->bulk && (rc = read_column_file(in + "bulk.dat", nx, ny, nz, S.Pi.data(), 0.0))) return rc;
  return GVH_OK;
}

// setInitialTmunuFromFile (InitialConditions.cpp:29-95): one Tmunu.dat row of
// 20 floats per cell, x y z e p ut ux uy un pitt..pinn Pi; e += 1e-3.
int read_tmunu_file(gvh_context *c, const std::string &root, InitialState &S) {
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const size_t n = (size_t)nx * ny * nz;
  const std::string path = root + "/input/Tmunu.dat";
  File F;
  F.f = fopen(path.c_str(), "r");
  if (!F.f) return gvh_fail(GVH_EIO, "cannot open %s", path.c_str());
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        float v[20];
        const int got = fscanf(F.f,
                               "%f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f\n",
                               &v[0], &v[1], &v[2], &v[3], &v[4], &v[5], &v[6], &v[7], &v[8],
                               &v[9], &v[10], &v[11], &v[12], &v[13], &v[14], &v[15], &v[16],
                               &v[17], &v[18], &v[19]);
        if (got != 20)
          return gvh_fail(GVH_EIO, "%s: malformed or short at cell (%d, %d, %d)", path.c_str(), i,
                          j, k);
        const size_t s = i + (size_t)nx * (j + (size_t)ny * k);
        S.e[s] = (double)v[3] + 1.0e-3;
        S.p[s] = v[4];
        S.ut[s] = v[5];
        S.ux[s] = v[6];
        S.uy[s] = v[7];
        S.un[s] = v[8];
        for (int m = 0; m < 10; ++m) S.pi10[m * n + s] = v[9 + m];
        S.Pi[s] = c->bulk ? v[19] : 0.0;
      }
  return GVH_OK;
}

// Whether the host view of u and pi is the previous step's: the reference
// copies them through the first-allocated buffers, which hold the previous
// step after an odd number of steps (DynamicalVariables.cu:310-331, 400-408).
bool stale_view(const gvh_context *c, int flags) {
  return !(flags & GVH_FO_CURRENT_FIELDS) && (c->steps_done & 1);
}

// output() (FileIO.cpp:15-60): one ghosted field h as "x y z value" lines of
// the interior cells into <outdir>/<name>_<t>.dat.
int write_field_file(const gvh_context *c, const std::vector<double> &h, const char *name,
                     double t, const std::string &outdir) {
  const int nx = c->nx, ny = c->ny, nz = c->nz, ncx = c->ncx, ncy = c->ncy;
  const double dx = c->L.latticeSpacingX, dy = c->L.latticeSpacingY,
               dz = c->L.latticeSpacingRapidity;
  char fname[1024];
  snprintf(fname, sizeof fname, "%s/%s_%.3f.dat", outdir.c_str(), name, t);
  File F;
  F.f = fopen(fname, "w");
  if (!F.f) return gvh_fail(GVH_EIO, "cannot write %s", fname);
  for (int k = 2; k < nz + 2; ++k) {
    const double z = (k - 2 - (nz - 1) / 2.) * dz;
    for (int j = 2; j < ny + 2; ++j) {
      const double y = (j - 2 - (ny - 1) / 2.) * dy;
      for (int i = 2; i < nx + 2; ++i) {
        const double x = (i - 2 - (nx - 1) / 2.) * dx;
        fprintf(F.f, "%.3f\t%.3f\t%.3f\t%.8f\n", x, y, z,
                h[i + (size_t)ncx * (j + (size_t)ncy * k)]);
      }
    }
  }
  return GVH_OK;
}

// outputDynamicalQuantities (HydroPlugin.cpp:44-85) with output()
// (FileIO.cpp:15-60): e, p, u, pi and Pi, one file per variable.
int write_snapshot(gvh_context *c, double t, const std::string &outdir, int flags) {
  const long long G = c->G;
  const bool stale = stale_view(c, flags);
  const double *ud = c->ubuf[stale ? c->iup : c->iu];
  const double *qd = c->qbuf[stale ? c->iQ() : c->iq];
  // u from the device buffer the host view reads; conserved fields by id
  struct Item {
    const char *name;
    const double *dev;  // velocity / e / p block, or nullptr for a conserved field
    int field;
  };
  std::vector<Item> items = {{"e", c->ebuf[c->ie], -1},   {"p", c->pbuf[c->ie], -1},
                             {"ux", ud + 1 * G, -1},       {"uy", ud + 2 * G, -1},
                             {"un", ud + 3 * G, -1},       {"ut", ud + 0 * G, -1},
                             {"pixx", nullptr, GVH_PIXX}, {"pixy", nullptr, GVH_PIXY},
                             {"pixn", nullptr, GVH_PIXN}, {"piyy", nullptr, GVH_PIYY},
                             {"piyn", nullptr, GVH_PIYN}, {"pitt", nullptr, GVH_PITT},
                             {"pitx", nullptr, GVH_PITX}, {"pity", nullptr, GVH_PITY},
                             {"pitn", nullptr, GVH_PITN}, {"pinn", nullptr, GVH_PINN}};
  if (c->bulk) items.push_back({"Pi", nullptr, GVH_PI});
  std::vector<double> h(G);
  for (const Item &it : items) {
    if (it.dev) {
      HIP_TRY(hipMemcpyAsync(h.data(), it.dev, G * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
      HIP_TRY(hipStreamSynchronize(c->stream));
    } else {
      int rc = gvh_copy_q_field(c, qd, it.field, h.data());
      if (rc) return rc;
    }
    int rc = write_field_file(c, h, it.name, t, outdir);
    if (rc) return rc;
  }
  return GVH_OK;
}

// The validity fields under the names of the reference's (commented-out)
// outputs, HydroPlugin.cpp:72-83, computed from the state at t as the
// reference's last checkValidity call left them.
int write_validity(gvh_context *c, double t, const std::string &outdir) {
  static const struct {
    const char *name;
    int field;
  } V[] = {{"regulations", GVH_V_REGULATIONS}, {"Rpi", GVH_V_RE_PI},
           {"R2pi", GVH_V_RE2_PI},             {"RPi", GVH_V_RE_BULK},
           {"R2Pi", GVH_V_RE2_BULK},           {"KnTaupi", GVH_V_KN_TAUPI},
           {"KnTauPi", GVH_V_KN_TAUBULK},      {"taupi", GVH_V_TAUPI},
           {"dxux", GVH_V_DXUX},               {"dyuy", GVH_V_DYUY},
           {"theta", GVH_V_THETA}};
  int rc = gvh_check_validity(c, t, 0);
  if (rc) return rc;
  std::vector<double> h(c->G);
  for (const auto &v : V) {
    if ((rc = gvh_get_validity(c, v.field, 1, h.data()))) return rc;
    if ((rc = write_field_file(c, h, v.name, t, outdir))) return rc;
  }
  return GVH_OK;
}

// Surface rows as text: each value as std::ostream prints a double by
// default (%g, 6 significant digits), space separated, one row per line
// (HydroPlugin.cpp:317-386).  Formatted in parallel, written in order.
int write_rows(FILE *f, const std::vector<double> &rows, long long nrows) {
  if (nrows == 0) return GVH_OK;
  const int nthreads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const long long chunk = (nrows + nthreads - 1) / nthreads;
  std::vector<std::string> out(nthreads);
  std::vector<std::thread> pool;
  for (int w = 0; w < nthreads; ++w)
    pool.emplace_back([&, w]() {
      const long long a = w * chunk, b = std::min(nrows, a + chunk);
      std::string &s = out[w];
      s.reserve((size_t)std::max(0ll, b - a) * 20 * 12);
      char buf[64];
      for (long long r = a; r < b; ++r)
        for (int k = 0; k < GVH_FO_COLUMNS; ++k) {
          const int len = snprintf(buf, sizeof buf, "%g", rows[(size_t)r * GVH_FO_COLUMNS + k]);
          s.append(buf, len);
          s.push_back(k + 1 < GVH_FO_COLUMNS ? ' ' : '\n');
        }
    });
  for (auto &t : pool) t.join();
  for (const std::string &s : out)
    if (!s.empty() && fwrite(s.data(), 1, s.size(), f) != s.size())
      return gvh_fail(GVH_EIO, "writing the surface file failed");
  return GVH_OK;
}

int center_values(gvh_context *c, double *e, double *p) {
  const int ictr = (c->nx % 2 == 0) ? c->ncx / 2 : (c->ncx - 1) / 2;
  const int jctr = (c->ny % 2 == 0) ? c->ncy / 2 : (c->ncy - 1) / 2;
  const int kctr = (c->nz % 2 == 0) ? c->ncz / 2 : (c->ncz - 1) / 2;
  const size_t s = ictr + (size_t)c->ncx * (jctr + (size_t)c->ncy * kctr);
  HIP_TRY(hipMemcpyAsync(e, c->ebuf[c->ie] + s, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(p, c->pbuf[c->ie] + s, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return GVH_OK;
}

}  // namespace

extern "C" void gvh_run_options_default(gvh_run_options *ro) {
  if (!ro) return;
  ro->freq = 10;
  ro->fofreq = 10;
  ro->fo_flags = 0;
  ro->snapshots = 1;
  ro->surface = 1;
  ro->verbose = 0;
  ro->validity = 0;
}

extern "C" int gvh_run(const gvh_lattice_parameters *lattice, const gvh_ic_parameters *ic,
                       const gvh_hydro_parameters *hydro, const char *root_dir,
                       const char *output_dir, const gvh_options *opts,
                       const gvh_run_options *ro_in, gvh_run_stats *stats) {
  const Clock::time_point t_begin = Clock::now();
  if (!lattice || !ic || !hydro || !opts || !root_dir)
    return gvh_fail(GVH_EINVAL, "gvh_run: NULL argument");
  gvh_run_options R;
  gvh_run_options_default(&R);
  if (ro_in) R = *ro_in;
  if (R.freq < 1) return gvh_fail(GVH_EINVAL, "freq must be >= 1");
  if (R.snapshots && !output_dir)
    return gvh_fail(GVH_EINVAL, "snapshots need an output directory");
  if (ic->initialConditionType != 10 && ic->initialConditionType != 11)
    return gvh_fail(GVH_EINVAL,
                    "initial condition type %d: only the file readers (10: input/e.dat etc., "
                    "11: input/Tmunu.dat) are supported",
                    ic->initialConditionType);
  Ctx C;
  int rc = gvh_create(lattice, hydro, opts, &C.c);
  if (rc) return rc;
  gvh_context *c = C.c;
  const int nt = lattice->numProperTimePoints;
  const double t0 = hydro->initialProperTimePoint;
  const double dt = lattice->latticeSpacingProperTime;
  const double e_fo = gvh::host_e_of_T(c->eos, hydro->freezeoutTemperatureGeV / HBARC);

  // initial state
  {
    InitialState S((size_t)c->nx * c->ny * c->nz);
    rc = ic->initialConditionType == 10 ? read_ic_files(c, root_dir, S)
                                        : read_tmunu_file(c, root_dir, S);
    if (rc) return rc;
    rc = gvh_set_initial_state(c, t0, S.e.data(), S.p.data(), S.ut.data(), S.ux.data(),
                               S.uy.data(), S.un.data(), S.pi10.data(), S.Pi.data());
    if (rc) return rc;
  }
  const bool validity = R.validity && R.snapshots;
  if (validity && (rc = gvh_validity_enable(c))) return rc;

  // freeze-out finder (HydroPlugin.cpp:152-212)
  const bool fo_dims = (c->nx > 1 && c->ny > 1);
  const bool surface = R.surface && fo_dims;
  if (R.surface && !fo_dims && R.verbose)
    printf("simulation is not in 3+1D or 2+1D; freezeout finder will not work!\n");
  File surf;
  if (surface) {
    if ((rc = gvh_freezeout_init(c, R.fofreq, e_fo, R.fo_flags))) return rc;
    const std::string path = std::string(root_dir) + "/output/surface.dat";
    surf.f = fopen(path.c_str(), "w");
    if (!surf.f) return gvh_fail(GVH_EIO, "cannot write %s", path.c_str());
  }
  if (R.verbose)
    printf("Grid size = %d x %d x %d\nfreezeout temperature = %.3f [fm^-1] (eF = %.3f [fm^-4])\n",
           c->nx, c->ny, c->nz, hydro->freezeoutTemperatureGeV / HBARC, e_fo);

  double t = t0;
  double ms_steps = 0, ms_fo = 0;
  long long rows_total = 0;
  int nsteps = 0, frozen_count = 0, frozen = 0;
  std::vector<double> rows;
  for (int n = 1; n <= nt + 1; ++n) {
    if ((n - 1) % R.freq == 0) {
      if (R.verbose) {
        double ec = 0, pc = 0;
        if ((rc = center_values(c, &ec, &pc))) return rc;
        printf("n = %d:%d (t = %.3f),\t (e, p) = (%.3f, %.3f) [GeV/fm^3]\n", n - 1, nt, t,
               ec * HBARC, pc * HBARC);
      }
      if (R.snapshots && (rc = write_snapshot(c, t, output_dir, R.fo_flags))) return rc;
      if (validity && (rc = write_validity(c, t, output_dir))) return rc;
    }
    if (surface) {
      const Clock::time_point a = Clock::now();
      if ((rc = gvh_freezeout_record(c, n))) return rc;
      if (n % R.fofreq == R.fofreq - 1) {
        long long ne = 0;
        if ((rc = gvh_freezeout_find(c, n, t, &ne))) return rc;
        rows.resize((size_t)ne * GVH_FO_COLUMNS);
        if ((rc = gvh_freezeout_rows(c, rows.data(), ne))) return rc;
        if ((rc = write_rows(surf.f, rows, ne))) return rc;
        rows_total += ne;
      }
      ms_fo += ms_since(a);
    }
    // termination (HydroPlugin.cpp:396-413): the count of frozen steps is
    // never reset
    long long above = 0;
    if ((rc = gvh_count_above(c, e_fo, &above))) return rc;
    if (above == 0) ++frozen_count;
    if (frozen_count >= R.fofreq + 1) {
      frozen = 1;
      if (R.verbose) printf("\nAll cells have dropped below freezeout energy density\n");
      break;
    }
    const Clock::time_point a = Clock::now();
    if ((rc = gvh_step(c, t))) return rc;
    if ((rc = gvh_synchronize(c))) return rc;
    ms_steps += ms_since(a);
    ++nsteps;
    t = t0 + n * dt;
  }
  if (surf.f && fflush(surf.f) != 0) return gvh_fail(GVH_EIO, "writing the surface file failed");
  if (R.verbose && nsteps) printf("Average time/step: %.3f ms\n", ms_steps / nsteps);
  if (stats) {
    stats->steps = nsteps;
    stats->frozen = frozen;
    stats->t_final = t;
    stats->surface_rows = rows_total;
    stats->ms_steps = ms_steps;
    stats->ms_freezeout = ms_fo;
    stats->ms_total = ms_since(t_begin);
  }
  return GVH_OK;
}

// Command line of CommandLineArguments.cpp:17-23: -t/--test and -h/--hydro
// take an optional attached argument, -o/--output and -c/--config a value.
extern "C" int gvh_run_hydro(int argc, char **argv, const gvh_options *opts,
                             const gvh_run_options *ro) {
  const char *config = nullptr, *output = nullptr;
  bool run_hydro = false;
  for (int i = 1; i < argc; ++i) {
    const char *a = argv[i];
    auto value = [&](const char *shortf, const char *longf, const char *&dst) -> int {
      const size_t ls = strlen(longf);
      if (!strcmp(a, shortf) || !strcmp(a, longf)) {
        if (i + 1 >= argc) return gvh_fail(GVH_EINVAL, "%s needs a value", a);
        dst = argv[++i];
        return 1;
      }
      if (!strncmp(a, longf, ls) && a[ls] == '=') {
        dst = a + ls + 1;
        return 1;
      }
      if (!strncmp(a, shortf, 2) && a[2]) {
        dst = a + 2;
        return 1;
      }
      return 0;
    };
    int r;
    if ((r = value("-c", "--config", config))) {
      if (r < 0) return r;
    } else if ((r = value("-o", "--output", output))) {
      if (r < 0) return r;
    } else if (!strncmp(a, "-h", 2) || !strncmp(a, "--hydro", 7)) {
      run_hydro = true;
    } else if (!strncmp(a, "-t", 2) || !strncmp(a, "--test", 6)) {
      // software tests of the reference's CLI: nothing to run here
    } else {
      return gvh_fail(GVH_EINVAL, "unknown argument %s", a);
    }
  }
  if (!run_hydro) return GVH_OK;
  if (!config) return gvh_fail(GVH_EINVAL, "--config DIR is required");
  gvh_lattice_parameters L;
  gvh_ic_parameters IC;
  gvh_hydro_parameters H;
  int rc;
  if ((rc = gvh_read_lattice_parameters(config, &L)) || (rc = gvh_read_ic_parameters(config, &IC)) ||
      (rc = gvh_read_hydro_parameters(config, &H)))
    return rc;
  char cwd[4096];
  if (!getcwd(cwd, sizeof cwd)) return gvh_fail(GVH_EIO, "getcwd failed");
  return gvh_run(&L, &IC, &H, cwd, output, opts, ro, nullptr);
}
