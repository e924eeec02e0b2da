// gvh_internal.h -- context layout shared by the library's host sources.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <vector>

#include "../../include/gpuvh.h"
#include "gvh_device.h"

namespace gvh {

// Host copy of the equation of state of a context.
struct HostEos {
  int kind = EOS_WB;
  int n = 0, M = 1, emin_exp = 0;
  std::vector<double> p, cs2, T;
};

double host_e_of_T(const HostEos &H, double T);
int eos_from_options(const gvh_options *o, HostEos &H);
// Argument checks shared by gvh_create and gvh_batch_submit.
int check_event_params(const gvh_lattice_parameters *L, const gvh_hydro_parameters *H);
int upload_eos_table(const HostEos &H, void **dev_pc, void **dev_tt, EosTable &T, hipStream_t st);
double host_eos_p(const HostEos &H, double e);

// Freeze-out state of a context (gvh_freezeout.hip).
struct FoState;
void fo_free(FoState *f);

}  // namespace gvh

int gvh_fail(int code, const char *fmt, ...);

struct gvh_context;
// Conserved field f (GVH_TTT..GVH_PI) of conserved buffer qbase as doubles on
// the full ghosted grid, whatever its storage precision.
int gvh_copy_q_field(gvh_context *c, const double *qbase, int f, double *host_ghosted);
// A ghosted block of G doubles on the device to the host: all of it
// (ghosted != 0) or the interior cells, x fastest.
int gvh_copy_ghosted_field(gvh_context *c, const double *d, int ghosted, double *host_out);

namespace gvh {
// Stage launch arguments of a context for stage 1 or 2 of the step at t.
StageArgs make_stage_args(gvh_context *c, int stage, double t);
Buffers context_buffers(const gvh_context *c);
Roles context_roles(const gvh_context *c);
double *fo_ring(FoState *f);
// The window search in phases (gvh_freezeout.hip): each phase may start once
// the previous one's stream work is complete.
int fo_init(gvh_context *c, int fofreq, double e_fo, int flags, int halves);
int fo_set_window_half(gvh_context *c, int half);
int fo_find_count(gvh_context *c, int n, hipStream_t st);
int fo_find_elements(gvh_context *c, hipStream_t st);
int fo_find_emit(gvh_context *c, double t, hipStream_t st, long long *nelements);
int fo_rows_async(gvh_context *c, double *rows, long long nrows, hipStream_t st);
// Fill n 64-bit words in a stream with a kernel (no hipMemsetAsync node).
hipError_t launch_fill_u64(unsigned long long *p, int n, unsigned long long v, hipStream_t st);
}  // namespace gvh

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t _e = (expr);                                                        \
    if (_e != hipSuccess)                                                          \
      return gvh_fail(GVH_EHIP, "%s failed: %s", #expr, hipGetErrorString(_e));  \
  } while (0)

struct gvh_context {
  gvh_lattice_parameters L{};
  gvh_hydro_parameters Hp{};
  int nx = 0, ny = 0, nz = 0, ncx = 0, ncy = 0, ncz = 0;
  long long G = 0;
  bool bulk = true;
  int precision = GVH_FP64;
  double zeta_tpeak = 1.01355;
  double e_reg = 0.0;
  gvh::HostEos eos;
  gvh::EosTable tab{};
  void *tab_pc = nullptr, *tab_tt = nullptr;

  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};

  void *arena = nullptr;
  // conserved buffers: T^{tau mu} as 4 blocks of G doubles, then pi and Pi as
  // 11 blocks of G doubles (FP64) or floats (mixed); qstride doubles apart
  double *qbuf[3] = {nullptr, nullptr, nullptr};  // [0],[1]: q^n / Q (swap), [2]: qS
  size_t qstride = 0;
  double *ubuf[3] = {nullptr, nullptr, nullptr};  // u^n, u^{n-1}, u* (rotate)
  double *ebuf[2] = {nullptr, nullptr};
  double *pbuf[2] = {nullptr, nullptr};
  unsigned long long *diag = nullptr;
  unsigned long long *count = nullptr;

  int iq = 0;   // buffer holding q^n
  int iu = 0;   // buffer holding u^n
  int iup = 1;  // buffer holding u^{n-1}
  int ie = 0;   // (e, p) buffer holding the current state
  int variant = 3;  // stage kernel arrangement (gvh_stage.hip launch_z)
  int strip = 16;   // y rows per traversal strip
  int max_iters = gvh::MAX_ITERS;  // Newton iteration budget per root-find
  bool initialized = false;
  long long steps_done = 0;     // KT steps since the initial state was set
  gvh::FoState *fo = nullptr;   // freeze-out history and surface buffers
  double *validity = nullptr;   // GVH_VALIDITY_FIELDS blocks of G doubles, on request

  int iQ() const { return 1 - iq; }
  int iqs() const { return 2; }
  int ius() const { return 3 - iu - iup; }
};
